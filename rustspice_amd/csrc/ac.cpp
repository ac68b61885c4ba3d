// ac.cpp - host side of the AC sweep (ac.hpp): frequency list, the stamp
// plan in the reference's insertion order, the excitation, launches in
// frequency chunks, and the (dB, degrees) conversion.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/rsgpu.h"
#include "ac.hpp"
#include "solver_impl.hpp"

using rsgpu::DevBuf;

struct rs_ac {
    std::vector<rs_device> dev;
    std::vector<rs_bsim_model> models;
    std::vector<double> ac_mag, ac_phase;
    int64_t node_count = 0;
    DevBuf<rs_device> d_dev;
    DevBuf<rs_bsim_model> d_models;
    DevBuf<double> d_x, d_vals, d_rhs, d_freqs, d_a, d_xs;
    DevBuf<uint8_t> d_mask;
    DevBuf<int32_t> d_pos, d_ptr, d_con, d_status;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_solve_ms = 0.0;
    int64_t last_launches = 0, last_n = 0, last_nent = 0, last_ncon = 0;
    std::string err;
    int32_t create_rc = 0;  // a failed rs_ac_create: every run returns it
};

namespace {

constexpr double kPi = 3.14159265358979323846;  // std::f64::consts::PI
constexpr size_t kScratchBudget = size_t(512) << 20;  // dense matrices per launch

int32_t afail(rs_ac* h, int32_t code, const std::string& m) {
    h->err = m;
    return code;
}
int32_t ahip(rs_ac* h, hipError_t e, const char* what) {
    return afail(h, RS_ERR_DEVICE, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

struct Con {
    int64_t row, col;
    int32_t code;
};

}  // namespace

extern "C" {

int64_t rs_ac_frequencies(int32_t sweep, int64_t points, double fstart, double fstop, double* out,
                          int64_t cap) {
    int64_t total = 0;
    auto put = [&](int64_t i, double f) {
        if (out && i < cap) out[i] = f;
    };
    if (sweep == 0 || sweep == 1) {
        const double span = sweep == 0 ? std::log10(fstop) - std::log10(fstart) : std::log2(fstop / fstart);
        const double c = std::ceil(span * (double)points);
        total = (c != c || c <= 0.0) ? 1 : (int64_t)c + 1;  // `as usize` saturates at 0
        const double den = (double)std::max<int64_t>(total - 1, 1);
        for (int64_t i = 0; i < total; i++) {
            if (sweep == 0) put(i, std::pow(10.0, std::log10(fstart) + (double)i * span / den));
            else put(i, fstart * std::pow(2.0, (double)i * span / den));
        }
    } else if (sweep == 2) {
        total = std::max<int64_t>(points, 0);
        const double step = (fstop - fstart) / (double)std::max<int64_t>(points - 1, 1);
        for (int64_t i = 0; i < total; i++) put(i, fstart + (double)i * step);
    } else {
        return -1;
    }
    return total;
}

rs_ac* rs_ac_create(const rs_device* dev, int64_t ndev, int64_t node_count, const rs_bsim_model* models,
                    int64_t nmodels, const double* ac_mag, const double* ac_phase) {
    rs_ac* h = new rs_ac();
    h->dev.assign(dev, dev + std::max<int64_t>(ndev, 0));
    h->node_count = node_count;
    h->ac_mag.assign(ac_mag, ac_mag + std::max<int64_t>(ndev, 0));
    h->ac_phase.assign(ac_phase, ac_phase + std::max<int64_t>(ndev, 0));
    // stamp_mos_ac evaluates through evaluate_mos, whose BSIM4 arm passes
    // sa = sb = 0 (crates/sim-devices/src/bsim/mod.rs:463-469)
    if (models && nmodels > 0) h->models.assign(models, models + nmodels);
    for (auto& m : h->models) m.sa = m.sb = 0.0;
    for (const auto& d : h->dev) {
        if (d.kind < RS_DEV_R || d.kind > RS_DEV_M) {
            h->create_rc = afail(h, RS_ERR_INVALID_MATRIX, "Invalid matrix: unknown device kind in the AC device list");
            return h;
        }
        if (d.kind == RS_DEV_M && (int)d.p[RS_MP_LEVEL] != 1 &&
            (d.p[RS_MP_MODEL] < 0 || (int64_t)d.p[RS_MP_MODEL] >= (int64_t)h->models.size())) {
            h->create_rc = afail(h, RS_ERR_INVALID_MATRIX, "Invalid matrix: BSIM device without a model table entry");
            return h;
        }
        for (int q = 0; q < 4; q++)
            if (d.node[q] < 0 || d.node[q] >= node_count) {
                h->create_rc = afail(h, RS_ERR_INVALID_MATRIX, "Invalid matrix: device node out of range");
                return h;
            }
    }
    hipError_t e;
    if ((e = h->d_dev.ensure(h->dev.size())) ||
        (e = hipMemcpy(h->d_dev.p, h->dev.data(), h->dev.size() * sizeof(rs_device), hipMemcpyHostToDevice)))
        h->create_rc = ahip(h, e, "rs_ac_create");
    else if (!h->models.empty() &&
             ((e = h->d_models.ensure(h->models.size())) ||
              (e = hipMemcpy(h->d_models.p, h->models.data(), h->models.size() * sizeof(rs_bsim_model),
                             hipMemcpyHostToDevice))))
        h->create_rc = ahip(h, e, "rs_ac_create");
    else if ((e = hipEventCreate(&h->ev0)) || (e = hipEventCreate(&h->ev1)))
        h->create_rc = ahip(h, e, "hipEventCreate");
    return h;
}

void rs_ac_destroy(rs_ac* h) {
    if (!h) return;
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    h->d_dev.release(); h->d_models.release(); h->d_x.release(); h->d_vals.release();
    h->d_rhs.release(); h->d_freqs.release(); h->d_a.release(); h->d_xs.release();
    h->d_mask.release(); h->d_pos.release(); h->d_ptr.release(); h->d_con.release();
    h->d_status.release();
    delete h;
}

int32_t rs_ac_last_error(const rs_ac* h, char* buf, int64_t len) {
    if (len > 0) std::snprintf(buf, (size_t)len, "%s", h->err.c_str());
    return (int32_t)h->err.size();
}

int32_t rs_ac_run(rs_ac* h, const double* x_dc, int64_t nx, const double* freqs, int64_t nfreq,
                  double* x_out, double* magph, int64_t cap_n, int64_t* n_out, int64_t* nsolved) {
    if (h->create_rc) return h->create_rc;
    h->err.clear();
    *nsolved = 0;
    const int64_t ndev = (int64_t)h->dev.size(), nc = h->node_count;
    int64_t naux = 0;
    for (const auto& d : h->dev) naux += d.kind == RS_DEV_V;
    const int64_t n = nc + naux;
    *n_out = n;
    if (n > 16384) return afail(h, RS_ERR_INVALID_MATRIX, "Invalid matrix: AC matrix larger than 16384");
    if (cap_n < n) return afail(h, RS_ERR_INVALID_MATRIX, "Invalid matrix: x_out row shorter than the MNA size");
    hipError_t e;
    // 1. small-signal values at the DC point, then the constant 1.0
    const int64_t nv = 3 * ndev + 1;
    if ((e = h->d_x.ensure(std::max<int64_t>(nx, 1))) || (e = h->d_vals.ensure(nv)) ||
        (e = h->d_mask.ensure(std::max<int64_t>(ndev, 1))))
        return ahip(h, e, "hipMalloc(ac eval)");
    if (nx > 0 && (e = hipMemcpy(h->d_x.p, x_dc, nx * sizeof(double), hipMemcpyHostToDevice)))
        return ahip(h, e, "hipMemcpy(x_dc)");
    const double one = 1.0;
    if ((e = hipMemcpy(h->d_vals.p + 3 * ndev, &one, sizeof(double), hipMemcpyHostToDevice)))
        return ahip(h, e, "hipMemcpy(vals)");
    if ((e = rsgpu::launch_ac_eval(h->d_dev.p, ndev, h->d_models.p, h->d_x.p, nx, h->d_vals.p, h->d_mask.p, 0)))
        return ahip(h, e, "k_ac_eval");
    std::vector<uint8_t> mask(std::max<int64_t>(ndev, 1));
    if (ndev > 0 && (e = hipMemcpy(mask.data(), h->d_mask.p, ndev, hipMemcpyDeviceToHost)))
        return ahip(h, e, "hipMemcpy(mask)");
    // 2. contributions in stamp order (stamp.rs:1525-1747), excitation on the host
    std::vector<Con> cons;
    std::vector<double> rhs(2 * n, 0.0);
    auto add = [&](int64_t r, int64_t c, int64_t vi, bool neg, bool imag) {
        cons.push_back(Con{r, c, (int32_t)((vi << 2) | (neg ? 2 : 0) | (imag ? 1 : 0))});
    };
    auto rhs_add = [&](int64_t i, double re, double im) {
        rhs[2 * i] += re;
        rhs[2 * i + 1] += im;
    };
    const int64_t kOne = 3 * ndev;
    int64_t k = nc;
    for (int64_t d = 0; d < ndev; d++) {
        const rs_device& v = h->dev[d];
        const int64_t a = v.node[0], b = v.node[1];
        switch (v.kind) {
        case RS_DEV_R:
        case RS_DEV_D:
        case RS_DEV_C: {
            const bool im = v.kind == RS_DEV_C;
            add(a, a, 3 * d, false, im); add(b, b, 3 * d, false, im);
            add(a, b, 3 * d, true, im); add(b, a, 3 * d, true, im);
            break;
        }
        case RS_DEV_V: {
            add(a, k, kOne, false, false); add(b, k, kOne, true, false);
            add(k, a, kOne, false, false); add(k, b, kOne, true, false);
            const double th = h->ac_phase[d] * kPi / 180.0;  // Complex64::from_polar
            rhs_add(k, h->ac_mag[d] * std::cos(th), h->ac_mag[d] * std::sin(th));
            k++;
            break;
        }
        case RS_DEV_I: {
            const double th = h->ac_phase[d] * kPi / 180.0;
            const double re = h->ac_mag[d] * std::cos(th), im = h->ac_mag[d] * std::sin(th);
            rhs_add(a, -re, -im);
            rhs_add(b, re, im);
            break;
        }
        case RS_DEV_M: {
            const int64_t dr = v.node[0], ga = v.node[1], so = v.node[2], bu = v.node[3];
            add(dr, dr, 3 * d, false, false); add(so, so, 3 * d, false, false);
            add(dr, so, 3 * d, true, false); add(so, dr, 3 * d, true, false);
            add(dr, ga, 3 * d + 1, false, false); add(dr, so, 3 * d + 1, true, false);
            add(so, ga, 3 * d + 1, true, false); add(so, so, 3 * d + 1, false, false);
            if (mask[d]) {
                add(dr, bu, 3 * d + 2, false, false); add(dr, so, 3 * d + 2, true, false);
                add(so, bu, 3 * d + 2, true, false); add(so, so, 3 * d + 2, false, false);
            }
            break;
        }
        default:
            break;
        }
    }
    // ground fix (engine.rs:882-891)
    cons.erase(std::remove_if(cons.begin(), cons.end(), [](const Con& c) { return c.row == 0; }), cons.end());
    cons.push_back(Con{0, 0, (int32_t)(kOne << 2)});
    rhs[0] = rhs[1] = 0.0;
    // 3. group by entry in first-appearance order; an entry's terms keep stamp order
    std::unordered_map<int64_t, int32_t> slot;
    std::vector<int32_t> pos, cnt;
    std::vector<int32_t> which(cons.size());
    for (size_t q = 0; q < cons.size(); q++) {
        const int64_t key = cons[q].row * n + cons[q].col;
        auto it = slot.find(key);
        int32_t s;
        if (it == slot.end()) {
            s = (int32_t)pos.size();
            slot.emplace(key, s);
            pos.push_back((int32_t)key);
            cnt.push_back(0);
        } else {
            s = it->second;
        }
        cnt[s]++;
        which[q] = s;
    }
    const int64_t nent = (int64_t)pos.size();
    std::vector<int32_t> ptr(nent + 1, 0), con(cons.size()), fill(nent, 0);
    for (int64_t s = 0; s < nent; s++) ptr[s + 1] = ptr[s] + cnt[s];
    for (size_t q = 0; q < cons.size(); q++) con[ptr[which[q]] + fill[which[q]]++] = cons[q].code;
    // 4. upload, then the frequencies in chunks that keep the dense scratch bounded
    const size_t per = (size_t)n * (size_t)n * 2 * sizeof(double);
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nfreq, (int64_t)(kScratchBudget / std::max<size_t>(per, 1))));
    if ((e = h->d_pos.ensure(nent)) || (e = h->d_ptr.ensure(nent + 1)) || (e = h->d_con.ensure(con.size())) ||
        (e = h->d_rhs.ensure(2 * n)) || (e = h->d_freqs.ensure(std::max<int64_t>(nfreq, 1))) ||
        (e = h->d_a.ensure((size_t)chunk * n * n * 2)) || (e = h->d_xs.ensure((size_t)std::max<int64_t>(nfreq, 1) * n * 2)) ||
        (e = h->d_status.ensure(std::max<int64_t>(nfreq, 1))))
        return ahip(h, e, "hipMalloc(ac)");
    if ((e = hipMemcpy(h->d_pos.p, pos.data(), nent * 4, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(h->d_ptr.p, ptr.data(), (nent + 1) * 4, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(h->d_con.p, con.data(), con.size() * 4, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(h->d_rhs.p, rhs.data(), 2 * n * sizeof(double), hipMemcpyHostToDevice)) ||
        (nfreq > 0 && (e = hipMemcpy(h->d_freqs.p, freqs, nfreq * sizeof(double), hipMemcpyHostToDevice))))
        return ahip(h, e, "hipMemcpy(ac plan)");
    rsgpu::AcDev p;
    p.n = (int32_t)n;
    p.nent = (int32_t)nent;
    p.ent_pos = h->d_pos.p;
    p.ent_ptr = h->d_ptr.p;
    p.con = h->d_con.p;
    p.vals = h->d_vals.p;
    p.rhs = h->d_rhs.p;
    p.freqs = h->d_freqs.p;
    h->last_launches = 0;
    if ((e = hipEventRecord(h->ev0, 0))) return ahip(h, e, "hipEventRecord");
    for (int64_t f0 = 0; f0 < nfreq; f0 += chunk) {
        if ((e = rsgpu::launch_ac_solve(p, f0, std::min(chunk, nfreq - f0), h->d_a.p, h->d_xs.p,
                                        h->d_status.p, 0)))
            return ahip(h, e, "k_ac_solve");
        h->last_launches++;
    }
    if ((e = hipEventRecord(h->ev1, 0)) || (e = hipEventSynchronize(h->ev1))) return ahip(h, e, "k_ac_solve");
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
    h->last_solve_ms = ms;
    h->last_n = n;
    h->last_nent = nent;
    h->last_ncon = (int64_t)con.size();
    // 5. results up to the first failed frequency (engine.rs:902-929)
    std::vector<int32_t> status(std::max<int64_t>(nfreq, 1));
    std::vector<double> xs((size_t)std::max<int64_t>(nfreq, 1) * n * 2);
    if (nfreq > 0 &&
        ((e = hipMemcpy(status.data(), h->d_status.p, nfreq * 4, hipMemcpyDeviceToHost)) ||
         (e = hipMemcpy(xs.data(), h->d_xs.p, (size_t)nfreq * n * 2 * sizeof(double), hipMemcpyDeviceToHost))))
        return ahip(h, e, "hipMemcpy(ac results)");
    int64_t ok = 0;
    while (ok < nfreq && status[ok]) ok++;
    for (int64_t f = 0; f < ok; f++) {
        if (x_out)
            std::memcpy(x_out + f * cap_n * 2, xs.data() + f * n * 2, n * 2 * sizeof(double));
        if (magph)
            for (int64_t i = 0; i < nc; i++) {
                const double re = xs[(f * n + i) * 2], im = xs[(f * n + i) * 2 + 1];
                const double mag = std::hypot(re, im);  // Complex::norm
                magph[(f * nc + i) * 2] = mag > 1e-30 ? 20.0 * std::log10(mag) : -600.0;
                magph[(f * nc + i) * 2 + 1] = std::atan2(im, re) * 180.0 / kPi;  // arg in degrees
            }
    }
    *nsolved = ok;
    return RS_OK;
}

int32_t rs_ac_timing(const rs_ac* h, double* out) {
    out[0] = h->last_solve_ms;
    out[1] = (double)h->last_launches;
    out[2] = (double)h->last_n;
    out[3] = (double)h->last_nent;
    out[4] = (double)h->last_ncon;
    return RS_OK;
}

}  // extern "C"
