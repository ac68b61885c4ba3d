// device_eval.hpp - per-device stamp evaluation shared by the single-circuit
// stamp kernel and the batched DC kernel. Values are written in the order
// stamp.rs adds them: R stamp.rs:242-259, I :261-276/:331-349, V :278-325,
// C (BE) :759-787, C (TRAP) :871-915, D :351-387, M level 1 :389-553 with
// crates/sim-devices/src/bsim/evaluate.rs:285-356; PULSE waveform.rs:82-113.
#pragma once
#include <hip/hip_runtime.h>

#include "bsim_eval.hpp"
#include "stamp.hpp"

namespace rsgpu {

__device__ __forceinline__ double xget(const double* x, int xl, int i) {
    return (i >= 0 && i < xl) ? x[i] : 0.0;
}

__device__ __forceinline__ double fmax_rust(double a, double b) {
    // f64::max: returns the non-NaN operand
    return (a >= b || b != b) ? a : b;
}

__device__ __forceinline__ double pulse_eval(const double* p, double t) {
    const double v1 = p[0], v2 = p[1], td = p[2], tr = p[3], tf = p[4], pw = p[5], per = p[6];
    if (t < td) return v1;
    const double trl = per > 0.0 ? fmod(t - td, per) : t - td;
    if (trl < tr) return v1 + (v2 - v1) * trl / tr;
    if (trl < tr + pw) return v2;
    if (trl < tr + pw + tf) {
        const double tfall = trl - tr - pw;
        return v2 - (v2 - v1) * tfall / tf;
    }
    return v1;
}

// BSIM MOSFET (level != 1): the stamp.rs:389-553 values into its kBsimSlots
// block, and the mask of the conditional groups that are stamped.
__device__ __forceinline__ uint8_t eval_bsim(const rs_device& v, const double* __restrict__ x,
                                             int xl, const StampArgs& a, double* __restrict__ o) {
    const double gmin = a.gmin > 0.0 ? a.gmin : 1e-12;
    const int level = (int)v.p[RS_MP_LEVEL];
    const rs_bsim_model& m = a.models[(int)v.p[RS_MP_MODEL]];
    const double vd = xget(x, xl, v.node[0]), vg = xget(x, xl, v.node[1]);
    const double vs = xget(x, xl, v.node[2]), vb = xget(x, xl, v.node[3]);
    const BsimOut r = bsim_eval(m, level, v.p[RS_MP_W], v.p[RS_MP_L], vd, vg, vs, vb,
                                v.p[RS_MP_TEMP]);
    const double gm = r.gm, gds = r.gds >= gmin ? r.gds : gmin, gmbs = r.gmbs;
    uint8_t mask = 0;
    o[0] = gds; o[1] = gds; o[2] = -gds; o[3] = -gds;
    o[4] = gm; o[5] = -gm; o[6] = -gm; o[7] = gm;
    if (fabs(gmbs) > gmin * 0.01) mask |= 1;
    o[8] = gmbs; o[9] = -gmbs; o[10] = -gmbs; o[11] = gmbs;
    o[12] = -r.ieq; o[13] = r.ieq;
    if (level == 54) {
        if (fabs(r.isub) > gmin && r.gsub > gmin * 0.01) mask |= 2;
        const double isub_eq = r.isub - r.gsub * (vd - vb);
        o[14] = r.gsub; o[15] = r.gsub; o[16] = -r.gsub; o[17] = -r.gsub;
        o[18] = -isub_eq; o[19] = isub_eq;
        if (fabs(r.igs) > gmin && r.gigs > gmin * 0.01) mask |= 4;
        const double igs_eq = r.igs - r.gigs * (vg - vs);
        o[20] = r.gigs; o[21] = r.gigs; o[22] = -r.gigs; o[23] = -r.gigs;
        o[24] = -igs_eq; o[25] = igs_eq;
        if (fabs(r.igd) > gmin && r.gigd > gmin * 0.01) mask |= 8;
        const double igd_eq = r.igd - r.gigd * (vg - vd);
        o[26] = r.gigd; o[27] = r.gigd; o[28] = -r.gigd; o[29] = -r.gigd;
        o[30] = -igd_eq; o[31] = igd_eq;
    }
    return mask;
}

// Evaluates device v (index d for its capacitor history) into o[0..9].
__device__ __forceinline__ void eval_device(const rs_device& v, int d, const double* __restrict__ x,
                                            int xl, const StampArgs& a,
                                            const double* __restrict__ cap_v,
                                            const double* __restrict__ cap_i,
                                            const uint8_t* __restrict__ has_v,
                                            const uint8_t* __restrict__ has_i,
                                            double* __restrict__ o) {
    const bool dc = a.mode == rsgpu::MODE_DC;
    switch (v.kind) {
    case RS_DEV_R: {
        const double g = 1.0 / v.value;
        o[0] = g; o[1] = g; o[2] = -g; o[3] = -g;
        break;
    }
    case RS_DEV_I:
    case RS_DEV_V: {
        double val;
        if (dc) val = v.value;  // waveform sources are not stamped in DC
        else val = (v.flags & RS_F_WAVEFORM) ? pulse_eval(v.p, a.t) : v.value;
        val = val * a.source_scale;
        if (v.kind == RS_DEV_I) {
            o[0] = -val; o[1] = val;
        } else {
            o[0] = 1.0; o[1] = -1.0; o[2] = 1.0; o[3] = -1.0; o[4] = val;
        }
        break;
    }
    case RS_DEV_C: {
        if (dc) break;
        const double c = v.value;
        const double vprev = has_v[d] ? cap_v[d] : 0.0;
        if (a.mode == rsgpu::MODE_TRAN_BE) {
            const double g = c / a.dt;
            const double ieq = g * vprev;
            o[0] = g; o[1] = g; o[2] = -g; o[3] = -g; o[4] = -ieq; o[5] = ieq;
        } else {
            const double g = 2.0 * c / a.dt;
            const double iprev = has_i[d] ? cap_i[d] : 0.0;
            const double ieq = g * vprev + iprev;
            o[0] = g; o[1] = g; o[2] = -g; o[3] = -g; o[4] = ieq; o[5] = -ieq;
        }
        break;
    }
    case RS_DEV_D: {
        const double gmin = a.gmin > 0.0 ? a.gmin : 1e-12;
        const double vt = 0.02585 * v.p[1];
        const double vd = xget(x, xl, v.node[0]) - xget(x, xl, v.node[1]);
        const double e = exp(vd / vt);
        const double id = v.p[0] * (e - 1.0);
        const double gd = (v.p[0] / vt) * e;
        const double g = fmax_rust(gd, gmin);
        const double ieq = id - gd * vd;
        o[0] = g; o[1] = g; o[2] = -g; o[3] = -g; o[4] = -ieq; o[5] = ieq;
        break;
    }
    case RS_DEV_M: {
        const double gmin = a.gmin > 0.0 ? a.gmin : 1e-12;
        const double GMIN = 1e-12;
        const double EPS_OX = 3.9 * 8.854e-12;
        const bool pm = (v.flags & RS_F_PMOS) != 0;
        const double vd = xget(x, xl, v.node[0]), vg = xget(x, xl, v.node[1]);
        const double vs = xget(x, xl, v.node[2]);
        const double cox = EPS_OX / v.p[RS_MP_TOX];
        const double beta = v.p[RS_MP_U0] * 1e-4 * cox;
        double vdi, vgi, vsi, sign;
        if (pm) { vdi = -vs; vgi = -vg; vsi = -vd; sign = -1.0; }
        else { vdi = vd; vgi = vg; vsi = vs; sign = 1.0; }
        double vgs = vgi - vsi, vds = vdi - vsi;
        if (vds < 0.0) { vds = -vds; vgs = vgi - vdi; }
        const double vth0 = v.p[RS_MP_VTH0];
        const double vth = pm ? -fabs(vth0) : fabs(vth0);
        const double beta_eff = beta * v.p[RS_MP_W] / v.p[RS_MP_L];
        const double lambda = 0.02;
        double ids, gm, gds;
        if (vgs <= vth) {
            ids = 0.0; gm = 0.0; gds = GMIN;
        } else if (vds < vgs - vth) {
            ids = beta_eff * ((vgs - vth) * vds - 0.5 * vds * vds);
            gm = beta_eff * vds;
            gds = fmax_rust(beta_eff * ((vgs - vth) - vds), GMIN);
        } else {
            const double vov = vgs - vth;
            ids = 0.5 * beta_eff * (vov * vov) * (1.0 + lambda * vds);
            gm = beta_eff * (vgs - vth) * (1.0 + lambda * vds);
            gds = fmax_rust(0.5 * beta_eff * (vov * vov) * lambda, GMIN);
        }
        const double ids_s = ids * sign;
        double ieq = ids_s - gm * (vg - vs) - gds * (vd - vs);
        const double gdsm = fmax_rust(gds, gmin);
        if (a.flags & rsgpu::kStampJacobian) {
            // d(ids_s)/d(vd, vg, vs) of the current above: internal-frame
            // partials (reversed when the internal vds was negative), mapped
            // back through the PMOS flip (vd, vg, vs) -> (-vs, -vg, -vd).
            const bool rev = (vdi - vsi) < 0.0;
            const double fd = rev ? -gm - gds : gds, fg = gm, fs = rev ? gds : -gm - gds;
            const double jd = pm ? fs : fd, jg = fg, js = pm ? fd : fs;
            const double lk = gdsm - gds;  // the gmin floor stamped on (d, s)
            o[0] = jd + lk; o[1] = -js + lk; o[2] = js - lk; o[3] = -jd - lk;
            o[4] = jg; o[5] = 0.0; o[6] = -jg; o[7] = 0.0;
            // right-hand side consistent with these entries: J x - rhs equals
            // the Newton stamp's current ids_s + (gdsm - gds)(vd - vs), so a
            // solve of J x = rhs is a Newton step on the exact Jacobian
            ieq = ids_s - (jd * vd + jg * vg + js * vs);
        } else {
            o[0] = gdsm; o[1] = gdsm; o[2] = -gdsm; o[3] = -gdsm;
            o[4] = gm; o[5] = -gm; o[6] = -gm; o[7] = gm;
        }
        o[8] = -ieq; o[9] = ieq;
        break;
    }
    default:
        break;
    }
}

}  // namespace rsgpu
