// bsim_eval.hpp - BSIM3 (evaluate_bsim_dc, level 49 and any level other than
// 1 and 54) and BSIM4 (evaluate_bsim4_dc, level 54) DC evaluation on the
// device, following crates/sim-devices/src/bsim/evaluate.rs:77-279 and
// :387-620 with threshold.rs:52-160, mobility.rs:55-110, channel.rs:59-225,
// bsim4/stress.rs:48-90, bsim4/substrate.rs:41-80, bsim4/tunneling.rs:55-130.
// Operations keep the reference's operand order (FP64, no contraction);
// exp/pow/sqrt are the device library's.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/rs_circuit.h"

namespace rsgpu {

struct BsimOut {
    double ids, gm, gds, gmbs, ieq;
    int region;  // 0 cutoff, 1 linear, 2 saturation
    double vth, isub, gsub, igs, igd, gigs, gigd;
};

// f64::max / f64::min (NaN-ignoring)
__host__ __device__ __forceinline__ double rmax(double a, double b) { return (a >= b || b != b) ? a : b; }
__host__ __device__ __forceinline__ double rmin(double a, double b) { return (a <= b || b != b) ? a : b; }

__device__ __forceinline__ void bsim_tunnel(double v, double area, double j, double vts, double n,
                                            double& i, double& g) {
    i = 0.0;
    g = 0.0;
    if (!(j > 0.0 && vts > 0.0)) return;
    const double vt_eff = vts * n;
    if (v > 0.0) {
        i = area * j * exp(rmin(v / vt_eff, 40.0));
        g = i / vt_eff;
    } else {
        i = area * j * exp(rmax(v / vt_eff, -40.0));
        g = fabs(i / vt_eff);
    }
}

__device__ inline BsimOut bsim_eval(const rs_bsim_model& m, int level, double w, double l,
                                    double vd, double vg, double vs, double vb, double temp) {
    constexpr double kEpsSi = 11.7 * 8.854e-12, kEpsOx = 3.9 * 8.854e-12;
    constexpr double kQ = 1.602e-19, kK = 1.381e-23, kGmin = 1e-12;
    const bool b4 = level == 54;
    BsimOut o{};
    // PMOS: (vd, vg, vs, vb) -> (-vs, -vg, -vd, -vb), current sign -1
    const double sign = m.pmos ? -1.0 : 1.0;
    const double vdi = m.pmos ? -vs : vd, vgi = m.pmos ? -vg : vg, vsi = m.pmos ? -vd : vs;
    const double vbi = m.pmos ? -vb : vb;
    double vgs = vgi - vsi, vds = vdi - vsi;
    const double vbs = vbi - vsi;
    if (vds < 0.0) {
        vds = -vds;
        vgs = vgi - vdi;
    }
    const double leff = rmax(l - 2.0 * m.lint, 1e-9), weff = rmax(w - 2.0 * m.wint, 1e-9);
    const double cox = kEpsOx / m.tox;
    const double vt = kK * temp / kQ;
    // stress (BSIM4 with SA/SB given)
    double u0m = 1.0, dvth_stress = 0.0;
    if (b4 && (m.sa > 0.0 || m.sb > 0.0) && !(fabs(m.ku0) < 1e-20 && fabs(m.kvth0) < 1e-20)) {
        const double sa = rmax(rmax(m.sa, m.saref), 1e-9), sb = rmax(rmax(m.sb, m.sbref), 1e-9);
        const double sar = rmax(m.saref, 1e-9), sbr = rmax(m.sbref, 1e-9);
        const double st = ((1.0 / sa - 1.0 / sar) + (1.0 / sb - 1.0 / sbr)) / 2.0;
        u0m = rmax(1.0 + m.ku0 * st * (1.0 + m.tku0 * (temp - m.tnom)), 0.1);
        dvth_stress = m.kvth0 * st;
    }
    // threshold voltage
    double vth, dvth;
    {
        const double phi = 0.7;
        const double spv = sqrt(rmax(phi - vbs, 0.01)), sp = sqrt(phi);
        vth = m.vth0;
        vth += m.k1 * (spv - sp) + m.k2 * vbs;
        dvth = 0.0 + (-m.k1 / (2.0 * spv) + m.k2);
        const double xdep = sqrt(2.0 * kEpsSi * phi / (kQ * 1e17));
        const double lt = rmax(sqrt(kEpsSi * m.tox * xdep / (kEpsSi * 3.9 / 11.7)), 1e-9);
        const double ratio = leff / (2.0 * lt);
        const double sce = -m.dvt0 * (ratio > 20.0 ? 0.0 : exp(-m.dvt1 * ratio)) * vt;
        vth += sce * (1.0 + m.dvt2 * vbs);
        dvth += sce * m.dvt2;
        const double va = fabs(vds);
        vth += -m.eta0 * va - m.dsub * exp(-m.drout * leff / lt) * va;
        if (m.nlx > 0.0 && weff > 0.0) vth += m.nlx / weff * 0.1;
        vth += m.kt1 * (temp / m.tnom - 1.0) + m.kt1l / leff * (temp / m.tnom - 1.0) +
               m.kt2 * vbs * (temp / m.tnom - 1.0);
        dvth += m.kt2 * (temp / m.tnom - 1.0);
    }
    if (b4) vth = vth + dvth_stress;
    const double vgst = vgs - vth;
    double ids, gm, gds, gmbs, vdsat = 0.0;
    if (vgst <= 0.0) {
        o.region = 0;
        const double n = rmax(m.nfactor, 1.0);
        const double i0 = b4 ? weff / leff * m.u0 * 1e-4 * u0m * cox * vt * vt * (n - 1.0)
                             : weff / leff * m.u0 * 1e-4 * cox * vt * vt * (n - 1.0);
        const double eg = exp(vgst / (n * vt)), ed = exp(-vds / vt);
        ids = rmax(i0 * eg * (1.0 - ed), 0.0);
        gm = ids / (n * vt);
        gds = i0 * eg * ed / vt;
        gmbs = -gm * dvth;
        gds = rmax(gds, kGmin);
        gm = rmax(gm, kGmin * 0.01);
    } else {
        // mobility
        const double vgp = rmax(vgs - vth, 0.0);
        const double eeff = (vgp + 2.0 * vt) / (6.0 * m.tox);
        const double den = 1.0 + m.ua * eeff + m.ub * eeff * eeff + m.uc * vbs * eeff;
        double ueff = rmax(m.u0 * pow(temp / m.tnom, m.ute) / rmax(den, 0.1), 1.0);
        if (b4) ueff = ueff * u0m;
        // saturation voltage
        const double esl = 2.0 * m.vsat / (ueff * 1e-4) * leff;
        const double vgx = rmax(vgs - vth, 1e-6);
        const double dn = vgx + esl;
        vdsat = vgx * esl / dn;
        const double dvdsat = (esl * esl) / (dn * dn);
        const double beta = weff / leff * (ueff * 1e-4) * cox;
        if (vds < vdsat) {
            o.region = 1;
            ids = beta * (vgst * vds - 0.5 * vds * vds);
            gm = beta * vds;
            gds = rmax(beta * (vgst - vds), kGmin);
            gmbs = -gm * dvth;
        } else {
            o.region = 2;
            double clm = 1.0, dclm = 0.0;
            const double ex = rmax(vds - vdsat, 0.0);
            if (ex > 0.0) {
                const double va_eff = 2.0 * m.vsat / (ueff * 1e-4) * leff / rmax(m.pclm, 0.1);
                clm = 1.0 / (1.0 - rmin(ex / va_eff, 0.9));
                dclm = clm * clm / va_eff;
            }
            const double ids_sat = 0.5 * beta * vdsat * vdsat;
            ids = ids_sat * clm;
            gm = beta * vdsat * dvdsat * clm;
            gds = rmax(ids_sat * dclm, kGmin);
            gds += gm * m.eta0;
            gmbs = -gm * dvth;
        }
    }
    ids = rmax(ids, 0.0);
    if (m.rdsw > 0.0) {  // series resistance (channel.rs:207-225)
        const double rds = rmax(m.rdsw / rmax(weff * 1e6, 0.1) * (1.0 + m.prt * (temp - m.tnom)), 0.0);
        if (rds > 0.0 && ids > 0.0 && ids * rds < vds * 0.5) gds = gds / (1.0 + rds * gds);
    }
    if (b4) {
        if (m.alpha0 > 0.0 && o.region == 2) {  // impact ionization
            double ex = vds - vdsat;
            if (ex >= 0.01) {
                ex = rmax(ex, 0.01);
                const double ae = m.alpha0 * (1.0 + m.alpha1 * vds);
                const double be = rmax(m.beta0 + m.beta1 * vbs, 1.0);
                const double et = exp(rmax(-be / ex, -40.0));
                o.isub = rmax(ae * ex * fabs(ids) * et, 0.0);
                o.gsub = rmax(ae * fabs(ids) * et * (1.0 + be / ex), 0.0);
            }
        }
        if (m.jtss > 0.0 || m.jtsd > 0.0) {
            const double area = weff * leff, n = rmax(m.nsti, 1.0);
            bsim_tunnel(vgs, area, m.jtss, m.vtss, n, o.igs, o.gigs);
            bsim_tunnel(vgs - vds, area, m.jtsd, m.vtsd, n, o.igd, o.gigd);
        }
        o.isub = o.isub * fabs(sign);
        o.igs = o.igs * fabs(sign);
        o.igd = o.igd * fabs(sign);
    }
    ids *= sign;
    o.ids = ids;
    o.gm = gm;
    o.gds = gds;
    o.gmbs = gmbs;
    o.ieq = ids - gm * (vg - vs) - gds * (vd - vs) - gmbs * (vb - vs);
    o.vth = vth;
    return o;
}

}  // namespace rsgpu
