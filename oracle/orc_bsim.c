/*
 * orc_bsim.c - BSIM3 / BSIM4 DC evaluation. TEST INFRASTRUCTURE (oracle).
 *
 * Restates crates/sim-devices/src/bsim/:
 *   evaluate.rs:77-279   evaluate_bsim_dc
 *   evaluate.rs:387-620  evaluate_bsim4_dc
 *   threshold.rs:52-160  calculate_vth
 *   mobility.rs:55-110   calculate_mobility
 *   channel.rs:59-220    calculate_vdsat, calculate_clm_factor, calculate_rds
 *   bsim4/substrate.rs:41-80, bsim4/tunneling.rs:55-130, bsim4/stress.rs:48-90
 * with params.rs:8-12 constants and params.rs:458-475 leff/weff/cox.
 * Rust f64::max/min are fmax/fmin (NaN-ignoring), powf is pow. Every
 * expression keeps the reference's operand order.
 */
#include "orc.h"
#include "orc_util.h"

static const double EPS_SI = 11.7 * 8.854e-12;
static const double EPS_OX = 3.9 * 8.854e-12;
static const double Q_E = 1.602e-19;
static const double K_B = 1.381e-23;
static const double BSIM_GMIN = 1e-12; /* evaluate.rs:51 */

static const rs_bsim_model *g_models = NULL;
static int64_t g_nmodels = 0;

void orc_set_bsim_models(const rs_bsim_model *m, int64_t n) { g_models = m; g_nmodels = n; }

const rs_bsim_model *orc_bsim_model(int64_t i) {
    return (i >= 0 && i < g_nmodels) ? &g_models[i] : NULL;
}

static double leff_of(const rs_bsim_model *m, double l) { return fmax(l - 2.0 * m->lint, 1e-9); }
static double weff_of(const rs_bsim_model *m, double w) { return fmax(w - 2.0 * m->wint, 1e-9); }

/* threshold.rs:52-160 */
static void calc_vth(const rs_bsim_model *m, double vbs, double vds, double leff, double weff,
                     double temp, double *vth_o, double *dvth_o) {
    double vt = K_B * temp / Q_E;
    double phi = 0.7;
    double vth = m->vth0;
    double dvth = 0.0;
    double phi_minus_vbs = fmax(phi - vbs, 0.01);
    double sqrt_phi = sqrt(phi);
    double sqrt_phi_vbs = sqrt(phi_minus_vbs);
    double body1 = m->k1 * (sqrt_phi_vbs - sqrt_phi);
    double dbody1 = -m->k1 / (2.0 * sqrt_phi_vbs);
    double body2 = m->k2 * vbs;
    double dbody2 = m->k2;
    vth += body1 + body2;
    dvth += dbody1 + dbody2;
    double xdep = sqrt(2.0 * EPS_SI * phi / (Q_E * 1e17));
    double lt = sqrt(EPS_SI * m->tox * xdep / (EPS_SI * 3.9 / 11.7));
    double lt_safe = fmax(lt, 1e-9);
    double sce_ratio = leff / (2.0 * lt_safe);
    double sce_exp = sce_ratio > 20.0 ? 0.0 : exp(-m->dvt1 * sce_ratio);
    double sce = -m->dvt0 * sce_exp * vt;
    double sce_vbs = sce * (1.0 + m->dvt2 * vbs);
    double dsce = sce * m->dvt2;
    vth += sce_vbs;
    dvth += dsce;
    double vds_abs = fabs(vds);
    double dibl = -m->eta0 * vds_abs;
    double dsub_term = m->dsub * exp(-m->drout * leff / lt_safe);
    double dibl_length = dsub_term * vds_abs;
    vth += dibl - dibl_length;
    if (m->nlx > 0.0 && weff > 0.0) {
        double nwe = m->nlx / weff;
        vth += nwe * 0.1;
    }
    double temp_ratio = temp / m->tnom;
    double vth_temp = m->kt1 * (temp_ratio - 1.0);
    double vth_temp_l = m->kt1l / leff * (temp_ratio - 1.0);
    double vth_temp_vbs = m->kt2 * vbs * (temp_ratio - 1.0);
    vth += vth_temp + vth_temp_l + vth_temp_vbs;
    dvth += m->kt2 * (temp_ratio - 1.0);
    *vth_o = vth;
    *dvth_o = dvth;
}

/* mobility.rs:55-110 */
static double calc_mobility(const rs_bsim_model *m, double vgs, double vbs, double vth,
                            double temp) {
    double temp_ratio = temp / m->tnom;
    double temp_factor = pow(temp_ratio, m->ute);
    double u0_temp = m->u0 * temp_factor;
    double vt = K_B * temp / Q_E;
    double vgst = fmax(vgs - vth, 0.0);
    double eeff = (vgst + 2.0 * vt) / (6.0 * m->tox);
    double ua_term = m->ua * eeff;
    double ub_term = m->ub * eeff * eeff;
    double uc_term = m->uc * vbs * eeff;
    double denom = 1.0 + ua_term + ub_term + uc_term;
    return fmax(u0_temp / fmax(denom, 0.1), 1.0);
}

/* channel.rs:59-90 */
static void calc_vdsat(const rs_bsim_model *m, double vgs, double vth, double ueff, double leff,
                       double *vdsat, double *dvdsat) {
    double vgst = fmax(vgs - vth, 1e-6);
    double ueff_m2 = ueff * 1e-4;
    double esat = 2.0 * m->vsat / ueff_m2;
    double esat_leff = esat * leff;
    double denom = vgst + esat_leff;
    *vdsat = vgst * esat_leff / denom;
    *dvdsat = (esat_leff * esat_leff) / (denom * denom);
}

/* channel.rs:116-150 */
static void calc_clm(const rs_bsim_model *m, double vds, double vdsat, double leff, double ueff,
                     double *clm, double *dclm) {
    double vds_excess = fmax(vds - vdsat, 0.0);
    if (vds_excess <= 0.0) { *clm = 1.0; *dclm = 0.0; return; }
    double ueff_m2 = ueff * 1e-4;
    double esat = 2.0 * m->vsat / ueff_m2;
    double va_eff = esat * leff / fmax(m->pclm, 0.1);
    double ratio = vds_excess / va_eff;
    double clamped = fmin(ratio, 0.9);
    double f = 1.0 / (1.0 - clamped);
    *clm = f;
    *dclm = f * f / va_eff;
}

/* channel.rs:207-225 */
static double calc_rds(const rs_bsim_model *m, double weff, double temp) {
    if (m->rdsw <= 0.0) return 0.0;
    double weff_um = weff * 1e6;
    double rds_base = m->rdsw / fmax(weff_um, 0.1);
    double delta_t = temp - m->tnom;
    double rds = rds_base * (1.0 + m->prt * delta_t);
    return fmax(rds, 0.0);
}

/* bsim4/stress.rs:48-90 */
static void calc_stress(double sa, double sb, double saref, double sbref, double ku0,
                        double kvth0, double temp, double tnom, double tku0, double *u0_mult,
                        double *dvth) {
    if (fabs(ku0) < 1e-20 && fabs(kvth0) < 1e-20) { *u0_mult = 1.0; *dvth = 0.0; return; }
    double sa_eff = fmax(sa, 1e-9), sb_eff = fmax(sb, 1e-9);
    double saref_eff = fmax(saref, 1e-9), sbref_eff = fmax(sbref, 1e-9);
    double inv_sa = 1.0 / sa_eff - 1.0 / saref_eff;
    double inv_sb = 1.0 / sb_eff - 1.0 / sbref_eff;
    double stress = (inv_sa + inv_sb) / 2.0;
    double temp_factor = 1.0 + tku0 * (temp - tnom);
    *u0_mult = fmax(1.0 + ku0 * stress * temp_factor, 0.1);
    *dvth = kvth0 * stress;
}

/* bsim4/substrate.rs:41-80 */
static void calc_isub(double ids, double vds, double vdsat, double alpha0, double alpha1,
                      double beta0, double beta1, double vbs, double *isub, double *gsub) {
    double vds_excess = vds - vdsat;
    if (alpha0 <= 0.0 || vds_excess < 0.01) { *isub = 0.0; *gsub = 0.0; return; }
    vds_excess = fmax(vds_excess, 0.01);
    double alpha_eff = alpha0 * (1.0 + alpha1 * vds);
    double beta_eff = fmax(beta0 + beta1 * vbs, 1.0);
    double exp_arg = fmax(-beta_eff / vds_excess, -40.0);
    double exp_term = exp(exp_arg);
    double is = alpha_eff * vds_excess * fabs(ids) * exp_term;
    double dis = alpha_eff * fabs(ids) * exp_term * (1.0 + beta_eff / vds_excess);
    *isub = fmax(is, 0.0);
    *gsub = fmax(dis, 0.0);
}

/* bsim4/tunneling.rs:55-130 (one side) */
static void tunnel_side(double v, double area, double j, double vts, double n, double *i,
                        double *g) {
    if (!(j > 0.0 && vts > 0.0)) { *i = 0.0; *g = 0.0; return; }
    double vt_eff = vts * n;
    if (v > 0.0) {
        double exp_arg = fmin(v / vt_eff, 40.0);
        double e = exp(exp_arg);
        *i = area * j * e;
        *g = *i / vt_eff;
    } else {
        double exp_arg = fmax(v / vt_eff, -40.0);
        double e = exp(exp_arg);
        *i = area * j * e;
        *g = fabs(*i / vt_eff);
    }
}

/* evaluate_bsim_dc (level != 54) and evaluate_bsim4_dc (level 54).
 * out: ids, gm, gds, gmbs, ieq, region (0 cutoff, 1 linear, 2 saturation),
 * vth, isub, gsub, igs, igd, gigs, gigd. */
void orc_bsim_eval(const rs_bsim_model *m, int level, double w, double l, double vd, double vg,
                   double vs, double vb, double temp, double out[13]) {
    const int b4 = level == 54;
    double vdi, vgi, vsi, vbi, sign;
    if (m->pmos) { vdi = -vs; vgi = -vg; vsi = -vd; vbi = -vb; sign = -1.0; }
    else { vdi = vd; vgi = vg; vsi = vs; vbi = vb; sign = 1.0; }
    double vgs = vgi - vsi, vds = vdi - vsi;
    double vbs = vbi - vsi;
    if (vds < 0.0) { vds = -vds; vgs = vgi - vdi; }
    double leff = leff_of(m, l), weff = weff_of(m, w);
    double cox = EPS_OX / m->tox;
    double vt = K_B * temp / Q_E;
    double u0_mult = 1.0, vth_shift = 0.0;
    if (b4 && (m->sa > 0.0 || m->sb > 0.0))
        calc_stress(fmax(m->sa, m->saref), fmax(m->sb, m->sbref), m->saref, m->sbref, m->ku0,
                    m->kvth0, temp, m->tnom, m->tku0, &u0_mult, &vth_shift);
    double vth, dvth;
    calc_vth(m, vbs, vds, leff, weff, temp, &vth, &dvth);
    if (b4) vth = vth + vth_shift;
    double vgst = vgs - vth;
    int region;
    double ids, gm, gds, gmbs, vdsat = 0.0;
    if (vgst <= 0.0) {
        region = 0;
        double n = fmax(m->nfactor, 1.0);
        double i0 = b4 ? weff / leff * m->u0 * 1e-4 * u0_mult * cox * vt * vt * (n - 1.0)
                       : weff / leff * m->u0 * 1e-4 * cox * vt * vt * (n - 1.0);
        double exp_vgst = exp(vgst / (n * vt));
        double exp_vds = exp(-vds / vt);
        ids = i0 * exp_vgst * (1.0 - exp_vds);
        ids = fmax(ids, 0.0);
        gm = ids / (n * vt);
        gds = i0 * exp_vgst * exp_vds / vt;
        gmbs = -gm * dvth;
        gds = fmax(gds, BSIM_GMIN);
        gm = fmax(gm, BSIM_GMIN * 0.01);
    } else {
        double ueff = calc_mobility(m, vgs, vbs, vth, temp);
        if (b4) ueff = ueff * u0_mult;
        double dvdsat;
        calc_vdsat(m, vgs, vth, ueff, leff, &vdsat, &dvdsat);
        double ueff_m2 = ueff * 1e-4;
        double beta = weff / leff * ueff_m2 * cox;
        if (vds < vdsat) {
            region = 1;
            ids = beta * (vgst * vds - 0.5 * vds * vds);
            gm = beta * vds;
            gds = fmax(beta * (vgst - vds), BSIM_GMIN);
            gmbs = -gm * dvth;
        } else {
            region = 2;
            double clm, dclm;
            calc_clm(m, vds, vdsat, leff, ueff, &clm, &dclm);
            double ids_sat = 0.5 * beta * vdsat * vdsat;
            ids = ids_sat * clm;
            gm = beta * vdsat * dvdsat * clm;
            gds = fmax(ids_sat * dclm, BSIM_GMIN);
            double gds_dibl = gm * m->eta0;
            gds += gds_dibl;
            gmbs = -gm * dvth;
        }
    }
    ids = fmax(ids, 0.0);
    double rds = calc_rds(m, weff, temp);
    if (rds > 0.0 && ids > 0.0) {
        double v_rds = ids * rds;
        if (v_rds < vds * 0.5) gds = gds / (1.0 + rds * gds);
    }
    double isub = 0.0, gsub = 0.0, igs = 0.0, igd = 0.0, gigs = 0.0, gigd = 0.0;
    if (b4) {
        if (m->alpha0 > 0.0 && region == 2)
            calc_isub(ids, vds, vdsat, m->alpha0, m->alpha1, m->beta0, m->beta1, vbs, &isub, &gsub);
        double vgd = vgs - vds;
        if (m->jtss > 0.0 || m->jtsd > 0.0) {
            double area = weff * leff;
            double n = fmax(m->nsti, 1.0);
            tunnel_side(vgs, area, m->jtss, m->vtss, n, &igs, &gigs);
            tunnel_side(vgd, area, m->jtsd, m->vtsd, n, &igd, &gigd);
        }
        isub = isub * fabs(sign);
        igs = igs * fabs(sign);
        igd = igd * fabs(sign);
    }
    ids *= sign;
    double vgs_o = vg - vs, vds_o = vd - vs, vbs_o = vb - vs;
    double ieq = ids - gm * vgs_o - gds * vds_o - gmbs * vbs_o;
    out[0] = ids; out[1] = gm; out[2] = gds; out[3] = gmbs; out[4] = ieq;
    out[5] = (double)region; out[6] = vth; out[7] = isub; out[8] = gsub;
    out[9] = igs; out[10] = igd; out[11] = gigs; out[12] = gigd;
}
