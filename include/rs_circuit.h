/*
 * rs_circuit.h - flat device records shared by the product library and the
 * CPU oracle.
 *
 * The reference hands devices to the stamp loop as `Instance` values: a kind,
 * a node list and string-valued `value`/`params` that are parsed on every
 * stamp (crates/sim-core/src/circuit.rs, crates/sim-core/src/stamp.rs:94-553).
 * Here the parsing is done once, up front: every number a stamp reads is stored
 * as the double the reference's parser would produce, and the two parse
 * outcomes that change behaviour are kept as flags:
 *   - RS_F_WAVEFORM: `value` is a PULSE(...) string. stamp_voltage/stamp_current
 *     (stamp.rs:261-300) cannot parse it as a number, so the DC stamp is skipped
 *     (no aux row, no RHS); only stamp_*_at_time (stamp.rs:303-349) evaluates it.
 *   - RS_F_PMOS: the instance carries `type=pmos` (stamp.rs:399-410).
 *
 * Node numbers are MNA indices: 0 is ground and stays in the matrix
 * (circuit.rs:37-46). Aux unknowns (voltage-source branch currents) are not
 * listed here; they are numbered node_count + k in the order sources are first
 * stamped (mna.rs:472-480).
 */
#ifndef RS_CIRCUIT_H
#define RS_CIRCUIT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum rs_device_kind {
    RS_DEV_R = 1, /* resistor: value = ohms                         */
    RS_DEV_C = 2, /* capacitor: value = farads                      */
    RS_DEV_V = 3, /* voltage source: value = DC volts or PULSE      */
    RS_DEV_I = 4, /* current source: value = DC amps or PULSE       */
    RS_DEV_D = 5, /* diode: p[0] = is, p[1] = emission coefficient  */
    RS_DEV_M = 6  /* MOSFET: see RS_MP_* below                      */
};

enum rs_device_flags {
    RS_F_PMOS = 1,
    RS_F_WAVEFORM = 2
};

/* MOSFET parameter slots in rs_device.p (values after build_bsim_params,
 * crates/sim-devices/src/bsim/mod.rs:83-388). Level 1 reads VTH0, U0, TOX,
 * W and L; any other level (49 BSIM3, 54 BSIM4; others evaluate as BSIM3,
 * bsim/mod.rs:430-472) reads its model parameters from the circuit's
 * rs_bsim_model table entry p[RS_MP_MODEL], and W, L and TEMP. */
enum rs_mos_param {
    RS_MP_LEVEL = 0, /* model level (`level`, default 49: stamp.rs:400)    */
    RS_MP_VTH0 = 1,  /* vth0 / vto (level 1)                               */
    RS_MP_U0 = 2,    /* u0 [cm^2/Vs]; level 1 ignores KP (bsim/mod.rs:441) */
    RS_MP_TOX = 3,   /* tox [m] (level 1)                                  */
    RS_MP_W = 4,     /* instance W [m]                                     */
    RS_MP_L = 5,     /* instance L [m]                                     */
    RS_MP_MODEL = 6, /* BSIM: index into the rs_bsim_model table           */
    RS_MP_TEMP = 7   /* BSIM: instance `temp` [K] (default 300.15)         */
};

/* The BsimParams fields that evaluate_bsim_dc and evaluate_bsim4_dc read
 * (crates/sim-devices/src/bsim/params.rs:13-303), after build_bsim_params,
 * plus the instance's sa and sb (stamp.rs:419-421). pmos: type=pmos. */
typedef struct rs_bsim_model {
    double vth0, k1, k2, dvt0, dvt1, dvt2, eta0, dsub, drout, nlx, nfactor;
    double u0, ua, ub, uc, vsat, pclm, tox, lint, wint, rdsw, prt;
    double tnom, ute, kt1, kt1l, kt2;
    double saref, sbref, ku0, kvth0, tku0;
    double alpha0, alpha1, beta0, beta1;
    double jtss, jtsd, vtss, vtsd, nsti;
    double sa, sb;
    int32_t pmos, pad;
} rs_bsim_model;

/* PULSE parameter slots in rs_device.p when RS_F_WAVEFORM is set
 * (waveform.rs:680-709 field order). */
enum rs_pulse_param {
    RS_PP_V1 = 0, RS_PP_V2 = 1, RS_PP_TD = 2, RS_PP_TR = 3,
    RS_PP_TF = 4, RS_PP_PW = 5, RS_PP_PER = 6
};

typedef struct rs_device {
    int32_t kind;
    int32_t flags;
    int32_t node[4]; /* R/C/V/I/D: node[0]=a (+), node[1]=b (-); M: d, g, s, b */
    double value;
    double p[8];
} rs_device;

/* Newton options, engine.rs:987-999 mapping of options.rs:79-145 defaults. */
typedef struct rs_newton_opts {
    int32_t max_iters;   /* itl1, default 100         */
    int32_t gmin_steps;  /* always 0 in the reference */
    int32_t source_steps;/* always 0 in the reference */
    int32_t pad;
    double abs_tol;      /* vntol, default 1e-6       */
    double rel_tol;      /* reltol, default 1e-3      */
    double gmin;         /* gmin, default 1e-12       */
    double damping;      /* 1.0                       */
    double damping_min;  /* 0.1                       */
} rs_newton_opts;

#ifdef __cplusplus
}
#endif
#endif
