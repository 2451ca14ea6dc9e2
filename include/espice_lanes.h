/* espice_lanes: sweep points as MI355X (gfx950) lanes behind a C ABI.
 *
 * This is the drop-in for ESPice's GPU-lane path. It replaces two boundaries
 * of the reference (OmarSiwy/ESPice 1.2.0):
 *
 *   - the internal device-lane context `GpuLanes` (src/analysis/gpu_lanes.zig:
 *     251-1027) and its raw kernels `arp_lane_*` / `arp_l*_<model>`
 *     (src/solver/lane_device.zig:26-258, src/device/eval.zig:2998-3203),
 *     which a caller such as `tran.runGpuLanes` (src/analysis/tran/tran.zig:
 *     1654-1820) drives tick by tick; and
 *   - the points API of include/espice.h:244-360 (`espice_append_points`,
 *     `espice_copy_point_measures`, `espice_release_points`,
 *     `espice_lane_width`), whose rows of live-parameter values become lanes.
 *
 * Conventions follow include/espice.h:139-158: every call returns an
 * esl_status; handles are generation-checked tokens, never addresses; calls
 * on one handle must be serialized, different handles are independent; no
 * call keeps a caller pointer after it returns; caller-owned output buffers
 * are sized by the documented counts. There is no CPU fallback: a handle is
 * created only when the HIP device and the kernels are present, and every
 * numerical call runs on the GPU.
 *
 * The circuit is described as a deck of devices (esl_deck) rather than
 * netlist text: netlist parsing is not part of the lane path. Units are SI.
 */
#ifndef ESPICE_LANES_H
#define ESPICE_LANES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ESL_ABI_VERSION 5u

/* esl_status: the values of espice_status (include/espice.h:21-38). */
typedef uint32_t esl_status;
enum {
    ESL_OK = 0, ESL_INVALID_ARGUMENT = 1, ESL_BUFFER_TOO_SMALL = 2,
    ESL_OUT_OF_MEMORY = 3, ESL_INVALID_QUERY = 4, ESL_RESULT_UNAVAILABLE = 5,
    ESL_FAILED = 6, ESL_ABI_MISMATCH = 7, ESL_INVALID_HANDLE = 8,
    /* No HIP device, or the kernels failed to load or launch: the caller
     * runs its points elsewhere, as tran.zig:1702-1716 does on a refused
     * GpuLanes.init. */
    ESL_DEVICE_UNAVAILABLE = 9
};

/* Per-point reasons: the values of espice_reason (include/espice.h:259-284). */
enum {
    ESL_REASON_OK = 0, ESL_REASON_NO_CONVERGENCE = 1,
    ESL_REASON_SINGULAR_MATRIX = 3, ESL_REASON_INVALID_PARAMETER = 9,
    ESL_REASON_OTHER = 255
};

/* Device kinds. The stamp order of a deck is kind-major in this order, then
 * deck order within a kind (the reference stamps batch by batch, instances
 * in tape order: src/device/eval.zig:782-922). */
enum {
    ESL_DEV_R = 0,     /* models/resistor.va  nodes p n          */
    ESL_DEV_C = 1,     /* models/capacitor.va nodes p n          */
    ESL_DEV_V = 2,     /* models/vsource.va   nodes p n (DC)     */
    ESL_DEV_I = 3,     /* models/isource.va   nodes p n (DC)     */
    ESL_DEV_D = 4,     /* models/diode.va     nodes a k (RS = 0) */
    ESL_DEV_MOS1 = 5,  /* models/mos1.va      nodes d g s b      */
    ESL_DEV_BJT = 6,   /* models/gummel_poon.va nodes c b e s (RB = RC = RE = 0) */
    ESL_DEV_KINDS = 7
};

#define ESL_NPARAM 32   /* doubles per device in esl_deck.params */
#define ESL_NTERM 4     /* terminals per device in esl_deck.nodes */

/* Parameter slots. R/C/V/I use slot 0 (r, c, dc, dc). */
enum {
    ESL_P_VALUE = 0,
    /* MOS1 (models/mos1.va:17-80): */
    ESL_M_TYPE = 0, ESL_M_VTO = 1, ESL_M_KP = 2, ESL_M_GAMMA = 3,
    ESL_M_PHI = 4, ESL_M_LAMBDA = 5, ESL_M_W = 6, ESL_M_L = 7, ESL_M_IS = 8,
    ESL_M_LD = 9, ESL_M_TOX = 10, ESL_M_U0 = 11, ESL_M_M = 12,
    ESL_M_CBD = 13, ESL_M_CBS = 14, ESL_M_PB = 15, ESL_M_MJ = 16,
    ESL_M_FC = 17, ESL_M_CGSO = 18, ESL_M_CGDO = 19, ESL_M_CGBO = 20,
    ESL_M_GMIN = 21, ESL_M_TEMP = 22, ESL_M_TNOM = 23,
    /* flag bits of which optional parameters were given (ESL_MG_*) */
    ESL_M_GIVEN = 24,
    /* Diode level 1 (models/diode.va): */
    ESL_D_IS = 0, ESL_D_N = 1, ESL_D_CJO = 2, ESL_D_VJ = 3, ESL_D_M = 4,
    ESL_D_TT = 5, ESL_D_FC = 6, ESL_D_AREA = 7, ESL_D_GMIN = 8,
    ESL_D_TEMP = 9, ESL_D_TNOM = 10, ESL_D_EG = 11, ESL_D_XTI = 12,
    /* breakdown BV (0: off), IBV, NBV (0: N); high-injection knee IKF (0:
     * off); recombination ISR (0: off), NR. A series RS is its own
     * ESL_DEV_R to an internal anode node (the builder adds it). */
    ESL_D_BV = 13, ESL_D_IBV = 14, ESL_D_NBV = 15, ESL_D_IKF = 16,
    ESL_D_ISR = 17, ESL_D_NR = 18,
    /* V source waveform (transient): slot 0 is the value at t <= td (and
     * the operating point's); ESL_V_WAVE 1 = step to ESL_V_V2 after
     * ESL_V_TD seconds. */
    ESL_V_WAVE = 1, ESL_V_V2 = 2, ESL_V_TD = 3,
    /* Gummel-Poon (models/gummel_poon.va:24-94, DC subset: no RB/RC/RE,
     * caps or excess phase; ESL_Q_SUBS +1 vertical only). Zero VAF, VAR,
     * IKF, IKR mean infinite. */
    ESL_Q_TYPE = 0, ESL_Q_IS = 1, ESL_Q_BF = 2, ESL_Q_NF = 3, ESL_Q_VAF = 4,
    ESL_Q_IKF = 5, ESL_Q_ISE = 6, ESL_Q_NE = 7, ESL_Q_BR = 8, ESL_Q_NR = 9,
    ESL_Q_VAR = 10, ESL_Q_IKR = 11, ESL_Q_ISC = 12, ESL_Q_NC = 13,
    ESL_Q_NKF = 14, ESL_Q_ISS = 15, ESL_Q_NS = 16, ESL_Q_AREA = 17,
    ESL_Q_M = 18, ESL_Q_GMIN = 19, ESL_Q_TEMP = 20, ESL_Q_TNOM = 21,
    ESL_Q_EG = 22, ESL_Q_XTI = 23, ESL_Q_XTB = 24, ESL_Q_AREAB = 25,
    ESL_Q_AREAC = 26, ESL_Q_SUBS = 27
};
/* Bits of params[ESL_M_GIVEN] (stored as a double holding a small integer):
 * the reference's $param_given tests (models/mos1.va:177-217). */
enum { ESL_MG_TOX = 1, ESL_MG_KP = 2, ESL_MG_PHI = 4, ESL_MG_GAMMA = 8 };

/* A deck: n_nodes nodes (node 0 is ground), n_dev devices, and P live
 * parameters. Every binding sets params[dev][param] = live[point][live_idx]
 * for each point, the role of a `.param` named in live_params
 * (include/espice.h:86-109). Arrays are row-major. */
typedef struct {
    uint32_t n_nodes;
    uint32_t n_dev;
    const uint32_t *kind;    /* [n_dev] ESL_DEV_* */
    const uint32_t *nodes;   /* [n_dev][ESL_NTERM], unused terminals 0 */
    const double *params;    /* [n_dev][ESL_NPARAM] */
    uint32_t n_live;
    uint32_t n_bind;
    const uint32_t *bind_dev;    /* [n_bind] */
    const uint32_t *bind_param;  /* [n_bind] parameter slot */
    const uint32_t *bind_live;   /* [n_bind] live index < n_live */
} esl_deck;

/* Newton tolerances: src/core/numerics.zig:377-404 defaults. */
typedef struct {
    double reltol, abstol, vntol, residual_tol;
    uint32_t max_iter;      /* max(itl1, 100), converger.zig:85-96 */
    uint32_t reserved;
} esl_tolerances;

/* esl_op_options.precision */
enum {
    ESL_FP64 = 0,          /* FP64 factors throughout (the drop-in) */
    ESL_FP32_REFINED = 1   /* FP32 factors, FP64 residual + refinement */
};
/* esl_op_options.tiling: how a lane maps onto the GPU. AUTO times both
 * tilings on the engine's first large batch and keeps the faster. */
enum {
    ESL_TILING_AUTO = 0,
    ESL_TILING_WAVE = 1,    /* one wavefront per lane, factors in LDS */
    ESL_TILING_THREAD = 2,  /* one thread per lane, state lane-major in HBM;
                               FP64 from cold or given starts, every model */
    ESL_TILING_BLOCK = 3    /* esl_batch_solve: one workgroup of 4 wavefronts per
                               lane; device evaluation spread over the four */
};
/* esl_op_options.start */
enum {
    ESL_START_COLD = 0,    /* zero x + junction seeds, INITFIX (dc/op.zig:24-27) */
    ESL_START_GIVEN = 1,   /* x0 supplied per point, no INITFIX */
    /* x0 from the network attached by esl_engine_set_warm_start, evaluated
     * on the device from each point's live values; a lane whose residual
     * does not fall in its first iteration, or that has not converged after
     * max_iter iterations, restarts from the attached nominal operating
     * point (max_iter more iterations). No INITFIX. */
    ESL_START_LEARNED = 2
};

typedef struct {
    size_t struct_size;    /* sizeof(esl_op_options) */
    esl_tolerances tol;
    uint32_t precision;    /* ESL_FP64 | ESL_FP32_REFINED */
    uint32_t start;        /* ESL_START_* */
    /* ESL_START_GIVEN: [count][n] start vectors (n = esl_engine_info.n). */
    const double *x0;
    /* FP32_REFINED: each Newton solve dx = -J^-1 F is refined in FP64
     * until max|r| <= refine_tol * max(max|F|, max_i sum_j |J_ij dx_j|)
     * with r = -F - J dx (defaults 1e-12, 8 steps). A step that does not
     * halve max|r|, or refine_max steps without success, sends the lane to
     * FP64 factors for that iteration and the rest of its Newton. */
    double refine_tol;
    uint32_t refine_max;
    uint32_t tiling;       /* ESL_TILING_* */
    /* Measures: with n_probes > 0 the solve returns x [count][n_probes],
     * x[p][k] = unknown probes[k] of point p (the role of the points API's
     * per-point measures, include/espice.h:338), instead of [count][n]. */
    const uint32_t *probes;
    uint32_t n_probes;
    /* ESL_OP_HOST_REPIVOT: a lane whose replay fails leaves for the host,
     * which factors its Jacobian into a new tape (the default for FP64 is
     * the lane's own factor on the device; both give the same bits).
     * ESL_OP_LADDER: FP64 cold or given starts that the plain rung leaves
     * unconverged (or singular) climb the .op ladder (dc/op.zig:160-288),
     * every run a Newton run on the device: the gmin rung (a conductance g
     * from every node row to ground, g from 1e-3 down to 1e-12 by an
     * adaptive factor, at most 50 iterations a run, a failed run backing up
     * toward the last good g from the last converged point; then a clean
     * run without it), and when that fails the source rung (every
     * independent source scaled by lambda, lambda from 0 to 1 in steps that
     * grow 1.5x after a converged run and halve after a failed one) and a
     * final run at full sources. The point's iterations are the sum over
     * its runs. Each lane climbs its own ladder inside the same launch (the
     * rung control runs on the device), so hard points cost only their own
     * runs; lanes fitting a workgroup's LDS (wavefront tiling). */
    uint32_t flags;
} esl_op_options;
enum { ESL_OP_HOST_REPIVOT = 1u, ESL_OP_LADDER = 2u };

/* Sizes of a compiled engine. */
typedef struct {
    uint32_t n;            /* MNA unknowns, ground (0) included */
    uint32_t nnz;          /* pattern entries */
    uint32_t n_nodes;      /* node unknowns 0..n_nodes-1; branches follow */
    uint32_t n_live;
    uint32_t nnz_l, nnz_u; /* strict L and U of the lead tape */
    uint64_t flops;        /* F: refactor multiply-adds of the lead tape */
    uint32_t levels;       /* column levels of the lead tape's refactor */
    uint32_t lds_bytes;    /* LDS per lane of the wavefront kernel */
    uint32_t tapes;        /* distinct pivot tapes in use */
    uint32_t reserved;
    uint32_t refactor_steps;  /* rows of the lead tape's refactor chain stream */
    uint32_t solve_steps;     /* rows of its forward + backward solve stream */
    uint32_t tiling;          /* ESL_TILING_* the autotuner settled on, 0 if none */
    uint32_t reserved2;
} esl_engine_info;

typedef struct esl_engine esl_engine;

uint32_t esl_abi_version(void);
/* The calling thread's last failure, NUL-terminated (empty when none);
 * *required (may be NULL) receives the size with the terminator. Like
 * espice_error_message (include/espice.h:239), per thread instead of per
 * problem. */
esl_status esl_error_message(char *buf, size_t capacity, size_t *required);
void esl_default_tolerances(esl_tolerances *tol);
void esl_default_op_options(esl_op_options *opt);

/* Compiles the deck into MNA form on the host (pattern, slots, the fill-
 * reducing column order), and allocates nothing on the device yet.
 * INVALID_ARGUMENT for a malformed deck (unknown kind, node out of range,
 * binding out of range, an unsupported parameter such as MOS1 rd/rs > 0);
 * DEVICE_UNAVAILABLE when no HIP device is present. diag receives a
 * NUL-terminated reason when capacity > 0. */
esl_status esl_engine_create(const esl_deck *deck, esl_engine **out,
                             char *diag, size_t diag_capacity);
esl_status esl_engine_destroy(esl_engine *e);
esl_status esl_engine_info_get(esl_engine *e, esl_engine_info *out);
/* The MNA pattern (CSC, n + 1 column pointers, nnz row indices) and the
 * column order q (n entries) the lead factor used. Sizes from info. */
esl_status esl_engine_pattern(esl_engine *e, uint32_t *col_ptr,
                              uint32_t *row_idx, uint32_t *q);

/* Solves the operating point of `count` points: values is
 * [count][n_live] row-major. Outputs: x [count][n] (the converged iterate,
 * or the last one; [count][n_probes] with probes), iterations [count] (Newton iterations, counted as
 * converger.Result.iterations), reasons [count] (ESL_REASON_*). Lanes run
 * on the GPU as one wavefront each. The first point's first Jacobian fixes
 * the pivot tape (gpu_lanes.zig:361-374); a lane whose replay fails the
 * growth, pivot or wide-multiplier test gets its own tape from a host
 * symbolic re-pivot of its matrix and resumes on the GPU. */
esl_status esl_op_solve(esl_engine *e, const double *values, size_t count,
                        const esl_op_options *opt, double *x,
                        uint32_t *iterations, uint32_t *reasons);

/* ---- Transient lanes (gpu_lanes.zig / tran.runGpuLanes, tran.zig:
 * 1654-1820, with fixed steps). Each point first solves its operating point
 * (cold start, as esl_op_solve), then integrates `steps` fixed steps of dt:
 * step 0 backward Euler, later steps the trapezoid (xmu = 0.5: ag0 = 2/dt,
 * ag1 = 1, integrator.zig:40-53), each a Newton solve on the device with the
 * .op gate order and the residual gate off (lane_device.zig:74-98), starting
 * from the last two accepted points extrapolated (x + (x - x_prev), the
 * predictor of tran.zig:1131-1137 at a fixed step) and limited against the
 * last accepted point. The
 * lanes share one tape from the first converged point's first transient
 * Jacobian; a lane whose replay fails is re-pivoted on the host and resumes.
 * The launches, each `steps_per_launch` steps, are captured in one hipGraph
 * and replayed. A step that does not converge in max_iter iterations ends
 * the lane (ESL_REASON_NO_CONVERGENCE): fixed-step lanes do not cut the
 * step. ---- */
/* A .meas card of the lanes' subset (the cards of src/output/measure.zig's
 * evaluate, :575-660, that read one transient vector: core.Measure with its
 * clauses, core/deck.zig:311-340): WHEN, TRIG/TARG, FIND (at AT=, or at
 * WHEN b), AVG, MIN, MAX, PP, INTEG, RMS. A clause names an unknown (node);
 * WHEN v(a)=v(b) names the level's unknown in level_node with
 * has_level_node = 1. Counts rise/fall/cross: ESL_MEAS_UNSET, ESL_MEAS_LAST
 * or k > 0; at: ESL_MEAS_NO_AT when unset; val 1e99 when unset (a clause
 * without VAL never fires); td, from, to: 0 when unset. Card b: TARG, or the
 * WHEN of FIND a WHEN b. */
enum { ESL_MEAS_WHEN = 1, ESL_MEAS_TRIG_TARG, ESL_MEAS_FIND, ESL_MEAS_AVG, ESL_MEAS_MIN, ESL_MEAS_MAX,
       ESL_MEAS_PP, ESL_MEAS_INTEG, ESL_MEAS_RMS };
#define ESL_MEAS_UNSET (-1)
#define ESL_MEAS_LAST (-2)
#define ESL_MEAS_NO_AT 1e99
#define ESL_MEAS_MAX_CARDS 64
typedef struct {
    uint32_t node, level_node, has_level_node, reserved;
    double val, td, from, to, at;
    int32_t rise, fall, cross, reserved2;
} esl_meas_clause;
typedef struct {
    uint32_t func, reserved[3];
    esl_meas_clause a, b;
} esl_meas_card;

typedef struct {
    size_t struct_size;        /* sizeof(esl_tran_options) */
    esl_tolerances tol;        /* per step (max_iter default 50) */
    double dt;                 /* step, s */
    uint32_t steps;
    uint32_t steps_per_launch; /* default 10 */
    const uint32_t *probes;    /* unknowns recorded after every step */
    uint32_t n_probes;
    uint32_t use_graph;        /* 1 (default): capture the launches in a hipGraph */
    /* Progress and cancellation (the reference ends a running batch on a
     * cancelled query, tran.zig:1824-1826): when set, the launches run in
     * groups of at least progress_steps steps (0: 100), each group captured
     * in its own graph, and after each group progress(ctx, steps done,
     * steps) is called on the calling thread; a nonzero return cancels the
     * solve (ESL_FAILED, esl_error_message says so). Results do not depend
     * on the grouping. */
    int (*progress)(void *ctx, uint32_t steps_done, uint32_t steps);
    void *progress_ctx;
    uint32_t progress_steps;
    uint32_t reserved;
    /* .meas cards measured on the device as each point accepts its steps,
     * from its operating point's row (t = 0) on, as the reference's Stream
     * measures a running transient (src/output/measure.zig:160-330):
     * meas_out [count][n_meas] receives each card's value (NaN where its
     * event never came); at most ESL_MEAS_MAX_CARDS cards. meas_stop = 1: a
     * point stops stepping once every card has resolved (its event found,
     * or past its TO= / AT=), the early stop of CHANGELOGS.md:13-14;
     * steps_done [count] (may be NULL) receives each point's steps, and its
     * probe rows past them are left untouched. */
    const esl_meas_card *meas;
    uint32_t n_meas;
    uint32_t meas_stop;
    double *meas_out;
    uint32_t *steps_done;
} esl_tran_options;
void esl_default_tran_options(esl_tran_options *opt);
/* values [count][n_live]; outputs x [count][n] (the last accepted state),
 * probe_out [count][steps][n_probes] (may be NULL when n_probes = 0),
 * iterations [count] (transient Newton iterations), reasons [count]. */
esl_status esl_tran_solve(esl_engine *e, const double *values, size_t count,
                          const esl_tran_options *opt, double *x, double *probe_out,
                          uint32_t *iterations, uint32_t *reasons);

/* ---- Adaptive transients: step control per lane on the device. Replaces
 * the host step control of tran.runGpuLanes and its round calls
 * (src/analysis/tran/tran.zig:728-1086, 1654-1820; GpuLanes.predictRound /
 * probe / apply, src/analysis/gpu_lanes.zig:699-713,890-969) for the
 * trapezoid method: each lane runs Sim.begin/Sim.end itself -- the
 * predictor, Newton, the LTE bound of integrator.stepBound
 * (src/solver/lane_kernels.zig:230-269, ngspice CKTterr coefficients, over
 * the row plane of charges), rejection below 0.9 dt, growth at most 2x and
 * at most dt_max, order 1 -> 2 promotion, dt / 8 at order 1 after a failed
 * Newton, landing on step-source edges. The first dt is
 * min(min(t_step, t_stop/100)/10, dt_max), cut to a tenth of the first
 * edge, /10 (dctran.c's first step). Lanes run `attempts_per_launch`
 * attempts per launch until every lane reaches t_stop or fails. A lane
 * whose replay fails re-pivots on the device; one whose factors outgrow
 * that room ends with ESL_REASON_OTHER. Not for engines whose lanes exceed
 * a workgroup's LDS. ---- */
typedef struct {
    size_t struct_size;        /* sizeof(esl_tran_adaptive_options) */
    esl_tolerances tol;        /* each step's Newton (max_iter default 100) */
    double t_stop;             /* s */
    double t_step;             /* the print step: sets the first dt */
    double dt_max;             /* 0: t_stop / 50 (ngspice's tmax) */
    double dt_min;             /* a dt below it ends the lane (1e-18) */
    double trtol, chgtol;      /* 7, 1e-14 (core/numerics.zig:377-404) */
    uint32_t max_steps;        /* accepted steps per lane: rows of times and probe_out */
    uint32_t attempts_per_launch;  /* default 64 */
    const uint32_t *probes;    /* unknowns recorded at every accepted point */
    uint32_t n_probes;
    uint32_t reserved;
} esl_tran_adaptive_options;
void esl_default_tran_adaptive_options(esl_tran_adaptive_options *opt);
/* values [count][n_live]. Outputs x [count][n] (the last accepted state),
 * times [count][max_steps] and probe_out [count][max_steps][n_probes] (the
 * accepted points; rows past accepted[l] are untouched; probe_out may be
 * NULL when n_probes = 0), accepted / rejected [count] (steps),
 * iterations [count] (Newton iterations of every attempt), reasons [count]
 * (NO_CONVERGENCE: dt fell below dt_min or max_steps ran out). */
esl_status esl_tran_adaptive_solve(esl_engine *e, const double *values, size_t count,
                                   const esl_tran_adaptive_options *opt, double *x,
                                   double *times, double *probe_out, uint32_t *accepted,
                                   uint32_t *rejected, uint32_t *iterations,
                                   uint32_t *reasons);

/* ---- Persistent uneven batches (no reference counterpart: the reference
 * forms one batch per template and options, executor.zig:480-520,
 * session.zig:562-590). Points of several templates, .op and transient
 * mixed, run in ONE launch of a persistent kernel: the host orders every
 * point (a job) longest predicted first, and workgroups and wavefronts
 * claim the next job as they finish one. A job runs lane per wavefront (four
 * lanes per workgroup, each in a quarter of the LDS) or lane per block (one
 * lane on the workgroup: the device evaluation spread over its four
 * wavefronts, the factor and solves in the first); the autotuner picks the
 * tiling per template from measured job times and keeps it on the engine.
 * A point whose replay fails leaves its Jacobian; the host factors it into
 * a new tape and the point resumes, with every other such point of the
 * batch, in one further persistent launch per round. Results equal esl_op_solve /
 * esl_tran_solve on the same points bit for bit (same lead tapes, same
 * arithmetic). ---- */
enum {
    ESL_BATCH_OP = 0,     /* operating points (cold start, FP64) */
    ESL_BATCH_TRAN = 1    /* operating points, then a fixed-step transient */
};
typedef struct {
    esl_engine *engine;        /* each engine at most once per batch */
    const double *values;      /* [count][n_live] */
    size_t count;
    uint32_t analysis;         /* ESL_BATCH_* */
    uint32_t flags;            /* ESL_OP_LADDER: the operating points the plain
                                  Newton leaves unconverged or singular climb the
                                  .op ladder (dc/op.zig:160-288) in their job of the
                                  persistent launch (the gmin and source rungs, as
                                  esl_op_solve's ESL_OP_LADDER), in both analyses (a
                                  transient's points start their steps from the
                                  ladder's result); iterations sum every run.
                                  0: no ladder. Other bits must be 0. */
    esl_tolerances tol;        /* the operating points' Newton, both analyses (esl_tran_solve
                                  runs its points' with the transient's tolerances and at
                                  least 100 iterations) */
    const esl_tran_options *tran;  /* ESL_BATCH_TRAN: dt, steps, tol, probes
                                      (steps_per_launch, use_graph and progress
                                      unused) */
    double *x;                 /* out [count][n] */
    double *probe_out;         /* out, ESL_BATCH_TRAN with probes: [count][steps][n_probes] */
    uint32_t *iterations;      /* out [count] (.op: Newton iterations; transient:
                                  Newton iterations over all steps) */
    uint32_t *reasons;         /* out [count] */
} esl_batch_item;
typedef struct {
    size_t struct_size;        /* sizeof(esl_batch_options) */
    uint32_t tiling;           /* ESL_TILING_AUTO (the autotuner), _WAVE or _BLOCK for every job */
    uint32_t tune;             /* AUTO: 1 = measure both tilings on this batch
                                  (the first batch of an engine always does) */
    uint32_t reserved[6];
} esl_batch_options;
void esl_default_batch_options(esl_batch_options *opt);
typedef struct {
    size_t struct_size;        /* sizeof(esl_batch_report) */
    double kernel_ms;          /* the persistent launch, HIP events */
    double repivot_ms;         /* follow-up launches for re-pivoted points */
    double setup_ms;           /* host: lead tapes, uploads, schedule */
    double total_ms;           /* whole call */
    double predicted_ms;       /* the scheduler's makespan estimate (0 before tuning) */
    double lead_ms;            /* the transient lead pre-pass launch (part of setup_ms) */
    uint32_t grid;             /* workgroups of the persistent launch */
    uint32_t block_jobs;       /* jobs run lane per block */
    uint32_t wave_jobs;        /* jobs run lane per wavefront */
    uint32_t repivots;         /* re-pivots on the host (a point may need several) */
    uint32_t lds_bytes;        /* dynamic LDS per workgroup */
    uint32_t repivot_launches; /* persistent launches after the first (one per round) */
    uint32_t reserved[2];
} esl_batch_report;
/* report may be NULL. */
esl_status esl_batch_solve(const esl_batch_item *items, size_t n_items,
                           const esl_batch_options *opt, esl_batch_report *report);

/* Device-time split of the last esl_op_solve, milliseconds. */
typedef struct {
    double kernel_ms;      /* Newton kernel(s), HIP events */
    double total_ms;       /* upload + kernels + download */
    uint32_t launches;     /* Newton kernel launches (1 + re-pivot / fallback rounds) */
    uint32_t repivots;     /* lanes re-pivoted on the host */
    uint64_t lane_iterations; /* sum of Newton iterations over lanes */
    uint32_t tiling;       /* ESL_TILING_* the last solve ran (first launch) */
    uint32_t fp64_lanes;   /* FP32_REFINED: lanes that fell back to FP64 factors */
    uint32_t warm_fallbacks;  /* START_LEARNED: lanes restarted from the nominal point */
    uint32_t device_repivot_lanes; /* lanes that resumed with their own LU on the device */
    double warm_ms;        /* START_LEARNED: network inference, HIP events */
    double lead_ms;        /* host: lead point's Jacobian, factor and tape program */
    double download_ms;    /* host: results to the caller */
} esl_timing;
esl_status esl_last_timing(esl_engine *e, esl_timing *out);

/* Development hook (no counterpart in the reference): enable != 0 makes the
 * following esl_op_solve calls count shader cycles per lane in five phases
 * (assembly, refactor, solve, update + limiting + gates, the full factors of
 * lanes on their own LU) and, sixth, the lane's own-LU replays that ran with
 * their values in LDS; out receives the last solve's counters [count][6]
 * when capacity allows. */
esl_status esl_debug_profile(esl_engine *e, int enable, unsigned long long *out,
                             size_t capacity);

/* Development hook, not part of the drop-in: the refactor and solves of the
 * kernel for lanes beyond a workgroup's LDS (the replay cut into blocks of
 * at most block_slots slots, the solves in column chunks), run on the CPU
 * for `lanes` matrices a [lanes][nnz] and right-hand sides rhs [lanes][n]
 * on the tape factored from lead_vals with column order q. Writes per lane
 * L [lanes][nnz_l], U [lanes][nnz_u], the pivots [lanes][n], dx [lanes][n]
 * (row order), fail (1: the lane must re-pivot) and sizes[4] = {blocks,
 * nnz_l, nnz_u, rows of all blocks} (lanes == 0: only sizes). Needs no device; lets the block
 * programs be checked against the serial replay (lane_kernels.zig:49-143)
 * on any host. */
esl_status esl_debug_big_replay(uint32_t n, const uint32_t *col_ptr,
                                const uint32_t *row_idx,
                                const double *lead_vals, const uint32_t *q,
                                double pivot_tol, uint32_t block_slots,
                                const double *a, const double *rhs,
                                uint32_t lanes, double *lx, double *ux,
                                double *udiag, double *dx, uint32_t *fail,
                                uint32_t *sizes);

/* Development hook, not part of the drop-in: the refactor and solve row
 * programs of the wavefront kernels (chain rows for runs of at least
 * min_chain multiply-subtracts on one target; 0: none), run on the CPU row by
 * row for `lanes` matrices a [lanes][nnz] and right-hand sides rhs
 * [lanes][n] on the tape factored from lead_vals with column order q, each
 * row checked for independent ops. Outputs as esl_debug_big_replay;
 * sizes[10] = {refactor rows, nnz_l, nnz_u, solve rows, refactor chains,
 * solve chains, refactor chain blocks, solve chain blocks, refactor and
 * solve critical paths} (lanes == 0: only sizes). ESL_FAILED when a row's
 * ops are not independent. Needs no device. */
esl_status esl_debug_row_replay(uint32_t n, const uint32_t *col_ptr,
                                const uint32_t *row_idx,
                                const double *lead_vals, const uint32_t *q,
                                double pivot_tol, uint32_t min_chain,
                                const double *a, const double *rhs,
                                uint32_t lanes, double *lx, double *ux,
                                double *udiag, double *dx, uint32_t *fail,
                                uint32_t *sizes);

/* ---- The LU boundary alone (arp_lane_refactor + arp_lane_solve,
 * src/solver/lane_device.zig:26-38; bodies lane_kernels.zig:49-143). ----
 * A tape is built on the host from one matrix (the lead lane) by the
 * threshold-pivoting Gilbert-Peierls factor (sparse_lu.zig:260-622) over
 * the caller's column order q; each lane then replays it on the GPU. */
typedef struct esl_tape esl_tape;
/* pivot_tol 1e-3 (direct.zig:32). SINGULAR (as ESL_FAILED) when the lead
 * matrix has no usable pivot in some non-void column. */
esl_status esl_tape_create(uint32_t n, const uint32_t *col_ptr,
                           const uint32_t *row_idx, const double *lead_vals,
                           const uint32_t *q, double pivot_tol,
                           esl_tape **out);
esl_status esl_tape_destroy(esl_tape *t);
/* nnz_l, nnz_u, flops (F), levels. */
esl_status esl_tape_sizes(esl_tape *t, uint32_t *nnz_l, uint32_t *nnz_u,
                          uint64_t *flops, uint32_t *levels);
/* The frozen tape: pinv [n], lp [n+1], li [nnz_l] (rows as pivot steps),
 * up [n+1], ui [nnz_u], flags [n] (bit 0 void column, bit 1 scaled pivot),
 * prow [nnz]. */
esl_status esl_tape_export(esl_tape *t, uint32_t *pinv, uint32_t *lp,
                           uint32_t *li, uint32_t *up, uint32_t *ui,
                           uint32_t *flags, uint32_t *prow);
/* For `lanes` matrices a [lanes][nnz] (pattern order) and right-hand sides
 * rhs [lanes][n]: replays the tape on the GPU (FP64, growth limit `growth`,
 * 1e-12 in Newton: direct.zig:53; wide bound 1/pivot_tol) and solves
 * dx = -A^-1 rhs. Outputs lx [lanes][nnz_l], ux [lanes][nnz_u],
 * udiag [lanes][n], dx [lanes][n], fail [lanes] (1 where the lane must
 * re-pivot: dead/non-finite pivot, growth, non-finite L/U, wide multiplier,
 * nonzero void slot). Any output may be NULL. */
esl_status esl_tape_replay(esl_tape *t, const double *a, const double *rhs,
                           size_t lanes, double growth, double *lx,
                           double *ux, double *udiag, double *dx,
                           uint32_t *fail);
/* No counterpart in the reference: esl_tape_replay for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and results (bitwise); the device staging of
 * the inputs and outputs lives in one fixed 32 MiB array per GPU that the
 * library holds until the process ends, and the call allocates nothing.
 * A call whose lanes need more than that (lanes * (nnz + 3n + nnz_l + nnz_u)
 * doubles plus lanes flags, each region rounded up to 256 bytes) returns
 * ESL_BUFFER_TOO_SMALL with a message; split the lanes. Calls that overlap
 * on one GPU share the array and race: serialize them. */
esl_status esl_tape_replay_noalloc(esl_tape *t, const double *a,
                                   const double *rhs, size_t lanes,
                                   double growth, double *lx, double *ux,
                                   double *udiag, double *dx, uint32_t *fail);

/* ---- Learned warm start (no counterpart in the reference) ----
 * A multilayer perceptron n_in -> n_hidden (x n_hidden_layers, SiLU) -> n_out
 * that maps a point's live parameter values to its operating point
 * (unknowns 1..n-1; ground is 0). Trained and evaluated on the device: BF16
 * MFMA products with FP32 accumulation, FP32 master weights, Adam, mean
 * squared error on standardized outputs. Parameters are ordered layer by
 * layer, each as W [out][in] row-major then b [out]. */
typedef struct esl_mlp esl_mlp;

typedef struct {
    size_t struct_size;    /* sizeof(esl_mlp_train_options) */
    uint32_t epochs;       /* passes over the training rows */
    uint32_t batch;        /* rows per Adam step */
    float lr, lr_final;    /* learning rate, cosine-decayed to lr_final */
    float beta1, beta2, eps, weight_decay;
    uint64_t seed;         /* row shuffling */
} esl_mlp_train_options;

void esl_mlp_default_train_options(esl_mlp_train_options *opt);
/* Weights drawn uniform in +-sqrt(6 / (fan_in + fan_out)) from seed,
 * biases zero. Scaling defaults to identity. */
esl_status esl_mlp_create(uint32_t n_in, uint32_t n_hidden, uint32_t n_hidden_layers,
                          uint32_t n_out, uint64_t seed, esl_mlp **out);
esl_status esl_mlp_destroy(esl_mlp *m);
esl_status esl_mlp_param_count(esl_mlp *m, size_t *count);
esl_status esl_mlp_get_params(esl_mlp *m, float *params, size_t count);
esl_status esl_mlp_set_params(esl_mlp *m, const float *params, size_t count);
/* Inputs are standardized as (v - in_mean) / in_scale, outputs produced as
 * y * out_scale + out_mean. Scales must be positive and finite. */
esl_status esl_mlp_set_scaling(esl_mlp *m, const double *in_mean, const double *in_scale,
                               const double *out_mean, const double *out_scale);
/* One forward and backward pass over rows (inputs [rows][n_in], targets
 * [rows][n_out + 1] operating points with ground first, as esl_op_solve
 * writes them); loss = mean squared error of the standardized outputs;
 * grads [param_count]. No update. */
esl_status esl_mlp_gradient(esl_mlp *m, const double *inputs, const double *targets,
                            size_t rows, float *loss, float *grads);
/* Adam over the rows (same layouts); loss_per_epoch [epochs] may be NULL. */
esl_status esl_mlp_train(esl_mlp *m, const double *inputs, const double *targets,
                         size_t rows, const esl_mlp_train_options *opt,
                         float *loss_per_epoch);
/* out [rows][n_out + 1]: ground 0, then the network's unknowns. */
esl_status esl_mlp_predict(esl_mlp *m, const double *inputs, size_t rows, double *out);
/* Attaches a network (n_in = n_live, n_out = n - 1) and the nominal
 * operating point x_nominal [n] for ESL_START_LEARNED. The engine keeps
 * references: the network must outlive its use; NULL detaches. */
esl_status esl_engine_set_warm_start(esl_engine *e, esl_mlp *m, const double *x_nominal);

#ifdef __cplusplus
}
#endif
#endif
