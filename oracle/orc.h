/* The oracle: a CPU FP64 restatement of ESPice's sweep-lane algorithms.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the espice_amd package links, loads
 * or calls this library; only tests/, __graft_entry__.smoke() and the CPU
 * baseline leg of bench.py do, to check results against and to time the
 * CPU path.
 *
 * Restates (file:line in OmarSiwy/ESPice 1.2.0):
 *   - SparseLu.factor / refactorColumns / solve (src/solver/sparse_lu.zig:
 *     260-622, 916-1024, 1140-1184), without supernode panels, which the
 *     reference documents as bitwise the column loop (sparse_lu.zig:772-776);
 *   - the lane replay and its peel tests (src/solver/lane_kernels.zig:49-143);
 *   - Newton with the reference's gate order (src/solver/converger.zig:
 *     199-315, 454-522, 794-819, 336-346, 377-384);
 *   - MNA planes with a ground pin (src/analysis/Circuit.zig:570-573);
 *   - the device models of models/{resistor,vsource,isource,mos1,diode}.va
 *     with ngspice's DEVpnjlim / DEVfetlim / DEVlimvds limiters
 *     (dev/solvers/newton-raphson-convergence.md:36-69,151-162) applied as
 *     limited evaluation points plus the residual correction of
 *     src/device/eval.zig:837-841, and the MODEINITJCT seeds
 *     (src/device/eval.zig:1841-1853, models/mos1.va:282-304).
 *
 * Parity status of each part is stated beside its function below.
 */
#ifndef ORC_H
#define ORC_H
#include <stddef.h>
#include <stdint.h>

#define ORC_NONE 0xffffffffu

/* ---- math (orc_math.c) ---- */
double orc_exp(double x);
double orc_log(double x);

/* ---- sparse LU (orc_lu.c) ---- */
typedef struct {
    uint32_t n, nnz;
    uint32_t *q, *pinv, *prow;
    uint32_t *lp, *li, *up, *ui;   /* li/ui as pivot steps */
    double *lx, *ux, *udiag;
    uint8_t *void_col, *scaled;
    uint32_t *void_slots;
    uint32_t nnz_l, nnz_u, n_void_slots;
    uint32_t cap_l, cap_u;
    int factored;
    double *w, *y;                 /* n each, zero between calls (w) */
} orc_lu;

orc_lu *orc_lu_new(uint32_t n, uint32_t nnz);
void orc_lu_free(orc_lu *lu);
/* Full factor with threshold pivoting over column order q (copied).
 * 0 on success, -1 SingularMatrix. */
int orc_lu_factor(orc_lu *lu, const uint32_t *col_ptr, const uint32_t *row_idx,
                  const double *vals, const uint32_t *q, double pivot_tol);
/* refactorColumns: 0 ok, -1 SingularMatrix. */
int orc_lu_refactor(orc_lu *lu, const uint32_t *col_ptr, const double *vals,
                    double growth_limit);
/* The lane replay (lane_kernels.refactor): returns 1 when the lane must
 * re-pivot (bad pivot, growth, non-finite, wide multiplier, void slot). */
int orc_lu_lane_refactor(orc_lu *lu, const uint32_t *col_ptr, const double *vals,
                         double growth_limit, double wide);
/* x = A^-1 b (solve). */
void orc_lu_solve(orc_lu *lu, const double *b, double *x);

/* FP32 factors on a tape of orc_lu (the precision mode of the GPU lanes:
 * LU in FP32, residual and refinement in FP64). Parity pinned by tests: a
 * matrix whose entries are exact in FP32 and whose factorization rounds
 * nowhere (integer factors) gives the FP64 factors exactly; FP32 factors of
 * random systems solve to FP32 accuracy; refinement reaches the FP64
 * solution. */
typedef struct {
    uint32_t n, cap_l, cap_u;
    float *lx, *ux, *udiag, *w, *y;
} orc_lu32;
orc_lu32 *orc_lu32_new(uint32_t n);
void orc_lu32_free(orc_lu32 *f);
/* The lane replay of orc_lu_lane_refactor in FP32 on (float)vals; the
 * growth and wide tests are computed in FP64 on the FP32 values. Returns 1
 * when the lane must re-pivot. */
int orc_lu_lane_refactor32(const orc_lu *lu, orc_lu32 *f, const uint32_t *col_ptr,
                           const double *vals, double growth_limit, double wide);
/* x = A32^-1 b in FP32 (no negation). */
void orc_lu_solve32(const orc_lu *lu, orc_lu32 *f, const float *b, float *x);
/* FP32 replay of the tape built from lead_vals over a [lanes][nnz] and
 * x = A32^-1 (float)b. */
int orc_lu_replay32_export(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx,
                           const double *lead_vals, const uint32_t *q, double pivot_tol,
                           const double *a, const double *b, size_t lanes, double growth,
                           float *lx, float *ux, float *udiag, float *x, uint32_t *fail);
/* dx = -A^-1 rhs (lane_kernels.solveNeg). */
void orc_lu_solve_neg(orc_lu *lu, const double *rhs, double *dx);
uint64_t orc_lu_flops(const orc_lu *lu);

/* Flat C entry points for ctypes (tests): */
int orc_lu_factor_export(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx,
                         const double *vals, const uint32_t *q, double pivot_tol,
                         uint32_t *sizes /* nnz_l, nnz_u, n_void_slots */,
                         uint32_t *pinv, uint32_t *lp, uint32_t *li, uint32_t *up,
                         uint32_t *ui, uint32_t *flags, uint32_t *prow,
                         double *lx, double *ux, double *udiag,
                         size_t cap_l, size_t cap_u);
/* Replays the tape built from lead_vals on `lanes` matrices a [lanes][nnz]
 * and solves -rhs; outputs as esl_tape_replay. */
int orc_lu_replay_export(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx,
                         const double *lead_vals, const uint32_t *q, double pivot_tol,
                         const double *a, const double *rhs, size_t lanes, double growth,
                         double *lx, double *ux, double *udiag, double *dx, uint32_t *fail);
/* x = A^-1 b via a fresh full factor (for dense cross-checks). */
int orc_lu_solve_export(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx,
                        const double *vals, const uint32_t *q, double pivot_tol,
                        const double *b, double *x);

/* ---- circuits and Newton (orc_ckt.c) ---- */
typedef struct {
    uint32_t n_nodes, n_dev;
    const uint32_t *kind, *nodes;
    const double *params;
    uint32_t n_live, n_bind;
    const uint32_t *bind_dev, *bind_param, *bind_live;
} orc_deck;                        /* the layout of esl_deck */

typedef struct {
    double reltol, abstol, vntol, residual_tol;
    uint32_t max_iter, reserved;   /* reserved: flags (ORC_TOL_LADDER ...) */
} orc_tol;                         /* the layout of esl_tolerances */
/* orc_op_solve*: FP64 cold or given starts that fail the plain rung climb
 * the .op ladder's gmin and source-stepping rungs (orc_ckt.c op_ladder) */
#define ORC_TOL_LADDER 1u
/* for the tests: the ladder also after a converged plain rung, the ladder
 * without its gmin rung, the ladder without its source rung */
#define ORC_TOL_LADDER_ALWAYS 2u
#define ORC_TOL_LADDER_NO_GMIN 4u
#define ORC_TOL_LADDER_NO_SRC 8u
/* for the tests: transient steps start from the accepted point instead of
 * the predictor's extrapolation */
#define ORC_TOL_NO_PREDICT 16u

typedef struct orc_ckt orc_ckt;
/* NULL on a malformed deck (message in err). */
orc_ckt *orc_ckt_new(const orc_deck *deck, char *err, size_t err_cap);
void orc_ckt_free(orc_ckt *c);
uint32_t orc_ckt_n(const orc_ckt *c);
uint32_t orc_ckt_nnz(const orc_ckt *c);
void orc_ckt_pattern(const orc_ckt *c, uint32_t *col_ptr, uint32_t *row_idx,
                     uint32_t *diag_slots, uint8_t *current_row);

/* Assembles the Newton planes of one point at x with limiting state as a
 * cold start leaves it (seeded, iteration 0): g [nnz], rhs [n]. */
void orc_ckt_assemble_cold(orc_ckt *c, const double *live, double *g, double *rhs);
/* Assembles at x given (no limiting: a warm start's iteration 0). */
void orc_ckt_assemble_at(orc_ckt *c, const double *live, const double *x,
                         double *g, double *rhs);
/* Assembles at x (no limiting) with charges at source time t: g, rhs as
 * orc_ckt_assemble_at, cp [nnz] the charge Jacobian dq/dx (the c plane), q [n]
 * the row charges. */
void orc_ckt_assemble_charges(orc_ckt *c, const double *live, const double *x, double t, double *g, double *rhs,
                              double *cp, double *q);
/* The same with every independent source scaled by src (the source rung). */
void orc_ckt_assemble_scaled(orc_ckt *c, const double *live, const double *x, double src,
                             double *g, double *rhs);

/* Operating point of `count` points over `threads` host threads (0 = one).
 * q: the column order every lane's lead factor uses; lead tape from point 0's
 * first Jacobian, re-pivot per lane on replay failure, as the GPU lanes do.
 * start_x NULL: cold start (INITFIX + seeds); else [count][n] warm starts.
 * Outputs x [count][n], iterations, reasons (ESL_REASON_*), repivots. */
int orc_op_solve(orc_ckt *c, const double *values, size_t count,
                 const orc_tol *tol, const uint32_t *q, const double *start_x,
                 int threads, double *x, uint32_t *iterations, uint32_t *reasons,
                 uint32_t *repivots);

/* Options of orc_op_solve2. precision 0 FP64, 1 FP32 factors with FP64
 * iterative refinement of each Newton solve (refine until max|r| <=
 * refine_tol * max(max|F|, max_i sum_j |A_ij dx_j|); a stall (max|r| not
 * halving) or refine_max steps without success sends the lane to FP64 for
 * the rest of its Newton, redoing the current iteration). nominal_x [n]
 * (with start_x): start_x is a learned guess, and a lane whose residual max
 * does not fall in its first iteration, or that has not converged after
 * max_iter iterations, restarts from nominal_x (the first check's assembly
 * is discarded, not counted). fp64_lanes / fell_back [count] receive 1 for
 * lanes that fell back to FP64 / restarted from nominal_x (may be NULL). */
typedef struct {
    uint32_t precision, refine_max;
    double refine_tol;
    const double *nominal_x;
} orc_prec;
int orc_op_solve2(orc_ckt *c, const double *values, size_t count,
                  const orc_tol *tol, const orc_prec *prec, const uint32_t *q,
                  const double *start_x, int threads, double *x,
                  uint32_t *iterations, uint32_t *reasons, uint32_t *repivots,
                  uint32_t *fp64_lanes, uint32_t *fell_back);
/* Transient lanes, fixed step (the GPU's esl_tran_solve): each point's
 * operating point (cold, tol_op, lead tape from point 0), then `steps` steps
 * of dt (step 0 backward Euler, then the trapezoid with xmu = 0.5) with the
 * companion A = G + ag0 C, F += ag0 (q - q1) - ag1 i_prev
 * (integrator.zig:40-53, lane_kernels.zig:194-215), each a Newton solve with
 * the residual gate off; the lanes share the tape of the first converged
 * point's first transient Jacobian and re-pivot on a failed replay. After a
 * step: i = ag0 (q - q1) - ag1 i_prev, q1 = q, probes, limiting state reset
 * to the accepted point. probe_out [count][steps][n_probes]. Parity pinned by
 * tests: RC closed forms (a single RC charging, a ladder against the exact
 * trapezoid recurrence computed densely) and agreement with the GPU. */
int orc_tran_solve(orc_ckt *c, const double *values, size_t count, const orc_tol *tol_op,
                   const orc_tol *tol, double dt, uint32_t steps, const uint32_t *probes,
                   uint32_t n_probes, const uint32_t *q, int threads, double *x,
                   double *probe_out, uint32_t *iterations, uint32_t *reasons);
/* orc_tran_solve with traces of each step's first iterate (tests): first_x
 * [count][steps][n] the predicted x, first_lim [count][steps][n_inst][4] the
 * limiting state after the predictor's limit pass (diode: [0] the limited
 * junction voltage; MOS: vgs, vds, vbs, von); rows of step 0 stay as given */
int orc_tran_solve_traced(orc_ckt *c, const double *values, size_t count, const orc_tol *tol_op,
                          const orc_tol *tol, double dt, uint32_t steps, const uint32_t *probes,
                          uint32_t n_probes, const uint32_t *q, int threads, double *x, double *probe_out,
                          uint32_t *iterations, uint32_t *reasons, double *first_x, double *first_lim);
/* Adaptive transient lanes (tran.zig Sim.begin/end, 728-1086, trapezoid
 * method): per lane the operating point (tol_op), then steps whose dt the
 * lane chooses: the first dt = min(min(t_step, t_stop/100)/10, dt_max),
 * cut to a tenth of the first step-source edge, /10; Newton failure -> dt/8
 * at order 1; the LTE bound of lane_kernels.zig:230-269 (ngspice CKTterr,
 * trapezoid coefficient 0.08333333333, BE 0.5, over the row plane of
 * charges) rejects below 0.9 dt and sets the next dt (growth at most 2x,
 * at most dt_max); order 1 -> 2 when the trapezoid's bound beats 1.05 dt;
 * landing on a step-source edge drops to order 1. times [count][max_steps]
 * and probe_out [count][max_steps][n_probes] get the accepted points;
 * accepted / rejected count steps; reasons 1 when dt fell below dt_min or
 * max_steps ran out. Parity pinned by tests (RC closed forms, the
 * trapezoid's error law, the step-size law of the bound). */
typedef struct {
    double t_stop, t_step, dt_max, dt_min, trtol, chgtol;
    uint32_t max_steps, reserved;
} orc_tran_adapt;
int orc_tran_adaptive(orc_ckt *c, const double *values, size_t count, const orc_tol *tol_op, const orc_tol *tol,
                      const orc_tran_adapt *o, const uint32_t *probes, uint32_t n_probes, const uint32_t *q,
                      int threads, double *x, double *times, double *probe_out, uint32_t *accepted,
                      uint32_t *rejected, uint32_t *iterations, uint32_t *reasons);
/* .meas cards over a transient's rows (orc_meas.c): the subset WHEN,
 * TRIG/TARG, FIND (AT= or WHEN), AVG, MIN, MAX, PP, INTEG, RMS, each a pass
 * of src/output/measure.zig fed one row at a time as its Stream does.
 * Counts: ORC_MEAS_UNSET, ORC_MEAS_LAST or k > 0; at = 1e99 when unset. */
enum { ORC_MEAS_WHEN = 1, ORC_MEAS_TRIG_TARG, ORC_MEAS_FIND, ORC_MEAS_AVG, ORC_MEAS_MIN, ORC_MEAS_MAX,
       ORC_MEAS_PP, ORC_MEAS_INTEG, ORC_MEAS_RMS };
#define ORC_MEAS_UNSET (-1)
#define ORC_MEAS_LAST (-2)
#define ORC_MEAS_MAX_CARDS 64
typedef struct {
    uint32_t node, level_node, has_level_node, pad;   /* the vector; WHEN v=v2's level vector */
    double val, td, from, to, at;
    int32_t rise, fall, cross, pad2;
} orc_meas_clause;
typedef struct {
    uint32_t func, pad[3];
    orc_meas_clause a, b;   /* b: TARG, or FIND's WHEN */
} orc_meas_card;
typedef struct {
    double td, last, prev, prev2, prev_x, time;
    int32_t rise, fall, cross, want_rise, want_fall, want_cross, above, pending, first, found;
    uint32_t rows, pad;
} orc_meas_cross;
typedef struct {
    double px, pv, value;
    uint32_t rows, pad;
} orc_meas_at;
typedef struct {
    double m, m_at, xv, pv, px, span;
    int32_t first, pad;
} orc_meas_ext;
typedef struct {
    double px, py, p[4][2], first_x, cut_x, sum[3];
    uint32_t rows, np, count;
    int32_t inside, cut, pad;
} orc_meas_integ;
typedef struct {
    orc_meas_cross cr[2];
    orc_meas_at at;
    orc_meas_ext ext[2];
    orc_meas_integ integ;
    double ring[3][2], find_at;
    uint32_t ring_len, over[2], resolved;
} orc_meas_state;
void orc_meas_init(const orc_meas_card *c, orc_meas_state *s);
int orc_meas_feed(const orc_meas_card *c, orc_meas_state *s, double t, const double *x);
double orc_meas_value(const orc_meas_card *c, orc_meas_state *s);
size_t orc_meas_rows(const orc_meas_card *cards, uint32_t n_cards, const double *t, const double *x, size_t rows,
                     size_t stride, int early, double *out);
#endif
