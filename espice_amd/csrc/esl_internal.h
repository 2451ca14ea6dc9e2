// Layouts shared by the host compiler of a deck (host.cpp) and the lane
// kernels (lanes.hip). Everything here lives in device memory, read-only to
// the kernels and shared by every lane.
//
// A deck program describes how one lane evaluates its devices: the devices
// in stamp order (kind-major, deck order within a kind), cut into chunks of
// 64 instances, one per thread of the lane's wavefront. Each thread stages
// its instance's contributions in LDS; a reduce task then sums the cells of
// one target (a pattern slot or a residual row) in instance order, so every
// slot sees its additions in the same order as a serial stamp.
//
// A tape program describes one frozen pivot sequence: where each pattern
// slot lands among the factor's slots (U values, then diagonals, then L
// values, then a discard slot), the refactor as columns grouped by level,
// and the two triangular solves as rows grouped by level. Every column (row)
// is processed by one thread in the serial order of its operations, so the
// parallel replay is bitwise the serial one.
#pragma once
#include <stdint.h>

#include "../../include/espice_lanes.h"

namespace esl {

struct uint2_t {
    uint32_t x, y;
};
struct alignas(16) uint4_t {
    uint32_t x, y, z, w;
};

constexpr int kWave = 64;
constexpr int kStageEntries = 16;   // most cells one instance stages (Gummel-Poon: 12 + 4)
constexpr uint32_t kRhsBit = 0x80000000u;   // task target is a residual row
constexpr uint32_t kCBit = 0x40000000u;     // ... a charge-Jacobian slot (transient only)
constexpr uint32_t kQBit = 0x20000000u;     // ... a charge row (transient only)
constexpr uint32_t kTargetMask = 0x1fffffffu;
constexpr uint32_t kVoidBit = 1u, kScaledBit = 2u;

enum Kind : uint32_t { kR = 0, kC = 1, kV = 2, kI = 3, kD = 4, kMos = 5, kQ = 6, kKinds = 7 };

// MOS1 per-instance constants (structure of arrays, index [field * n_mos + j]).
enum MosField : uint32_t {
    MF_TYPE = 0, MF_VTO, MF_VBIT_A, MF_VBIT_B, MF_VBIT_C, MF_GAMMA, MF_PHIT,
    MF_LAMBDA, MF_VT, MF_ISATS, MF_ISATD, MF_VCRITS, MF_VCRITD, MF_GMIN,
    MF_KP, MF_RATIO4, MF_M, MF_W, MF_LEFF, MF_COUNT
};
// Gummel-Poon fields (prepared on the host): is_t * area = IS * QF_IS_SCALE,
// bf_t = BF * QF_BFAC.
enum BjtField : uint32_t {
    QF_TYPE = 0, QF_IS, QF_IS_SCALE, QF_BF, QF_BFAC, QF_BR_T, QF_VT, QF_NFVT, QF_NRVT, QF_ISE, QF_NEVT, QF_ISC,
    QF_NCVT, QF_ISS, QF_NSVT, QF_INV_VAF, QF_INV_VAR, QF_INV_IKF, QF_INV_IKR, QF_NKF, QF_MGMIN, QF_VERT, QF_COUNT
};
// Diode fields (prepared on the host): is_t = IS * DF_IS_SCALE.
enum DioField : uint32_t {
    DF_IS = 0, DF_IS_SCALE, DF_VTE, DF_GMIN, DF_CJO, DF_VJ, DF_MT, DF_QEXP, DF_VFC, DF_F1, DF_INVF2, DF_F3, DF_TT,
    DF_BV, DF_IBV, DF_VT, DF_VTEBRK, DF_IKF, DF_ISR, DF_VTEREC,
    DF_COUNT
};

struct Chunk {
    uint32_t kind, first, count, task_begin;
};

struct DeckProg {
    uint32_t n, nnz, n_nodes, n_live;
    uint32_t n_chunks, n_mos, diag0_slot, pad0;
    const Chunk *chunks;        // [n_chunks + 1]: task_begin of the last is the task count
    const uint32_t *task_target;
    const uint32_t *task_cell;  // [n_tasks + 1] offsets into cells
    const uint16_t *cells;      // staging index e * 64 + t
    const uint32_t *diag_slot;  // [n]
    const uint8_t *current_row; // [n]
    // R: t[2], value, bind
    const uint32_t *r_t; const double *r_val; const int32_t *r_bind;
    // V: t[3] (p, n, branch), value, bind
    const uint32_t *v_t; const double *v_val; const int32_t *v_bind;
    // I: t[2], value, bind
    const uint32_t *i_t; const double *i_val; const int32_t *i_bind;
    // MOS1: t[4] (d, g, s, b), fields, bindings of vto and kp
    const uint32_t *m_t; const double *m_par; const int32_t *m_bind_vto; const int32_t *m_bind_kp;
    // first instance index of each kind in the stamp order
    uint32_t kind_first[kKinds + 1];
    // pattern slot of each Jacobian entry (kNoSlot for a ground entry):
    // R [4 * nR], V [4 * nV], MOS1 [11 * n_mos], in the stamp entry order
    const uint32_t *r_slot, *v_slot, *m_slot;
    // the pattern by rows, columns ascending: [n + 1], [nnz], [nnz] (pattern slot)
    const uint32_t *csr_ptr, *csr_col, *csr_slot;
    // diodes: t[2] (a, k), fields [DF_COUNT][n_dio], binding of IS
    uint32_t n_dio, n_cap, n_cslot, pad1;
    const uint32_t *d_t; const double *d_par; const int32_t *d_bind_is;
    // Gummel-Poon BJTs: t[4] (c, b, e, s), fields [QF_COUNT][n_bjt], bindings of IS and BF
    uint32_t n_bjt, pad2;
    const uint32_t *q_t; const double *q_par; const int32_t *q_bind_is, *q_bind_bf;
    // capacitors: t[2], value, bind
    const uint32_t *c_t; const double *c_val; const int32_t *c_bind;
    // V source waveforms: kind (0 DC, 1 step), v2, td per source
    const uint32_t *v_wkind; const double *v_v2, *v_td;
    // charge-Jacobian slots: the k-th is pattern slot cslot_pat[k]
    const uint32_t *cslot_pat;
    // the pattern by columns ([n + 1], [nnz] rows) and the column order
    // of the factors: a lane's own re-pivot reads them
    const uint32_t *csc_ptr, *csc_row, *colq;
    // pattern slot of each Jacobian entry (kNoSlot for a ground entry):
    // diodes [4 * n_dio] (a-a, a-k, k-a, k-k), Gummel-Poon [12 * n_bjt] in
    // the entry order of their evaluation (the lane-per-thread kernel)
    const uint32_t *d_slot, *q_slot;
    // the refinement residual's stream [res_steps + kRing][64] and the
    // length of a lane's packed FP64 Jacobian (stream words, then the
    // chains' entries, each chain padded to 64)
    uint32_t res_steps, res_total;
    const uint32_t *res_w;
    // the residual's chain rows, run after the stream by the whole
    // wavefront: {row, entries, first column word, first a64 position},
    // and their columns (each chain padded to 64 words)
    const uint4_t *res_ch;
    const uint32_t *res_cc;
    uint32_t n_res_ch, pad_rc;
};
constexpr uint32_t kNoSlot = 0xffffffffu;
// Residual-stream words: kResRow | i starts row i, kResRow | kResPad pads,
// anything else is an entry's column.
constexpr uint32_t kResRow = 0x80000000u, kResPad = 0x40000000u;

// Words of the refactor and solve row programs, uint2 each:
// x = d | (b << 16), y = c | (kind << 16) | (flags << 20); d, b, c are
// 16-bit LDS indices into [factor slots | solve vector]. One row holds one
// word per thread and every word of a row has the same kind, so a row runs
// as one branch-free wavefront step; rows run in order, and a word's operands
// were final in an earlier row. Idle threads of a row work on a private pad
// slot.
// A chain row (R_CHAIN) hands each thread a run of multiply-subtracts on one
// target, all of whose operands were final before the row: the thread keeps
// the target in a register and takes the run's terms in their order, its
// operand reads issued a block ahead, so the run costs a few cycles a term
// instead of an LDS round trip a row. Its word: x = d | (len << 16), y = seg
// | (R_CHAIN << 16), seg indexing the program's chain segments (word offset,
// blocks); the segment's terms are u32 words b | (c << 16) laid out
// [blocks + kChainAhead][64][kChainBlk] (terms past a thread's len are
// skipped). Chain rows sit only at the last position of a kRing block of
// rows, so the row loop runs them between blocks.
enum RowKind : uint32_t {
    R_NOP = 0,    // whole row idle (stream padding)
    R_MAC = 1,    // M[d] = M[d] - M[b] * M[c]
    R_MACZ = 2,   // M[d] = M[c] != 0 ? M[d] - M[b] * M[c] : M[d]
    R_DIV = 3,    // M[d] = M[d] / M[b]
    R_CHAIN = 4,  // a run of R_MAC (refactor) or R_MACZ (solve) terms per thread
    R_KINDS = 5
};
// A chain row may instead run its chains one after another on the whole
// wavefront (flag kChainBcast in y): its segment (header word offset,
// chains) points at headers {d, terms, term word offset, 0} and each chain's
// terms b | (c << 16), padded to a multiple of 64 words. Thread k loads term
// k of a round of 64 and forms its product; the multiply-subtracts then run
// in order on the broadcast products, so the target rounds as the ops one by
// one (a skipped R_MACZ term subtracts +0, which leaves any value unchanged).
constexpr uint32_t kChainBcast = 1u << 22;
constexpr int kChainBlk = 8;     // terms of a chain block (one thread's 32 bytes)
constexpr int kChainAhead = 2;   // blocks of term words in flight (and padded at a segment's end)
constexpr uint32_t kFlagUnscaled = 1;   // R_DIV on an unscaled column: growth + wide tests
constexpr uint32_t kFlagScaled = 2;     // R_DIV on a scaled column: non-finite test
inline constexpr uint32_t row_kind(uint32_t y) { return (y >> 16) & 0xfu; }
inline constexpr uint32_t row_flags(uint32_t y) { return (y >> 20) & 0x3u; }
// Reduce-stream words: a target word sets the accumulator's slot (bit 30:
// residual row, else factor slot); a cell word adds one staged value.
constexpr uint32_t kRedTarget = 0x80000000u, kRedRhs = 0x40000000u;
// A reduce stream's padding word: a target word for the lane's scratch
// double at rhs + 2n (after x), which nothing else reads.
__host__ __device__ inline uint32_t red_pad(uint32_t n) { return kRedTarget | kRedRhs | (2 * n); }
constexpr int kRing = 16;      // rows in flight; streams carry kRing pad rows
// 16-bit cell indices of a chunk's reduce chains: they fill the 64 pad
// slots of the factor area
constexpr uint32_t kRedChainCells = 256;

// Doubles of a lane's working set in LDS: tv_cap factor slots, the Y block
// (y, sdiag, xold after the assembly; during it the stamp staging areas,
// one per wavefront of the lane's crew), rhs, x, and a scratch pair that
// the assembly reduce's padding words write (rhs + 2n). A transient lane's
// Y block holds y and xold, and q and the charge-Jacobian cells follow x.
constexpr uint32_t kStageDoubles = kWave * kStageEntries;
__host__ __device__ inline uint32_t lds_y_block(uint32_t n, uint32_t nw = 1)
{
    const uint32_t stage = nw * kStageDoubles;
    return 3 * n > stage ? 3 * n : stage;
}
__host__ __device__ inline uint32_t lds_doubles(uint32_t tv_cap, uint32_t n, uint32_t nw = 1)
{
    return tv_cap + lds_y_block(n, nw) + 2 * n + 2;
}
__host__ __device__ inline uint32_t tran_y_block(uint32_t n, uint32_t nw = 1)
{
    const uint32_t stage = nw * kStageDoubles;
    return 2 * n > stage ? 2 * n : stage;
}
__host__ __device__ inline uint32_t tran_doubles(uint32_t tv_cap, uint32_t n, uint32_t n_cslot, uint32_t nw = 1)
{
    return tv_cap + tran_y_block(n, nw) + 3 * n + n_cslot;
}

// One block of a big tape's replay: a run of consecutive ops of the row
// program order whose slots fit LDS. The block gathers its m slots
// (blk_slots[slot_off ...]) into LDS positions 0..m-1, runs its rows (words
// from row row0 of blk_words, indices into [m slots | 64 pads]), then
// scatters the first w positions (the slots it writes) back.
struct BigBlock {
    uint32_t slot_off, m, w, rows;   // rows: a multiple of kRing
    uint32_t row0, pad0, pad1, pad2;
};

struct TapeProg {
    uint32_t n_tv, n_u, n_l, lx0;   // tv = [U | diag | L | discard | void extras | 64 pads]
    uint32_t discard, n_levels, n_flev, n_blev;
    uint32_t n_void, n_rlev, n_slev, pad_slot;   // n_rlev/n_slev: critical paths in rows
    uint32_t rw_rows, sw_rows, pad3, pad4;       // stream rows (multiples of kRing)
    double growth, wide;
    const uint32_t *amap;      // [nnz] pattern slot -> tv slot
    const uint32_t *pinv;      // [n] row -> pivot step
    const uint32_t *qinv;      // [n] column -> pivot step
    const uint32_t *flags;     // [n] per pivot step
    const uint32_t *lev_ptr;   // [n_levels + 1]
    const uint32_t *lev_cols;  // [n] pivot steps by level
    const uint32_t *up;        // [n + 1]
    const uint32_t *ui;        // [n_u]
    const uint32_t *lp;        // [n + 1]
    const uint32_t *fptr;      // [n + 1] first flop of each column
    const uint32_t *dst;       // [F] tv slot each flop updates
    const uint32_t *flev_ptr, *flev_rows;            // forward solve rows by level
    const uint32_t *lrow_ptr, *lrow_k, *lrow_slot;   // L by rows, ascending column
    const uint32_t *blev_ptr, *blev_rows;            // backward solve rows by level
    const uint32_t *urow_ptr, *urow_k, *urow_slot;   // U by rows, descending column
    const uint32_t *void_slots;                      // pattern slots a void pivot needs zero
    // row programs: [rows + kRing][64] words (the kRing tail rows are R_NOP)
    const uint2_t *rw;          // refactor
    const uint2_t *sw;          // forward + backward solve
    const uint32_t *red;        // assembly reduce words, per chunk [len + kRing][64]
    const uint32_t *red_off;    // [n_chunks] word offset of each chunk's stream
    const uint32_t *red_len;    // [n_chunks] rows of each chunk's stream (multiple of kRing)
    const uint32_t *diag_tv;    // [n] tv slot of each row's diagonal
    const uint32_t *li;         // [n_l] row (pivot step) of each L entry
    // big tapes (lanes whose slots live in HBM): the replay in blocks of at
    // most blk_m slots; the solves in chunks of columns, each chunk's L
    // (forward) or U (backward) entries at most sol_cap: columns
    // fch[c] .. fch[c + 1] and bch[c] .. bch[c + 1]
    uint32_t n_blk, blk_m, n_fch, n_bch;
    uint32_t sol_cap, pad5, pad6, pad7;
    const BigBlock *blk;        // [n_blk]
    const uint32_t *blk_slots;  // tv slots of each block's LDS positions
    const uint2_t *blk_words;   // rows of all blocks, [row][64]
    const uint32_t *fch, *bch;  // [n_fch + 1], [n_bch + 1]
    // chain segments of the refactor and solve programs: (word offset into
    // the terms, blocks) each, and their terms
    const uint2_t *rseg, *sseg;
    const uint32_t *rcw, *scw;
    // [res_total]: the factor slot each position of a lane's packed FP64
    // Jacobian copies (a pad slot for a row start, whose value is unused)
    const uint32_t *rsrc;
    // assembly reduce chains: the longest targets of each chunk, one per
    // thread; headers {target word as in the reduce stream, cells | index
    // offset << 16} from rc_off[c] (rc_count[c] of them), the chunk's
    // kRedChainCells 16-bit cell indices from word rc_coff[c], staged in
    // the pad slots while the chains run, rc_blocks[c] blocks of kChainBlk
    const uint2_t *rc_head;
    const uint32_t *rc_cells;
    const uint32_t *rc_off, *rc_count, *rc_coff, *rc_blocks;
};

// Per-lane results of the Newton kernel.
enum LaneStatus : uint32_t {
    kConverged = 0, kNoConvergence = 1, kSingular = 3, kRepivot = 100,
    kToFp64 = 101,  // FP32 factors failed on a fresh tape, or refinement stalled
    kNoRoom = 104   // a lane on its own LU met a launch whose LDS cannot hold its pattern
};

// Where a lane that left the kernel early resumes. fresh != 0: the host just
// gave the lane a tape built from its matrix at iteration `iter`.
// k0: the iteration the lane's current start began at; warm: 0 no learned
// start, 1 learned start active (f0 = its first residual max), 2 restarted
// from the nominal point.
struct LaneResume {
    uint32_t iter, init_fix, tape, fresh;
    double limit;
    uint32_t k0, warm;
    double f0;
};

// A lane's own LU (the device re-pivot, esl_privlu.h). cap == 0: off, a
// lane whose replay fails leaves for the host.
// Per-engine arrays, lane-strided; `cap` entries of L and of U per lane.
struct PrivLU {
    uint32_t cap, n, nnz, pad;
    uint32_t *hdr;                                  // [N][4]: nnz_l, nnz_u, n_void_slots, 0
    uint32_t *pinv, *flags, *flag, *topo, *stack, *pstack, *want, *want_of;   // [N][n]
    uint32_t *lp, *up;                              // [N][n + 1]
    uint32_t *prow, *void_slots;                    // [N][nnz]
    uint32_t *li, *ui;                              // [N][cap]
    double *lx, *ux;                                // [N][cap]
    double *udiag, *rscale;                         // [N][n]
    double *ap;                                     // [N][nnz] the lane's Jacobian, pattern order
};

constexpr uint32_t kPrivTape = 0xfffffffeu;   // LaneResume::tape of a lane on its own LU

// Arguments of the Newton kernel (op_newton).
// A lane's .op ladder (dc/op.zig:160-288): the gmin rung (g stepped down
// from kLadGminStart / 10 to the target by a factor that grows after easy
// runs and shrinks after hard or failed ones, failed runs restarting from
// the last converged point; then a clean run without the shunts), the
// source rung (lambda from 0 by a step that grows 1.5x and halves on
// failure) and a final run at full sources. Runs of the stepping rungs have
// at most kLadItl2 iterations, the clean and final runs the plain budget.
enum LadderPhase : uint32_t { kLadPlain = 0, kLadGmin, kLadGminClean, kLadSrc, kLadFinal, kLadDone };
constexpr double kLadGminTarget = 1e-12, kLadGminStart = 1e-2;
constexpr uint32_t kLadItl2 = 50;
// flags: a good point saved, the next gmin run starts cold, and the last
// run's factor met a singular matrix (the next run factors afresh instead of
// replaying, as direct.zig:268-298 leaves no factorization after one)
constexpr uint32_t kLadHaveGood = 1, kLadCold = 2, kLadNeedFactor = 4;
struct LadderState {
    double factor, good, gv;            // gmin rung
    double lambda, lambda_good, delta;  // source rung
    uint32_t phase, total, solves, flags;
};

struct OpArgs {
    DeckProg dp;
    const TapeProg *tapes;
    const uint32_t *lane_ids;   // [grid] the lane each workgroup solves
    const double *live;         // [N][n_live]
    double *x;                  // [N][n]: start, resume and final iterate
    double *lim;                // [N][4][n_mos] limited (vgs, vds, vbs) and von
    LaneResume *resume;         // [N]
    uint32_t *status;           // [N]
    uint32_t *iterations;       // [N]
    double *a_out;              // dump: [grid][nnz] Jacobian in pattern order
    double *f_out;              // dump: [grid][n] residual
    double reltol, abstol, vntol, residual_tol;
    uint32_t max_iter, mode;    // mode: 0 cold, 1 given x, 2 resume, 3 learned x + nominal fallback
    uint32_t tv_cap, dump;
    const double *x_nom;        // [n] mode 3: the nominal operating point
    uint8_t *fell_back;         // [N] mode 3: 1 where the lane restarted from x_nom
    // FP32 factors with FP64 refinement: lanes with lane64[lane] == 0 factor
    // in FP32; a64 [N][res_total + 64] holds their FP64 Jacobian for the
    // residual, packed in the residual stream's order
    uint32_t fp32, refine_max;
    double refine_tol;
    const uint8_t *lane64;      // [N]
    double *a64;                // [N][dp.res_total + 64]
    // optional: [N][kProfPhases] shader cycles per phase (assemble, refactor,
    // solve, update + limits + gates, the own LU's full factors), summed over
    // the lane's iterations, and the count of own-LU replays whose values
    // were in LDS
    unsigned long long *prof;
    // the lanes' own LU for re-pivots on the device (FP64 lanes)
    PrivLU priv;
    // optional [N]: the .op ladder's gmin rung, a conductance gmin[lane]
    // from every node row to ground (A_ii += g, F_i += g x_i) after each
    // assembly (Jacobian dumps included)
    const double *gmin;
    // 1: a lane that ends its run (converged, unconverged, singular) writes
    // the LU it is on back to resume[lane].tape, so that a later run of the
    // lane (the ladder's next rung) continues on the same pivots
    uint32_t keep_tape;
    // LDS doubles of one lane in the standalone kernels (set at launch): a
    // launch that leaves LDS idle gives the lanes' own LU room past the
    // working set
    uint32_t lane_lds;
    // optional [N]: the .op ladder's source rung, every independent source
    // scaled by src[lane] in the device evaluation (Jacobian dumps included)
    const double *src;
    // optional: the .op ladder on the device (FP64 lanes). A lane whose
    // plain run ends unconverged or singular climbs the gmin rung, the
    // source rung and the final run of dc/op.zig:160-288 in the same launch,
    // its control state in lad[lane] across exits and resumes, its last
    // converged point in xgood [N][n]
    LadderState *lad;
    double *xgood;
};
constexpr int kProfPhases = 6;

// Transient lanes (tran_steps): fixed-step integration, one wavefront per
// lane, state between launches in HBM. Step s (0-based) ends at
// t = (s + 1) dt; step 0 is backward Euler (ag0 = 1/dt, ag1 = 0), later
// steps the trapezoid with xmu = 0.5 (ag0 = 2/dt, ag1 = 1).
enum : uint32_t { kTranRunning = 102 };

// Per-lane step control of an adaptive transient lane (tran.zig Sim: the
// time, the step, the last two accepted steps, the order, counts and the
// breakpoint clamp).
struct AdaptLane {
    double t, dt, dt_old0, dt_old1, bp_next, bp_save_dt;   // bp_next: inf when none
    uint32_t order, steps, rejects, bp_clamped;
};

// An adaptive transient (TranArgs::adaptive): the lanes choose their steps
// (tran.zig:728-1086 for the trapezoid: orders 1 and 2, the LTE bound over
// the row plane of charges, accept/reject, order promotion, step-source
// breakpoints). q2, q3 and the accepted point xc per lane in HBM; the
// accepted times and probe rows go to times [N][max_steps] and
// TranArgs::probes [N][max_steps][n_probes].
struct AdaptArgs {
    AdaptLane *st;                 // [N]
    double *q2, *q3, *xc;          // [N][n]
    double *times;                 // [N][max_steps]
    const double *bps;             // step-source edges after 0, ascending
    double t_stop, dt_max, dt_min, trtol, chgtol, delmin, min_break;
    uint32_t n_bps, max_steps, attempts, any_charge;   // attempts: per launch
};

struct MeasState;
struct TranArgs {
    OpArgs op;                  // deck, tapes, lane ids, live, x, lim, resume, status, iterations, tolerances
    double *q1, *ip;            // [N][n] charges of the last accepted step, dynamic currents
    double *xp;                 // [N][n] the accepted state before the last one (the predictor's)
    uint32_t *step;             // [N] next step to run
    const uint32_t *probe_ids;  // [n_probes]
    double *probes;             // [N][steps][n_probes]
    double dt;
    uint32_t steps, n_probes, s0, s1;   // this launch runs steps [s0, s1)
    uint32_t init;              // 1: the first launch (q1 from the operating point, ip = 0)
    uint32_t dump;              // 1: only write the first step's first Jacobian to op.a_out
    uint32_t tran_tape;         // batch: the tape a lane's transient starts on
    uint32_t adaptive;          // 1: steps chosen per lane by `ad` (s0, s1, dt unused past the lead)
    AdaptArgs ad;
    // .meas cards fed as fixed steps are accepted (esl_meas.h): the cards,
    // per lane and card its state mst [N][n_meas], per lane the rows fed
    // (mrows [N], 0 until the operating point's row), the values at the
    // lane's end meas_out [N][n_meas]; meas_stop: a lane whose cards all
    // resolved stops stepping
    const esl_meas_card *meas;
    MeasState *mst;
    uint32_t *mrows;
    double *meas_out;
    uint32_t n_meas, meas_stop;
};

// The persistent batch: lanes of several templates (engines) in one launch
// of workgroups of kCrew wavefronts. A job is one lane: its operating point
// (op >= 0 names the .op arguments), then, when tran names transient
// arguments, its transient from that point. Each workgroup claims block
// jobs one at a time while any is left, each on the whole workgroup (lane
// per block: wavefront 0 leads, the others evaluate devices), then each
// of its wavefronts claims wave jobs, each in a kCrew-th of the LDS (lane
// per wavefront). The host orders both lists longest predicted first.
constexpr uint32_t kCrew = 4;
constexpr uint32_t kNoArgs = 0xffffffffu;
struct BatchJob {
    uint32_t op, tran, lane, flags;
};
struct BatchArgs {
    const OpArgs *ops;          // [.op argument sets]
    const TranArgs *trans;      // [transient argument sets]
    const BatchJob *block_jobs; // [n_block], longest predicted first
    const BatchJob *wave_jobs;  // [n_wave], longest predicted first; the first n_lead run before the block jobs
    uint32_t n_block, n_wave, n_lead, pad0;
    uint32_t *next;             // [3] claim counters (block, wave, lead): 0, n_lead, 0 at launch
    uint32_t wave_doubles;      // LDS doubles of one wavefront in the wave phase
    uint32_t pad;
    uint64_t *block_ticks, *wave_ticks;   // optional: each job's duration (wall-clock ticks)
};
constexpr uint32_t kStepOp = 0xffffffffu;   // step marker: the lane has not reached its transient

// Arguments of the lane-per-thread Newton kernel (op_newton_lpt): thread t
// solves lane lane_ids[t]; its state is lane-major, element i at i * N + t.
struct LptArgs {
    DeckProg dp;
    TapeProg T;
    const uint32_t *lane_ids;   // [N]
    uint32_t N, n_mos;
    const double *live_t;       // [n_live][N]
    double *tvg;                // [n_tv][N]
    double *rhsg, *xg, *xold, *yg, *sdiag;   // [n][N]
    double *limg;               // [4 * n_mos][N]
    double *x;                  // [lanes][n]: start (mode 1) and final iterate
    double *lim;                // [lanes][4][n_mos]: state handed to a re-pivot
    LaneResume *resume;         // [lanes]
    uint32_t *status, *iterations;   // [lanes]
    double reltol, abstol, vntol, residual_tol;
    uint32_t max_iter, mode;    // mode: 0 cold, 1 given x
    unsigned long long *prof;   // optional [lanes][kProfPhases]
};

// Arguments of the kernel for lanes whose working set is beyond a
// workgroup's LDS (big_lanes): one wavefront per lane, workgroup g runs lane
// ta.op.lane_ids[g], .op (tran == 0: the OpArgs of ta, modes 0-2, dump) or
// transient steps (tran == 1); a workgroup holds `waves` lanes, one per
// wavefront. The lane's vectors and factor slots live in HBM at
// ws + g * ws_stride: tv [tv_cap], then rhs, x, xold, sdiag, q [n
// each] and the charge-Jacobian cells [n_cslot]; LDS (lds doubles) holds
// the staging cells, a block of the replay, or the solve vector and a chunk
// of factor entries.
constexpr uint32_t kBigWaves = 4;   // most lanes (wavefronts) per big-lane workgroup
struct BigArgs {
    TranArgs ta;
    double *ws;
    uint64_t ws_stride;
    uint32_t N, tran;
    uint32_t tv_cap, lds;       // lds: doubles of one lane's LDS area
    uint32_t waves, pad;        // lanes per workgroup (1..kBigWaves)
};
__host__ __device__ inline size_t big_stride(uint32_t tv_cap, uint32_t n, uint32_t n_cslot)
{
    const size_t d = (size_t)tv_cap + 5 * (size_t)n + n_cslot;
    return (d + 15) / 16 * 16;   // 128-byte aligned lanes
}
// the LDS of a big lane: the staging cells, a replay block of blk_m slots
// and 64 pads, or the solve vector with a chunk of sol_cap entries and
// their indices
__host__ __device__ inline uint32_t big_lds_doubles(uint32_t n, uint32_t blk_m, uint32_t sol_cap)
{
    uint32_t d = blk_m + kWave;
    const uint32_t s = n + sol_cap + (sol_cap + 1) / 2;
    d = s > d ? s : d;
    return kStageDoubles > d ? kStageDoubles : d;
}

// Arguments of the LU-only replay (tape_replay).
struct ReplayArgs {
    TapeProg T;
    uint32_t n, nnz;
    const double *a, *rhs;
    double *lx, *ux, *udiag, *dx;
    uint32_t *fail;
};

}  // namespace esl
