// Lane kernels for gfx950: one wavefront (64 threads) per sweep point.
//
// The reference runs one GPU thread per lane and walks the whole tape
// serially in every thread (src/solver/lane_device.zig:26-38, bodies
// lane_kernels.zig:49-143; eval per lane src/device/eval.zig:2998-3042),
// with the lane's factors streaming through global memory and the host
// syncing twice per Newton tick (src/analysis/gpu_lanes.zig:782-832). Here a
// lane is a wavefront: its factor slots, residual and vectors sit in LDS;
// devices evaluate 64 at a time (one per thread) and reduce in instance
// order; the refactor and the solves run as row programs (64 independent
// updates per wavefront step, each slot's updates in the serial order of the
// column replay); the Newton loop, its gates and the limiting stay on the
// device until the lane converges, fails, or needs a new pivot sequence. In
// the FP32 mode the factors and solves run in FP32 on the same programs and
// each Newton solve is refined against the FP64 Jacobian.
//
// Compiled with -ffp-contract=off: no fused multiply-add anywhere, so every
// value is bitwise the serial CPU restatement's.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "esl_device.h"
#include "esl_eval.h"
#include "esl_internal.h"
#include "esl_meas.h"
#include "esl_privlu.h"
#include "eslmath.h"

namespace esl {

// A lane's resume record, uniform across its wavefront.
__device__ inline LaneResume uni_resume(const LaneResume &g)
{
    LaneResume r;
    r.iter = uni(g.iter);
    r.init_fix = uni(g.init_fix);
    r.tape = uni(g.tape);
    r.fresh = uni(g.fresh);
    r.limit = uni(g.limit);
    r.k0 = uni(g.k0);
    r.warm = uni(g.warm);
    r.f0 = uni(g.f0);
    return r;
}

// A lane-per-block lane's crew: wavefront 0 leads the lane; at each
// assembly round it posts the round's first chunk c0, every wavefront k of
// the crew evaluates chunk c0 + k into its own staging area, and the leader
// then reduces the round's chunks in chunk order. Sums stay in instance
// order, so the result does not depend on the crew's size.
struct CrewCtl {
    uint32_t cmd, c0, charges, x_off;
    uint32_t st_off, pad;
    double t, src;
};
__shared__ CrewCtl g_crew;
constexpr uint32_t kCmdEval = 1, kCmdDone = 2;

// Where a crew's helpers find the lane: x and the staging areas as offsets
// from the lane's working-set base.
struct Crew {
    uint32_t nw, x_off, st_off;
};

__device__ inline void crew_post(const Crew &cr, uint32_t c0, bool charges, double t, uint32_t tid,
                                 double src = 1.0)
{
    if (tid == 0) {
        g_crew.cmd = kCmdEval;
        g_crew.c0 = c0;
        g_crew.charges = charges ? 1u : 0u;
        g_crew.x_off = cr.x_off;
        g_crew.st_off = cr.st_off;
        g_crew.t = t;
        g_crew.src = src;
    }
    __syncthreads();
}

// The leader ends the lane: after this the followers have left crew_follow.
__device__ inline void crew_done(uint32_t tid)
{
    if (tid == 0) g_crew.cmd = kCmdDone;
    __syncthreads();
    __syncthreads();
}

// A crew wavefront other than the leader: evaluates its chunk of each round
// the leader posts until the leader ends the lane. base: the lane's working
// set (LDS, or the workgroup's slab for a spilled lane).
__device__ inline void crew_follow(const DeckProg &dp, const double *live, double *lim, double *base, uint32_t wv,
                                   uint32_t tid)
{
    for (;;) {
        __syncthreads();
        const CrewCtl c = g_crew;
        if (c.cmd != kCmdEval) break;
        const uint32_t k = c.c0 + wv;
        if (k < dp.n_chunks)
            eval_chunk(dp, dp.chunks[k], live, base + c.x_off, lim, dp.n_mos, base + c.st_off + wv * kStageDoubles,
                       tid, c.t, c.charges != 0, c.src);
        __syncthreads();
    }
    __syncthreads();
}

// acc plus the staged cells [c0, c1) in their order: the cells are read
// four at a time ahead of the adds, so the chain of adds does not wait on
// an LDS round trip per cell (the sums are the serial ones).
__device__ inline double add_cells(double acc, const double *stc, const uint16_t *cells, uint32_t c0, uint32_t c1)
{
    uint32_t cc = c0;
    for (; cc + 4 <= c1; cc += 4) {
        const double a0 = stc[cells[cc]], a1 = stc[cells[cc + 1]], a2 = stc[cells[cc + 2]], a3 = stc[cells[cc + 3]];
        acc = acc + a0;
        acc = acc + a1;
        acc = acc + a2;
        acc = acc + a3;
    }
    for (; cc < c1; cc++) acc = acc + stc[cells[cc]];
    return acc;
}

// Assembly into pattern slots (the Jacobian dump for a host re-pivot): the
// deck's reduce tasks, one target per thread.
__device__ void assemble_pattern(const DeckProg &dp, const double *live, const double *x, double *lim,
                                 uint32_t lim_stride, double *a, double *rhs, double *st, uint32_t tid,
                                 double src = 1.0)
{
    const uint32_t n = dp.n;
    for (uint32_t i = tid; i < dp.nnz; i += kWave) a[i] = 0.0;
    for (uint32_t i = tid; i < n; i += kWave) rhs[i] = 0.0;
    wsync();
    for (uint32_t c = 0; c < dp.n_chunks; c++) {
        const Chunk ch = dp.chunks[c];
        if (ch.kind == kC) continue;   // no static stamps
        eval_chunk(dp, ch, live, x, lim, lim_stride, st, tid, 0.0, false, src);
        wsync();
        const uint32_t t1 = dp.chunks[c + 1].task_begin;
        for (uint32_t t = ch.task_begin + tid; t < t1; t += kWave) {
            const uint32_t tgt = dp.task_target[t];
            if (tgt & (kCBit | kQBit)) continue;
            double *d = (tgt & kRhsBit) ? &rhs[tgt & ~kRhsBit] : &a[tgt];
            *d = add_cells(*d, st, dp.cells, dp.task_cell[t], dp.task_cell[t + 1]);
        }
        wsync();
    }
    if (tid == 0) {
        a[dp.diag0_slot] = a[dp.diag0_slot] + 1.0;
        rhs[0] = rhs[0] + x[0];
    }
    wsync();
}

// The gmin rung's shunts (conductance g) on an assembled Jacobian M whose
// row-i diagonal is M[diag[i]]: A_ii += g, F_i += g x_i on every node row
// (not row 0, the ground pin, nor branch rows), as the oracle's newton_run.
__device__ inline void add_gmin(double g, const DeckProg &dp, const uint32_t *diag, double *M, double *rhs,
                                const double *x, uint32_t tid)
{
    if (!(g > 0)) return;
    for (uint32_t i = 1 + tid; i < dp.n; i += kWave)
        if (!dp.current_row[i]) {
            M[diag[i]] = M[diag[i]] + g;
            rhs[i] = rhs[i] + g * x[i];
        }
    wsync();
}

// The LDS double a reduce word reads, without a branch: a target word its
// factor slot or rhs row (padding words: the lane's scratch double at
// rhs + 2n), a cell word its staged cell. Branch-free, so the reads of the
// words ahead stay in flight together.
__device__ inline double *red_src(uint32_t w, double *tv, double *rhs, double *stc)
{
    double *tb = (w & kRedRhs) ? rhs : tv;
    return ((int32_t)w < 0 ? tb : stc) + (w & 0xfffffu);
}

// The reduce chains of a chunk: thread k < chains adds its chain's cells
// to its target in their order, the target in a register, eight cells a
// block; the sums are the ones the reduce stream would form. A block's
// values are read while the block before adds, its indices a block earlier
// still; cells past the chain's end read as -0, which leaves any sum
// unchanged, so only the adds are serial. The chunk's
// 16-bit cell indices sit in the pad slots (pw); hd: the thread's header
// (threads without a chain add nothing to the scratch double), nb: blocks
// of the longest chain.
__device__ inline void red_chains(esl::uint2_t hd, uint32_t nb, const uint32_t *pw, double *tv, double *rhs,
                                  double *stc)
{
    const uint32_t len = hd.y & 0xffffu;
    const __attribute__((address_space(3))) uint32_t *ip = lptr(pw) + (hd.y >> 17);
    const __attribute__((address_space(3))) double *sv = lptr(stc);
    double *tgt = red_src(hd.x, tv, rhs, stc);
    auto idx = [&](uint32_t b, uint32_t (&ws)[4]) {
#pragma unroll
        for (int q = 0; q < 4; q++) ws[q] = ip[4 * b + q];
    };
    auto vals = [&](uint32_t b, const uint32_t (&ws)[4], double (&v)[kChainBlk]) {
#pragma unroll
        for (int e = 0; e < kChainBlk; e++) {
            const double u = sv[(ws[e >> 1] >> ((e & 1) * 16)) & 0xffffu];
            v[e] = b * kChainBlk + e < len ? u : -0.0;
        }
    };
    uint32_t w[4];
    double v[kChainBlk];
    idx(0, w);
    vals(0, w, v);
    if (nb > 1) idx(1, w);
    double acc = *tgt;
    for (uint32_t b = 0; b < nb; b++) {
        double u[kChainBlk];
#pragma unroll
        for (int e = 0; e < kChainBlk; e++) u[e] = v[e];
        if (b + 1 < nb) {
            vals(b + 1, w, v);
            if (b + 2 < nb) idx(b + 2, w);
        }
#pragma unroll
        for (int e = 0; e < kChainBlk; e++) acc = acc + u[e];
    }
    *tgt = acc;
}

// Assembly into the tape's factor slots: every chunk evaluates, then each
// thread runs its reduce words (coalesced rows, 16 rows in flight), adding
// each target's cells in instance order into a register. A crew of nw
// wavefronts evaluates nw chunks per round.
__device__ void assemble(const DeckProg &dp, const TapeProg &T, const double *live, const double *x, double *lim,
                         uint32_t lim_stride, double *tv, double *rhs, double *st, uint32_t tid, const Crew &cr,
                         double src = 1.0)
{
    const uint32_t n = dp.n;
    for (uint32_t i = tid; i < T.n_tv; i += kWave) tv[i] = 0.0;
    for (uint32_t i = tid; i < n; i += kWave) rhs[i] = 0.0;
    wsync();
    for (uint32_t c0 = 0; c0 < dp.n_chunks; c0 += cr.nw) {
        if (cr.nw > 1) crew_post(cr, c0, false, 0.0, tid, src);
        eval_chunk(dp, dp.chunks[c0], live, x, lim, lim_stride, st, tid, 0.0, false, src);   // capacitors stage nothing here
        if (cr.nw > 1) __syncthreads();
        else wsync();
        const uint32_t c1 = c0 + cr.nw < dp.n_chunks ? c0 + cr.nw : dp.n_chunks;
        for (uint32_t c = c0; c < c1; c++) {
            if (dp.chunks[c].kind == kC) continue;   // no static stamps
            double *stc = st + (c - c0) * kStageDoubles;
            // the stream through a global pointer: its ring of loads must
            // not hold up the LDS waits of the reduce (flat loads would)
            const __attribute__((address_space(1))) uint32_t *rs = gptr(T.red) + T.red_off[c] + tid;
            const uint32_t len = T.red_len[c];
            // the chains' headers and cell indices load now, ahead of use
            const uint32_t nc = T.rc_count[c];
            uint32_t tw0 = 0u, tw1 = 0u;
            esl::uint2_t hd{red_pad(n), 0u};
            if (nc) {
                const __attribute__((address_space(1))) uint32_t *cg = gptr(T.rc_cells) + T.rc_coff[c] + tid;
                tw0 = cg[0];
                tw1 = cg[kWave];
                if (tid < nc) {
                    const __attribute__((address_space(1))) esl::uint2_t &h = gptr(T.rc_head)[T.rc_off[c] + tid];
                    hd = esl::uint2_t{h.x, h.y};
                }
            }
            uint32_t q[kRing];
#pragma unroll
            for (int k = 0; k < kRing; k++) {
                q[k] = rs[k * kWave];
                __builtin_amdgcn_sched_barrier(0);
            }
            // Each word's value is read kAhead words before its use, so a
            // long target's chain of adds waits on the adds, not on an LDS
            // round trip each. The running sum is stored to the current
            // target at every word (its final value at the switch to the
            // next target): a chunk gives each target to one thread, once,
            // so no other read sees the partial sums. The sums are the
            // instance-ordered ones of the serial assembly.
            constexpr int kAhead = kRing / 2;
            double va[kAhead];
            double *pa[kAhead];
#pragma unroll
            for (int k = 0; k < kAhead; k++) {
                pa[k] = red_src(q[k], tv, rhs, stc);
                va[k] = *pa[k];
            }
            double *cur = rhs + 2 * n;   // the scratch double
            double acc = 0.0;
            for (uint32_t w = 0; w < len; w += kRing) {
#pragma unroll
                for (int k = 0; k < kRing; k++) {
                    const double v = va[k % kAhead];
                    double *p = pa[k % kAhead];
                    const bool tgt = (int32_t)q[k] < 0;
                    *cur = acc;
                    cur = tgt ? p : cur;
                    acc = tgt ? v : acc + v;
                    // word w + k + kAhead: still in the ring (k < kAhead), or
                    // the refill of the slot used kAhead words ago
                    pa[k % kAhead] = red_src(q[(k + kAhead) % kRing], tv, rhs, stc);
                    va[k % kAhead] = *pa[k % kAhead];
                    q[k] = rs[(w + kRing + k) * kWave];
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            *cur = acc;
            if (nc) {
                // the pad slots take the indices (the stream above reads no pad)
                uint32_t *pw = reinterpret_cast<uint32_t *>(tv + T.pad_slot);
                pw[tid] = tw0;
                pw[kWave + tid] = tw1;
                wsync();
                red_chains(hd, T.rc_blocks[c], pw, tv, rhs, stc);
            }
            wsync();
        }
    }
    if (tid == 0) {
        const uint32_t s0 = T.diag_tv[0];
        tv[s0] = tv[s0] + 1.0;
        rhs[0] = rhs[0] + x[0];
    }
    wsync();
}

// ---- the lane replay (lane_kernels.zig:49-118) on tv = [U | diag | L | ...]
// in precision R; true when the lane must re-pivot ----
template <typename R>
__device__ bool refactor(const TapeProg &T, uint32_t n, R *tv, uint32_t tid)
{
    LuFlags f{false, 0, 0};
    for (uint32_t v = tid; v < T.n_void; v += kWave) f.bad |= tv[T.amap[T.void_slots[v]]] != R(0);
    for (uint32_t k = tid; k < n; k += kWave)
        if (T.flags[k] & kVoidBit) tv[T.n_u + k] = R(1);
    run_rows<R>(gptr(T.rw), T.rw_rows, tv, T.growth, f, tid, T.pad_slot, T.rseg, T.rcw);
    for (uint32_t p = tid; p < T.n_u; p += kWave) {
        const R u = tv[p];
        f.nonfinite |= dbits((double)(u - u));
    }
    for (uint32_t k = tid; k < n; k += kWave) {
        if (T.flags[k] & kVoidBit) continue;
        const R d = tv[T.n_u + k];
        f.bad |= d == R(0) || !(fabs((double)d) < __builtin_inf());
    }
    const uint64_t worst = wave_max_u64(f.worst);
    const bool fail = f.bad || f.nonfinite != 0 || !(bitsd(worst) <= T.wide);
    return wave_any(fail);
}

// ---- dx = -A^-1 rhs into y (pivot order); row i's dx is y[qinv[i]]
// (lane_kernels.zig:123-143) ----
__device__ void solve_neg(const TapeProg &T, uint32_t n, double *tv, const double *rhs, double *y, uint32_t tid)
{
    for (uint32_t r = tid; r < n; r += kWave) y[T.pinv[r]] = -rhs[r];
    wsync();
    LuFlags f{false, 0, 0};
    run_rows<double, true>(gptr(T.sw), T.sw_rows, tv, 0.0, f, tid, T.pad_slot, T.sseg, T.scw);
}

// ---- FP32 factors + FP64 refinement of one Newton solve: dx = -A^-1 F.
// tv32 = [FP32 factor slots | y32]; r64 and dx64 are FP64 vectors; the FP64
// Jacobian a64 is packed in the residual stream's order (dp.res_w, then the
// chain rows' entries). r = -F - A dx row by row over the pattern in
// ascending column order. Stops when
// max|r| <= refine_tol * max(max|F|, max_i sum_j |A_ij dx_j|); a step that
// does not halve max|r|, or refine_max steps without success, is a stall.
// scratch: an LDS double the stream's padding rows read and write. ----
__device__ bool refine_solve(const OpArgs &a, const TapeProg &T, uint32_t n, float *tv32, const double *rhs,
                             const double *a64, double *dx, double *r, double *scratch, uint32_t tid)
{
    const DeckProg &dp = a.dp;
    float *y32 = tv32 + T.n_tv;
    uint64_t fm = 0;
    for (uint32_t i = tid; i < n; i += kWave) {
        r[i] = -rhs[i];
        dx[i] = 0.0;
        const uint64_t b = dbits(fabs(rhs[i]));
        fm = b > fm ? b : fm;
    }
    if (tid == 0) *scratch = 0.0;
    fm = wave_max_u64(fm);
    double prev = __builtin_inf();
    for (uint32_t s = 0; s < a.refine_max; s++) {
        wsync();
        for (uint32_t i = tid; i < n; i += kWave) y32[T.pinv[i]] = (float)r[i];
        wsync();
        LuFlags f{false, 0, 0};
        run_rows<float, true>(gptr(T.sw), T.sw_rows, tv32, 0.0, f, tid, T.pad_slot, T.sseg, T.scw);
        wsync();
        for (uint32_t i = tid; i < n; i += kWave) dx[i] = dx[i] + (double)y32[T.qinv[i]];
        wsync();
        // r = -F - A dx over the residual stream: a row-start word flushes
        // the previous row's sum into its r slot (the running sum is stored
        // at every word; a row belongs to one thread) and starts from -F_i;
        // an entry subtracts A_ij dx_j. The words' LDS values and the packed
        // Jacobian's values are read kAhead words ahead of their use.
        uint64_t rm = 0, am = 0;
        {
            const __attribute__((address_space(1))) uint32_t *rw = gptr(dp.res_w) + tid;
            const __attribute__((address_space(1))) double *ag = gptr(a64) + tid;
            auto src = [&](uint32_t w) -> const double * {
                const double *row = (w & kResPad) ? scratch : rhs + (w & 0x3fffffffu);
                return (int32_t)w < 0 ? row : dx + w;
            };
            uint32_t q[kRing];
#pragma unroll
            for (int k = 0; k < kRing; k++) {
                q[k] = rw[k * kWave];
                __builtin_amdgcn_sched_barrier(0);
            }
            constexpr int kAhead = kRing / 2;
            double va[kAhead], ga[kAhead];
            uint32_t off = 0;   // packed index of the step read ahead, less tid
#pragma unroll
            for (int k = 0; k < kAhead; k++) {
                va[k] = *src(q[k]);
                ga[k] = ag[off];
                off += (uint32_t)__popcll(__ballot(q[k] != (kResRow | kResPad)));
            }
            double *cur = scratch;
            double acc = 0.0, aa = 0.0;
            for (uint32_t st = 0; st < dp.res_steps; st += kRing) {
#pragma unroll
                for (int k = 0; k < kRing; k++) {
                    const uint32_t w = q[k];
                    const double v = va[k % kAhead], g = ga[k % kAhead];
                    const bool row = (int32_t)w < 0;
                    const double t = g * v;
                    *cur = acc;
                    const uint64_t rb = dbits(fabs(acc)), ab = dbits(aa);
                    rm = row && rb > rm ? rb : rm;
                    am = row && ab > am ? ab : am;
                    double *next = (w & kResPad) ? scratch : r + (w & 0x3fffffffu);
                    cur = row ? next : cur;
                    const double nacc = acc - t, naa = aa + fabs(t);
                    acc = row ? -v : nacc;
                    aa = row ? 0.0 : naa;
                    // step st + k + kAhead: still in the ring, or the refill
                    // of the slot used kAhead steps ago
                    const uint32_t wa = q[(k + kAhead) % kRing];
                    va[k % kAhead] = *src(wa);
                    ga[k % kAhead] = ag[off];
                    off += (uint32_t)__popcll(__ballot(wa != (kResRow | kResPad)));
                    q[k] = rw[(st + kRing + k) * kWave];
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            *cur = acc;
            const uint64_t rb = dbits(fabs(acc)), ab = dbits(aa);
            rm = rb > rm ? rb : rm;
            am = ab > am ? ab : am;
        }
        // the chain rows: thread k forms the product of entry k of a round
        // of 64 and writes it to pad k of the FP64 factor area (above the
        // FP32 factors), then every thread subtracts the round's products
        // in column order (padding subtracts +0, which changes nothing)
        if (dp.n_res_ch) {
            double *pads = reinterpret_cast<double *>(tv32) + T.pad_slot;
            const __attribute__((address_space(4))) esl::uint4_t *hs =
                (const __attribute__((address_space(4))) esl::uint4_t *)dp.res_ch;
            wsync();
            for (uint32_t ch = 0; ch < dp.n_res_ch; ch++) {
                const uint32_t row = hs[ch].x, m = hs[ch].y, cw = hs[ch].z, aw = hs[ch].w;
                double acc = -rhs[row], aa = 0.0;
                for (uint32_t r0 = 0; r0 < m; r0 += kWave) {
                    const uint32_t col = gptr(dp.res_cc)[cw + r0 + tid];
                    const double g = gptr(a64)[aw + r0 + tid];
                    const double t = g * dx[col];
                    pads[tid] = r0 + tid < m ? t : 0.0;
                    const uint32_t groups = ((m - r0 < (uint32_t)kWave ? m - r0 : (uint32_t)kWave) + 7u) >> 3;
                    for (uint32_t k8 = 0; k8 < groups; k8++) {
                        double v[8];
#pragma unroll
                        for (int e = 0; e < 8; e++) v[e] = pads[k8 * 8 + e];
#pragma unroll
                        for (int e = 0; e < 8; e++) {
                            acc = acc - v[e];
                            aa = aa + fabs(v[e]);
                        }
                    }
                }
                r[row] = acc;
                const uint64_t rb = dbits(fabs(acc)), ab = dbits(aa);
                rm = rb > rm ? rb : rm;
                am = ab > am ? ab : am;
            }
        }
        wsync();
        const double rmax = bitsd(wave_max_u64(rm));
        const double target = a.refine_tol * fmax(bitsd(fm), bitsd(wave_max_u64(am)));
        if (rmax <= target) return true;
        if (!(rmax <= 0.5 * prev)) return false;
        prev = rmax;
    }
    return false;
}

// ---- the .op ladder's control on the device (dc/op.zig:160-288, the
// oracle's op_ladder): the state of a lane between its runs ----
__device__ inline LadderState uni_ladder(const LadderState &g)
{
    LadderState r;
    r.factor = uni(g.factor);
    r.good = uni(g.good);
    r.gv = uni(g.gv);
    r.lambda = uni(g.lambda);
    r.lambda_good = uni(g.lambda_good);
    r.delta = uni(g.delta);
    r.phase = uni(g.phase);
    r.total = uni(g.total);
    r.solves = uni(g.solves);
    r.flags = uni(g.flags);
    return r;
}

// The run a phase makes: budget, the gmin rung's shunt and the source scale.
__device__ inline void ladder_run(const LadderState &L, uint32_t max_iter, double &g, double &lam, uint32_t &budget)
{
    if (L.phase == kLadPlain) return;
    g = L.phase == kLadGmin ? L.gv : 0.0;
    lam = L.phase == kLadSrc ? L.lambda : 1.0;
    budget = (L.phase == kLadGmin || L.phase == kLadSrc) ? kLadItl2 : max_iter;
}

// What follows a run that ended with status and its iterations: flags of
// kLadEnd (the lane's result is status, its iterations L.total for a lane
// that climbed), kLadSaveGood (the run's point becomes the last good one),
// kLadFromCold / kLadFromGood (the next run's start; neither: from the
// run's point). status becomes the final run's outcome.
constexpr uint32_t kLadEnd = 1, kLadSaveGood = 2, kLadFromCold = 4, kLadFromGood = 8;
__device__ inline uint32_t ladder_next(LadderState &L, uint32_t &status, uint32_t &iters)
{
    const bool ok = status == kConverged;
    if (L.phase == kLadPlain) {
        if (ok) return kLadEnd;
        L.factor = 10.0;
        L.good = kLadGminStart;
        L.gv = kLadGminStart / 10.0;
        L.lambda = 0.0;
        L.lambda_good = -1.0;
        L.delta = 0.25;
        L.phase = kLadGmin;
        L.total = iters;
        L.solves = 0;
        L.flags = kLadCold | (status == kSingular ? kLadNeedFactor : 0u);
        return kLadFromCold;
    }
    L.total += iters;
    uint32_t act = 0;
    switch (L.phase) {
    case kLadGmin:
        L.solves++;
        if (ok) {
            if (L.gv <= kLadGminTarget) {
                L.phase = kLadGminClean;   // the clean run, from this point
                break;
            }
            act |= kLadSaveGood;
            L.flags |= kLadHaveGood;
            L.good = L.gv;
            if (iters <= kLadItl2 / 4) L.factor = fmin(L.factor * sqrt(L.factor), 10.0);
            else if (iters > 3 * kLadItl2 / 4) L.factor = fmax(sqrt(L.factor), 1.00005);
            if (L.gv < L.factor * kLadGminTarget) {
                L.factor = L.gv / kLadGminTarget;
                L.gv = kLadGminTarget;
            } else {
                L.gv = L.gv / L.factor;
            }
            L.flags &= ~kLadCold;
        } else {
            if (L.factor < 1.00005) {
                L.phase = kLadSrc;
                L.solves = 0;
                break;
            }
            L.factor = sqrt(sqrt(L.factor));
            L.gv = L.good / L.factor;
            if (L.flags & kLadHaveGood) {
                act |= kLadFromGood;
                L.flags &= ~kLadCold;
            } else {
                L.flags |= kLadCold;
            }
        }
        if (L.phase == kLadGmin && L.solves >= 100) {
            L.phase = kLadSrc;
            L.solves = 0;
        }
        break;
    case kLadGminClean:
        if (ok) {
            L.phase = kLadDone;
        } else {
            L.phase = kLadSrc;
            L.solves = 0;
        }
        break;
    case kLadSrc:
        L.solves++;
        if (ok) {
            if (L.lambda >= 1.0) {
                L.phase = kLadFinal;
                break;
            }
            L.lambda_good = L.lambda;
            act |= kLadSaveGood;
            L.delta = L.delta * 1.5;
            L.lambda = fmin(L.lambda + L.delta, 1.0);
        } else {
            L.delta = L.delta * 0.5;
            if (L.delta < 1e-4 || L.lambda_good < 0.0) {
                L.phase = kLadFinal;   // from where the last run stopped
                break;
            }
            act |= kLadFromGood;
            L.lambda = fmin(L.lambda_good + L.delta, 1.0);
        }
        if (L.solves >= 100) L.phase = kLadFinal;
        break;
    case kLadFinal:
        L.phase = kLadDone;
        if (!ok) status = kNoConvergence;
        break;
    default:
        break;
    }
    if (L.phase == kLadDone) {
        iters = L.total;
        return act | kLadEnd;
    }
    if (status == kSingular) L.flags |= kLadNeedFactor;
    // a run of the gmin rung starts cold until one converged; a source-rung
    // run starts cold until a lambda converged
    if ((L.phase == kLadGmin && (L.flags & kLadCold)) || (L.phase == kLadSrc && L.lambda_good < 0.0))
        act |= kLadFromCold;
    return act;
}

// A lane's Jacobian (pattern order) and residual at its start or saved
// iterate: the lead point's matrix for the host factor, or a failing lane's
// for its re-pivot.
__global__ __launch_bounds__(kWave) void op_dump(OpArgs a)
{
    extern __shared__ double sm[];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = a.lane_ids[blockIdx.x];
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n, nm = dp.n_mos;
    double *tv = sm;
    double *st = tv + a.tv_cap;
    double *rhs = tv + a.tv_cap + lds_y_block(n);
    double *x = rhs + n;
    const double *live = a.live + (size_t)lane * dp.n_live;
    double *lim = a.lim + (size_t)lane * lim_width(dp);
    lane_init(a, lane, live, x, lim, tid);
    // a lane on the device ladder: its run's shunt and source scale
    double g = a.gmin ? a.gmin[lane] : 0.0, lam = a.src ? a.src[lane] : 1.0;
    uint32_t budget = a.max_iter;
    if (a.lad) ladder_run(uni_ladder(a.lad[lane]), a.max_iter, g, lam, budget);
    assemble_pattern(dp, live, x, lim, nm, tv, rhs, st, tid, lam);
    add_gmin(g, dp, dp.diag_slot, tv, rhs, x, tid);
    for (uint32_t p = tid; p < dp.nnz; p += kWave) a.a_out[(size_t)blockIdx.x * dp.nnz + p] = tv[p];
    for (uint32_t i = tid; i < n; i += kWave) a.f_out[(size_t)blockIdx.x * n + i] = rhs[i];
}

// One lane's operating point, led by one wavefront (the body of op_newton
// and of the batch kernel's .op jobs); nw > 1: a crew of nw wavefronts
// shares the device evaluation. Returns the lane's status.
template <bool kOwnLU>
__device__ __forceinline__ uint32_t op_lane(const OpArgs &a, uint32_t lane, double *base, uint32_t nw,
                                            uint32_t avail)
{
    const uint32_t tid = threadIdx.x % kWave;
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n, nm = dp.n_mos;
    // Working set: [tv (tv_cap) | Y block | rhs | x]. The lane's factor
    // slots start at tv and its solve vector follows at tv + n_tv (one index
    // space for the row programs); the Y block holds y, sdiag and xold after
    // the assembly and the stamp staging areas during it.
    const uint32_t yb = lds_y_block(n, nw);
    double *tv = base;
    double *sdiag = tv + a.tv_cap + n;
    double *xold = sdiag + n;
    double *st = tv + a.tv_cap;
    double *rhs = tv + a.tv_cap + yb;
    double *x = rhs + n;
    const Crew cr{nw, a.tv_cap + yb + n, a.tv_cap};
    // LDS past the working set (avail doubles from base): the own LU's values
    const uint32_t used = lds_doubles(a.tv_cap, n, nw);
    double *tail = base + used;
    const uint32_t tail_n = avail > used ? avail - used : 0u;
    const double *live = a.live + (size_t)lane * dp.n_live;
    double *lim = a.lim + (size_t)lane * lim_width(dp);
    lane_init(a, lane, live, x, lim, tid);


    LaneResume rs = uni_resume(a.resume[lane]);
    const double src = a.src ? uni(a.src[lane]) : 1.0;   // the source rung's scale
    const bool fp32 = a.fp32 && !uni((uint32_t)a.lane64[lane]);
    // FP64 lanes re-pivot on the device, on their own LU, when it is there
    const bool can_priv = kOwnLU && a.priv.cap != 0 && !fp32;
    bool priv = kOwnLU && rs.tape == kPrivTape;
    const TapeProg T = a.tapes[priv ? 0 : rs.tape];   // unused while the lane is on its own LU
    double *pap = a.priv.ap + (size_t)lane * dp.nnz;  // the lane's Jacobian for its own LU (pattern order)
    // FP64 solve vector; in FP32 lanes and on the lane's own LU, dx in row order
    double *y = priv ? tv + a.tv_cap : tv + T.n_tv;
    bool staged = true;
    if (priv) {
        // own LU: its dense column and pattern live at tv
        staged = priv_enter(a.priv, lane, dp.csc_ptr, dp.colq, tv, a.tv_cap, tail, tail_n, tid);
        wsync();
    }
    if (!staged) {
        // cannot happen with launches of one solve (they share tv_cap); the
        // point fails visibly rather than continue on other pivots
        if (tid == 0) a.status[lane] = kNoRoom;
        return kNoRoom;
    }
    float *tv32 = reinterpret_cast<float *>(tv);
    double *a64 = a.a64 + (size_t)lane * (dp.res_total + kWave);
    uint32_t iter0 = 0, k0 = 0, warm = a.mode == 3 ? 1u : 0u;
    bool init_fix = a.mode == 0;
    double limit = 1e9, f0 = 0.0;
    if (a.mode == 2) {
        iter0 = rs.iter;
        init_fix = rs.init_fix != 0;
        limit = rs.limit;
        k0 = rs.k0;
        warm = rs.warm;
        f0 = rs.f0;
    }
    // a learned start that stalls restarts from the nominal point: its
    // residual did not fall in the first iteration, or max_iter passed
    auto restart_nominal = [&](uint32_t at) {
        for (uint32_t i = tid; i < n; i += kWave) x[i] = a.x_nom[i];
        wsync();
        seed_lim_at(dp, live, x, lim, tid);
        wsync();
        k0 = at;
        warm = 2;
    };
    // the .op ladder (FP64 lanes, when on): a run's budget, shunt and source
    // scale follow the ladder's phase; the plain run is the first
    const bool ladder = a.lad != nullptr && !fp32;
    LadderState L{};
    double g_run = a.gmin ? uni(a.gmin[lane]) : 0.0, lam = src;
    uint32_t budget = a.max_iter;
    if (ladder) {
        if (a.mode == 2) L = uni_ladder(a.lad[lane]);
        else L.phase = kLadPlain;
        ladder_run(L, a.max_iter, g_run, lam, budget);
    }
    uint32_t status = kNoConvergence, iters = budget;
    // a lane the host found singular in its re-pivot resumes only to go on
    // with its ladder: its run ended singular at the saved iterate
    bool host_singular = ladder && a.mode == 2 && uni(a.status[lane]) == kSingular;
    unsigned long long pc[kProfPhases] = {};
    unsigned long long tc = clock64();
    auto lap = [&](int ph) {
        if (a.prof) {
            const unsigned long long now = clock64();
            pc[ph] += now - tc;
            tc = now;
        }
    };
    uint32_t iter = iter0;
    for (;;) {   // the runs of the lane's ladder (one run without it)
    for (;; iter++) {
        if (host_singular) {
            host_singular = false;
            status = kSingular;
            iters = iter + 1;
            break;
        }
        if (iter - k0 >= budget) {
            if (warm != 1) break;
            restart_nominal(iter);
        }
        lap(3);
        if (priv) {
            assemble_pattern(dp, live, x, lim, nm, pap, rhs, st, tid, lam);
            add_gmin(g_run, dp, dp.diag_slot, pap, rhs, x, tid);
        } else {
            assemble(dp, T, live, x, lim, nm, tv, rhs, st, tid, cr, lam);
            add_gmin(g_run, dp, T.diag_tv, tv, rhs, x, tid);
        }
        if (warm == 1) {
            uint64_t fb = 0;
            for (uint32_t i = tid; i < n; i += kWave) {
                const uint64_t b = dbits(fabs(rhs[i]));
                fb = b > fb ? b : fb;
            }
            const double fmax = bitsd(wave_max_u64(fb));
            if (iter == k0) {
                f0 = fmax;
            } else if (iter == k0 + 1 && !(fmax <= f0)) {
                restart_nominal(iter);
                iter--;   // this assembly is discarded, not an iteration
                continue;
            }
        }
        if (iter == k0) {
            double gm = 1.0;
            uint64_t xm = 0, fm = 0;
            for (uint32_t i = tid; i < n; i += kWave) {
                gm = fmax(gm, fabs(priv ? pap[dp.diag_slot[i]] : tv[T.diag_tv[i]]));
                const uint64_t xb = dbits(fabs(x[i])), fb = dbits(fabs(rhs[i]));
                xm = xb > xm ? xb : xm;
                fm = fb > fm ? fb : fm;
            }
            gm = wave_fmax(gm);
            xm = wave_max_u64(xm);
            fm = wave_max_u64(fm);
            const double scale = fmin(fmax(bitsd(xm), bitsd(fm) * gm), 1e18);
            limit = !(1e3 * scale > 1e9) ? 1e9 : 1e3 * scale;
        }
        for (uint32_t i = tid; i < n; i += kWave) sdiag[i] = priv ? pap[dp.diag_slot[i]] : tv[T.diag_tv[i]];
        wsync();
        lap(0);
        // FP32 lanes: keep the FP64 Jacobian for the residual, round the
        // factor slots to FP32 in place (the FP32 slots and y32 occupy the
        // low half of the FP64 slot area), replay and refine.
        uint32_t exit_status = kConverged;
        bool have_dx = false;
        if (fp32) {
            for (uint32_t k = tid; k < dp.res_total; k += kWave) a64[k] = tv[T.rsrc[k]];
            __threadfence_block();   // the refinement reads a64 back from other threads
            for (uint32_t sl = tid; sl < T.n_tv; sl += kWave) {
                const float v = (float)tv[sl];
                tv32[sl] = v;
            }
            wsync();
            if (refactor<float>(T, n, tv32, tid)) {
                exit_status = (rs.fresh && iter == iter0) ? kToFp64 : kRepivot;
            } else {
                lap(1);
                if (refine_solve(a, T, n, tv32, rhs, a64, y, xold, rhs + 2 * n, tid)) have_dx = true;
                else exit_status = kToFp64;
            }
        } else if (priv) {
            // the lane's own LU: replay it, factor again when a peel test fails
            int rc = kPlOk;
            const bool fresh = (L.flags & kLadNeedFactor) != 0;
            L.flags &= ~kLadNeedFactor;
            bool res = false;
            const bool failed =
                fresh || priv_replay_at(a.priv, lane, dp.csc_ptr, dp.colq, tv, a.tv_cap, tail, tail_n, tid, &res);
            lap(1);
            if (res) pc[5]++;
            if (failed) {
                rc = priv_factor_at(a.priv, lane, dp.csc_ptr, dp.csc_row, dp.colq, tv, a.tv_cap, tail, tail_n, tid);
                lap(4);
            }
            wsync();
            if (rc == kPlSingular) exit_status = kSingular;
            else if (rc == kPlFull) exit_status = kRepivot;
        } else if ((L.flags & kLadNeedFactor) || refactor<double>(T, n, tv, tid)) {
            // (a lane whose last factor met a singular matrix factors afresh)
            L.flags &= ~kLadNeedFactor;
            exit_status = kRepivot;
            if (can_priv) {
                // re-pivot here: the Jacobian again in pattern order (the same
                // sums as the failed assembly), factored by the lane
                wsync();
                assemble_pattern(dp, live, x, lim, nm, pap, rhs, st, tid, lam);
                add_gmin(g_run, dp, dp.diag_slot, pap, rhs, x, tid);
                for (uint32_t i = tid; i < n; i += kWave) tv[i] = 0.0;   // the factor's dense column (LDS)
                wsync();
                const int rc =
                    priv_factor_at(a.priv, lane, dp.csc_ptr, dp.csc_row, dp.colq, tv, a.tv_cap, tail, tail_n, tid);
                for (uint32_t i = tid; i < n; i += kWave) sdiag[i] = pap[dp.diag_slot[i]];
                wsync();
                if (rc == kPlOk) {
                    exit_status = kConverged;
                    priv = true;
                    rs.tape = kPrivTape;
                    y = tv + a.tv_cap;
                } else if (rc == kPlSingular) {
                    exit_status = kSingular;
                }
            }
        }
        if (exit_status == kSingular) {
            status = kSingular;
            iters = iter + 1;
            break;
        }
        if (exit_status != kConverged) {
            lap(1);
            for (uint32_t i = tid; i < n; i += kWave) a.x[(size_t)lane * n + i] = x[i];
            if (tid == 0) {
                LaneResume o;
                o.iter = iter;
                o.init_fix = init_fix ? 1u : 0u;
                o.tape = rs.tape;
                o.fresh = 0;
                o.limit = limit;
                o.k0 = k0;
                o.warm = warm;
                o.f0 = f0;
                a.resume[lane] = o;
            }
            if (a.a_out && exit_status == kRepivot) {
                // the Jacobian at this iterate (pattern order) for the host's
                // re-pivot; the device evaluation reads x and the limiting
                // state exactly as the failed assembly did
                wsync();
                assemble_pattern(dp, live, x, lim, nm, a.a_out + (size_t)lane * dp.nnz, rhs, st, tid, lam);
                add_gmin(g_run, dp, dp.diag_slot, a.a_out + (size_t)lane * dp.nnz, rhs, x, tid);
            }
            status = exit_status;
            iters = iter;
            break;
        }
        if (!have_dx && priv) {
            lap(1);
            priv_solve_at(a.priv, lane, dp.colq, rhs, y, tv, a.tv_cap, tail, tail_n, tid);
            wsync();
            have_dx = true;
        } else if (!have_dx) {
            lap(1);
            solve_neg(T, n, tv, rhs, y, tid);
        }
        lap(2);
        wsync();
        uint64_t worst = 0;
        for (uint32_t i = tid; i < n; i += kWave) {
            const double xo = x[i];
            const double d = have_dx ? y[i] : y[T.qinv[i]];
            const double xn = xo + d;
            xold[i] = xo;
            x[i] = xn;
            const double tl = a.reltol * fmax(fabs(xn), fabs(xo)) + (dp.current_row[i] ? a.abstol : a.vntol);
            const bool finite = fabs(xn) < __builtin_inf() && fabs(d) < __builtin_inf();
            const double q = fabs(d) / tl;
            const double score = !finite ? __builtin_inf() : (q == q ? q : 0.0);
            const uint64_t sb = dbits(score);
            worst = sb > worst ? sb : worst;
        }
        const double scaled = bitsd(wave_max_u64(worst));
        wsync();
        const bool limited = wave_any(limit_pass(dp, live, x, lim, tid));
        bool ok = !limited && iter != k0 && scaled < 1.0;
        if (ok) {
            bool miss = false;
            for (uint32_t i = tid; i < n; i += kWave) {
                const double xi = x[i];
                miss |= !(fabs(xi) <= limit);
                const double s0 = fabs(sdiag[i]);
                const double s = s0 == 0.0 ? 1.0 : s0;
                const double rt = fmax(a.residual_tol, 10.0 * s * (a.reltol * fabs(xi) + a.vntol));
                miss |= !(fabs(rhs[i]) <= rt);
            }
            ok = !wave_any(miss);
        }
        if (ok && init_fix) {
            init_fix = false;
        } else if (ok) {
            for (uint32_t i = tid; i < n; i += kWave) x[i] = xold[i];
            status = kConverged;
            iters = iter + 1;
            break;
        }
        wsync();
    }
    if (status == kNoConvergence) iters = iter;
    if (!ladder || (status != kConverged && status != kNoConvergence && status != kSingular)) break;
    // the run ended: the ladder's next run, or the lane's end
    wsync();
    const uint32_t act = ladder_next(L, status, iters);
    if (act & kLadSaveGood)
        for (uint32_t i = tid; i < n; i += kWave) a.xgood[(size_t)lane * n + i] = x[i];
    if (act & kLadEnd) break;
    if (act & kLadFromGood) {
        wsync();
        for (uint32_t i = tid; i < n; i += kWave) x[i] = a.xgood[(size_t)lane * n + i];
    }
    wsync();
    if (act & kLadFromCold) {
        lane_cold(dp, live, x, lim, tid);
    } else {
        seed_lim_at(dp, live, x, lim, tid);
        wsync();
    }
    ladder_run(L, a.max_iter, g_run, lam, budget);
    status = kNoConvergence;
    iters = budget;
    iter = 0;
    k0 = 0;
    warm = 0;
    init_fix = (act & kLadFromCold) != 0;
    limit = 1e9;
    f0 = 0.0;
    }
    lap(3);
    wsync();
    if (a.fell_back && tid == 0) a.fell_back[lane] = warm == 2 ? 1 : 0;
    if (status != kRepivot && status != kToFp64)
        for (uint32_t i = tid; i < n; i += kWave) a.x[(size_t)lane * n + i] = x[i];
    if (tid == 0) {
        a.status[lane] = status;
        a.iterations[lane] = iters;
        if (ladder) a.lad[lane] = L;   // a lane that leaves mid-ladder resumes its rung
        if (a.keep_tape && status != kRepivot && status != kToFp64) a.resume[lane].tape = rs.tape;
        if (a.prof)
            for (int p = 0; p < kProfPhases; p++) a.prof[(size_t)lane * kProfPhases + p] += pc[p];
    }
    return status;
}

// The Newton kernel; op_newton_own adds the lanes' own LU (a lane whose
// replay fails factors its Jacobian itself). The plain kernel keeps the
// registers of the common path: lanes that fail there resume in the other.
__global__ __launch_bounds__(kWave, 2) void op_newton(OpArgs a)
{
    extern __shared__ double sm[];
    op_lane<false>(a, uni(a.lane_ids[blockIdx.x]), sm, 1, 0);
}

__global__ __launch_bounds__(kWave) void op_newton_own(OpArgs a)
{
    extern __shared__ double sm[];
    op_lane<true>(a, uni(a.lane_ids[blockIdx.x]), sm, 1, a.lane_lds);
}

// ---- transient lanes ----

// Assembly of a transient Newton iterate: static stamps into tv (through the
// tape's slot map) and rhs, charge-Jacobian cells into cq, row charges into q
// (src/device/eval.zig:782-922 with the charge planes of LaneSink), the
// ground pin, then A = G + ag0 C and F += ag0 (q - q1) - ag1 i_prev
// (lane_kernels.zig:194-215). q1 and i_prev are the lane's rows in global
// memory. Jacobian entries go to tv[amap[slot]] (amap null: tv[slot], the
// pattern order). A crew of nw wavefronts evaluates nw chunks per round.
__device__ void assemble_tran(const DeckProg &dp, const uint32_t *amap, uint32_t n_slots, const double *live,
                              const double *x, double *lim, uint32_t nm, double *tv, double *rhs, double *q,
                              double *cq, double *st, double t, double ag0, double ag1, const double *q1,
                              const double *ip, uint32_t tid, const Crew &cr)
{
    const uint32_t n = dp.n;
    for (uint32_t i = tid; i < n_slots; i += kWave) tv[i] = 0.0;
    for (uint32_t i = tid; i < n; i += kWave) {
        rhs[i] = 0.0;
        q[i] = 0.0;
    }
    for (uint32_t k = tid; k < dp.n_cslot; k += kWave) cq[k] = 0.0;
    wsync();
    for (uint32_t c0 = 0; c0 < dp.n_chunks; c0 += cr.nw) {
        if (cr.nw > 1) crew_post(cr, c0, true, t, tid);
        eval_chunk(dp, dp.chunks[c0], live, x, lim, nm, st, tid, t, true);
        if (cr.nw > 1) __syncthreads();
        else wsync();
        const uint32_t c1 = c0 + cr.nw < dp.n_chunks ? c0 + cr.nw : dp.n_chunks;
        for (uint32_t c = c0; c < c1; c++) {
            const double *stc = st + (c - c0) * kStageDoubles;
            const uint32_t t1 = dp.chunks[c + 1].task_begin;
            for (uint32_t k = dp.chunks[c].task_begin + tid; k < t1; k += kWave) {
                const uint32_t tgt = dp.task_target[k];
                const uint32_t idx = tgt & kTargetMask;
                double *d = (tgt & kRhsBit) ? &rhs[idx] : (tgt & kQBit) ? &q[idx] : (tgt & kCBit) ? &cq[idx]
                                                                                                   : &tv[amap ? amap[tgt] : tgt];
                *d = add_cells(*d, stc, dp.cells, dp.task_cell[k], dp.task_cell[k + 1]);
            }
            wsync();
        }
    }
    if (tid == 0) {
        const uint32_t d0 = amap ? amap[dp.diag0_slot] : dp.diag0_slot;
        tv[d0] = tv[d0] + 1.0;
        rhs[0] = rhs[0] + x[0];
    }
    wsync();
    for (uint32_t k = tid; k < dp.n_cslot; k += kWave) {
        const uint32_t sl = amap ? amap[dp.cslot_pat[k]] : dp.cslot_pat[k];
        tv[sl] = tv[sl] + ag0 * cq[k];
    }
    for (uint32_t i = tid; i < n; i += kWave) {
        const double d = ag0 * (q[i] - q1[i]);
        const double base = rhs[i] + d;
        rhs[i] = base - ag1 * ip[i];
    }
    wsync();
}

// Runs steps [s0, s1) of each still-running lane: per step a Newton solve
// with the .op gate order (residual gate off: only a NaN row refuses, as
// lane_device.zig:74-98), then acceptance (published x = x_old), the
// dynamic currents i = ag0 (q - q1) - ag1 i_prev, q1 = q, the probe row,
// and limiting state reset to the accepted point. A lane whose replay fails
// leaves its Jacobian in a_out for a host re-pivot and resumes at the same
// iterate.
// |a - b| within `ulps` units in the last place (same sign), a == b included.
__device__ inline bool almost_equal_ulps(double a, double b, int64_t ulps)
{
    if (a == b) return true;
    if ((a < 0) != (b < 0)) return false;
    const int64_t ia = (int64_t)dbits(fabs(a)), ib = (int64_t)dbits(fabs(b));
    return (ia > ib ? ia - ib : ib - ia) <= ulps;
}

// The first step-source edge at or after t_from, or inf.
__device__ inline double next_break(const AdaptArgs &ad, double t_from)
{
    for (uint32_t k = 0; k < ad.n_bps; k++)
        if (ad.bps[k] >= t_from) return ad.bps[k];
    return __builtin_inf();
}

// The LTE bound of one step (integrator.stepBound, lane_kernels.zig:230-269)
// over the row plane of charges: method 0 backward Euler (coefficient 0.5),
// 1 the trapezoid (0.08333333333); cur_method the method whose dynamic
// current the step used. q0 (LDS) the step's charges, q1..q3 the accepted
// ring, ip the dynamic currents (HBM). NaN when any entry read is not
// finite. Every thread returns the lane's bound.
__device__ inline double step_bound(int method, int cur_method, const double *q0, const double *q1,
                                   const double *q2, const double *q3, const double *ip, uint32_t n,
                                   const AdaptArgs &ad, double reltol, double abstol, double dt, double dt1,
                                   double dt2, double ag0, double ag1, uint32_t tid)
{
    const double coeff = method == 0 ? 0.5 : 0.08333333333;
    const double sum01 = dt + dt1, sum12 = dt1 + dt2, sum012 = dt + (dt1 + dt2);
    double vmin = __builtin_inf();
    bool poison = false;
    for (uint32_t j = tid; j < n; j += kWave) {
        const double qc = q0[j], qp = q1[j], qp2 = q2[j], qp3 = q3[j], i_p = ip[j];
        const double i_new = cur_method == 1 ? ag0 * (qc - qp) - ag1 * i_p : ag0 * (qc - qp);
        const double volttol = abstol + reltol * fmax(fabs(i_new), fabs(i_p));
        const double chargetol = reltol * fmax(fmax(fabs(qc), fabs(qp)), ad.chgtol) / dt;
        const double tl = fmax(volttol, chargetol);
        poison |= !(fabs(qc) < __builtin_inf()) || !(fabs(qp) < __builtin_inf()) ||
                  !(fabs(qp2) < __builtin_inf()) || !(fabs(i_p) < __builtin_inf());
        const double f12 = (qp - qp2) / dt1;
        double dd = ((qc - qp) / dt - f12) / sum01;
        if (method != 0) {
            poison |= !(fabs(qp3) < __builtin_inf());
            const double f123 = (f12 - (qp2 - qp3) / dt2) / sum12;
            dd = (dd - f123) / sum012;
        }
        const double del = ad.trtol * tl / fmax(abstol, coeff * fabs(dd));
        poison |= del != del;
        vmin = fmin(vmin, del);
    }
    vmin = -wave_fmax(-vmin);
    if (wave_any(poison)) return __builtin_nan("");
    return method == 0 ? vmin : sqrt(vmin);
}

// A row of the lane's .meas cards (t, the unknowns x in LDS): thread k
// feeds card k (its state in HBM; the first row initialises it). True when
// every card has resolved.
__device__ inline bool meas_row(const TranArgs &ta, uint32_t lane, double t, const double *x, uint32_t tid)
{
    const bool first = uni(ta.mrows[lane]) == 0;
    bool open = false;
    for (uint32_t k = tid; k < ta.n_meas; k += kWave) {
        const esl_meas_card &c = ta.meas[k];
        MeasState &st = ta.mst[(size_t)lane * ta.n_meas + k];
        if (first) meas_init(c, st);
        open |= !meas_feed(c, st, t, x);
    }
    const bool all = !wave_any(open);
    if (tid == 0) ta.mrows[lane] = ta.mrows[lane] + 1;
    return all;
}

// The lane's measures at its end: meas_out [lane][k].
__device__ inline void meas_finish(const TranArgs &ta, uint32_t lane, uint32_t tid)
{
    const bool fed = uni(ta.mrows[lane]) != 0;
    for (uint32_t k = tid; k < ta.n_meas; k += kWave)
        ta.meas_out[(size_t)lane * ta.n_meas + k] =
            fed ? meas_value(ta.meas[k], ta.mst[(size_t)lane * ta.n_meas + k]) : __builtin_nan("");
}

template <bool kOwnLU>
__device__ __forceinline__ void tran_lane(const TranArgs &ta, uint32_t lane, double *base, uint32_t nw,
                                          uint32_t avail)
{
    const OpArgs &a = ta.op;
    const uint32_t tid = threadIdx.x % kWave;
    if (a.status[lane] != kTranRunning) return;
    uint32_t s = ta.step[lane];
    if (s >= ta.s1) return;
    if (s < ta.s0) s = ta.s0;
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n, nm = dp.n_mos;
    // Working set: [tv (tv_cap) | Y block (y, xold) | rhs | x | q | cq]
    const uint32_t yb = tran_y_block(n, nw);
    double *tv = base;
    double *xold = tv + a.tv_cap + n;
    double *st = tv + a.tv_cap;
    double *rhs = tv + a.tv_cap + yb;
    double *x = rhs + n;
    double *q = x + n;
    double *cq = q + n;
    double *q1 = ta.q1 + (size_t)lane * n;   // charges and dynamic currents of the last accepted step
    double *ip = ta.ip + (size_t)lane * n;
    double *xp = ta.xp + (size_t)lane * n;   // the accepted state before the last one
    const Crew cr{nw, a.tv_cap + yb + n, a.tv_cap};
    // LDS past the working set (avail doubles from base): the own LU's values
    const uint32_t used = tran_doubles(a.tv_cap, n, dp.n_cslot, nw);
    double *tail = base + used;
    const uint32_t tail_n = avail > used ? avail - used : 0u;
    const double *live = a.live + (size_t)lane * dp.n_live;
    double *lim = a.lim + (size_t)lane * lim_width(dp);
    LaneResume rs = uni_resume(a.resume[lane]);
    const bool can_priv = kOwnLU && a.priv.cap != 0;
    bool priv = kOwnLU && rs.tape == kPrivTape;
    const TapeProg T = a.tapes[priv ? 0 : rs.tape];   // unused while the lane is on its own LU
    double *pap = a.priv.ap + (size_t)lane * dp.nnz;  // the lane's Jacobian for its own LU (pattern order)
    double *y = priv ? tv + a.tv_cap : tv + T.n_tv;
    bool staged = true;
    if (priv) {
        // own LU: its dense column and pattern live at tv
        staged = priv_enter(a.priv, lane, dp.csc_ptr, dp.colq, tv, a.tv_cap, tail, tail_n, tid);
        wsync();
    }
    if (!staged) {
        // cannot happen with launches of one solve (they share tv_cap); the
        // point fails visibly rather than continue on other pivots
        if (tid == 0) a.status[lane] = kNoRoom;
        return;
    }
    // the companion assembly into the tape's slots, or in pattern order for
    // the lane's own LU
    auto assemble_step = [&](bool own, double tt, double g0, double g1) {
        if (own) assemble_tran(dp, nullptr, dp.nnz, live, x, lim, nm, pap, rhs, q, cq, st, tt, g0, g1, q1, ip, tid, cr);
        else assemble_tran(dp, T.amap, T.n_tv, live, x, lim, nm, tv, rhs, q, cq, st, tt, g0, g1, q1, ip, tid, cr);
    };
    for (uint32_t i = tid; i < n; i += kWave) x[i] = a.x[(size_t)lane * n + i];
    wsync();
    if (ta.init) {
        // start at the operating point: limiting state at its own voltages,
        // charges from one evaluation there, no dynamic current
        seed_lim_at(dp, live, x, lim, tid);
        for (uint32_t i = tid; i < n; i += kWave) {
            q1[i] = 0.0;
            ip[i] = 0.0;
            xp[i] = x[i];
        }
        wsync();
        assemble_step(priv, 0.0, 0.0, 0.0);
        for (uint32_t i = tid; i < n; i += kWave) q1[i] = q[i];
        if (ta.adaptive) {
            // the charge ring starts full of the operating point's charges
            double *q2 = ta.ad.q2 + (size_t)lane * n, *q3 = ta.ad.q3 + (size_t)lane * n;
            double *xc = ta.ad.xc + (size_t)lane * n;
            for (uint32_t i = tid; i < n; i += kWave) {
                q2[i] = q[i];
                q3[i] = q[i];
                xc[i] = x[i];
            }
        }
        wsync();
        if (ta.dump) {
            // the point's first transient Jacobian (step 0, iteration 0),
            // pattern order, in its row of a_out: a transient lead candidate
            assemble_tran(dp, nullptr, dp.nnz, live, x, lim, nm, a.a_out + (size_t)lane * dp.nnz, rhs, q, cq, st,
                          ta.dt, 1.0 / ta.dt, 0.0, q1, ip, tid, cr);
            return;
        }
    }
    wsync();
    uint32_t iter0 = rs.iter;   // nonzero only when resuming inside a step
    double limit = rs.limit;
    uint32_t status = kTranRunning;
    // the measures read the operating point's row first (t = 0); a lane
    // whose cards all resolve there takes no step when it stops early
    if (ta.n_meas && !ta.adaptive && uni(ta.mrows[lane]) == 0) {
        const bool all = meas_row(ta, lane, 0.0, x, tid);
        if (all && ta.meas_stop) status = kConverged;
        wsync();
    }
    uint32_t total_it = a.iterations[lane];
    // One step's Newton at source time t with companion coefficients ag0,
    // ag1, from iteration `iter` on: true when it converged (x is then the
    // published point, x_old of the converging iteration, and q its
    // charges); a failed replay without room for the lane's own factor, or
    // a singular matrix, sets status.
    auto newton = [&](double t, double ag0, double ag1, uint32_t &iter) -> bool {
        for (; iter < a.max_iter; iter++) {
            assemble_step(priv, t, ag0, ag1);
            if (iter == 0) {
                double gm = 1.0;
                uint64_t xm = 0, fm = 0;
                for (uint32_t i = tid; i < n; i += kWave) {
                    gm = fmax(gm, fabs(priv ? pap[dp.diag_slot[i]] : tv[T.diag_tv[i]]));
                    const uint64_t xb = dbits(fabs(x[i])), fb = dbits(fabs(rhs[i]));
                    xm = xb > xm ? xb : xm;
                    fm = fb > fm ? fb : fm;
                }
                gm = wave_fmax(gm);
                xm = wave_max_u64(xm);
                fm = wave_max_u64(fm);
                const double scale = fmin(fmax(bitsd(xm), bitsd(fm) * gm), 1e18);
                limit = !(1e3 * scale > 1e9) ? 1e9 : 1e3 * scale;
            }
            wsync();
            // the replay, then on a failed peel test the lane's own factor
            // (or, without room for it, the host's re-pivot)
            int rc = kPlOk;
            bool to_host = false;
            if (priv) {
                rc = priv_update_at(a.priv, lane, dp.csc_ptr, dp.csc_row, dp.colq, tv, a.tv_cap, tail, tail_n, tid);
                wsync();
                to_host = rc == kPlFull;
            } else if (refactor<double>(T, n, tv, tid)) {
                wsync();
                assemble_step(can_priv, t, ag0, ag1);   // the Jacobian again, pattern order when own
                if (can_priv) {
                    for (uint32_t i = tid; i < n; i += kWave) tv[i] = 0.0;   // the factor's dense column (LDS)
                    wsync();
                    rc = priv_factor_at(a.priv, lane, dp.csc_ptr, dp.csc_row, dp.colq, tv, a.tv_cap, tail, tail_n,
                                        tid);
                    wsync();
                    if (rc == kPlOk) {
                        priv = true;
                        rs.tape = kPrivTape;
                        y = tv + a.tv_cap;
                    }
                    to_host = rc == kPlFull;
                } else {
                    to_host = true;
                }
            }
            if (rc == kPlSingular) {
                status = kSingular;
                break;
            }
            if (to_host) {
                // the Jacobian for the host's re-pivot; resume at this iterate
                if (priv || can_priv)
                    for (uint32_t p = tid; p < dp.nnz; p += kWave) a.a_out[(size_t)lane * dp.nnz + p] = pap[p];
                else
                    for (uint32_t p = tid; p < dp.nnz; p += kWave) a.a_out[(size_t)lane * dp.nnz + p] = tv[T.amap[p]];
                if (tid == 0) {
                    LaneResume o = rs;
                    o.iter = iter;
                    o.limit = limit;
                    o.fresh = 0;
                    a.resume[lane] = o;
                }
                status = kRepivot;
                break;
            }
            if (priv) {
                priv_solve_at(a.priv, lane, dp.colq, rhs, y, tv, a.tv_cap, tail, tail_n, tid);
            } else {
                solve_neg(T, n, tv, rhs, y, tid);
            }
            wsync();
            uint64_t worst = 0;
            for (uint32_t i = tid; i < n; i += kWave) {
                const double xo = x[i];
                const double d = priv ? y[i] : y[T.qinv[i]];
                const double xn = xo + d;
                xold[i] = xo;
                x[i] = xn;
                const double tl = a.reltol * fmax(fabs(xn), fabs(xo)) + (dp.current_row[i] ? a.abstol : a.vntol);
                const bool finite = fabs(xn) < __builtin_inf() && fabs(d) < __builtin_inf();
                const double qq = fabs(d) / tl;
                const double score = !finite ? __builtin_inf() : (qq == qq ? qq : 0.0);
                const uint64_t sb = dbits(score);
                worst = sb > worst ? sb : worst;
            }
            const double scaled = bitsd(wave_max_u64(worst));
            wsync();
            const bool limited = wave_any(limit_pass(dp, live, x, lim, tid));
            bool ok = !limited && iter != 0 && scaled < 1.0;
            if (ok) {
                bool miss = false;
                for (uint32_t i = tid; i < n; i += kWave) {
                    miss |= !(fabs(x[i]) <= limit);
                    miss |= !(fabs(rhs[i]) <= __builtin_inf());
                }
                ok = !wave_any(miss);
            }
            wsync();
            if (ok) {
                for (uint32_t i = tid; i < n; i += kWave) x[i] = xold[i];
                iter++;
                return true;
            }
        }
        return false;
    };
    if (!ta.adaptive)
    for (; s < ta.s1 && status == kTranRunning; s++) {
        const double t = (double)(s + 1) * ta.dt;
        const double ag0 = s == 0 ? 1.0 / ta.dt : 2.0 / ta.dt;
        const double ag1 = s == 0 ? 0.0 : 1.0;
        uint32_t iter = iter0;
        const bool converged = newton(t, ag0, ag1, iter);
        if (status == kRepivot || status == kSingular) {
            total_it += iter - iter0;
            break;
        }
        total_it += iter - iter0;
        iter0 = 0;
        if (!converged) {
            status = kNoConvergence;
            break;
        }
        // accept: q holds the charges of the converged point (the last
        // assembly ran at x_old)
        wsync();
        for (uint32_t i = tid; i < n; i += kWave) {
            const double d = ag0 * (q[i] - q1[i]);
            ip[i] = d - ag1 * ip[i];
            q1[i] = q[i];
        }
        for (uint32_t k = tid; k < ta.n_probes; k += kWave)
            ta.probes[((size_t)lane * ta.steps + s) * ta.n_probes + k] = x[ta.probe_ids[k]];
        wsync();
        // the measures read the accepted row; with every card resolved the
        // lane may stop (CHANGELOGS.md:13-14)
        if (ta.n_meas && meas_row(ta, lane, t, x, tid) && ta.meas_stop) status = kConverged;
        // the limiting state at the accepted point (fetlim thresholds from
        // its last evaluation)
        seed_lim_at(dp, live, x, lim, tid, true);
        wsync();
        // the next step's first iterate: the last two accepted points
        // extrapolated by dt/dt_prev = 1 (MODEINITPRED, tran.zig:1131-1137;
        // lane_device.zig:141-153), then one limit pass of it against the
        // accepted point (applyLimits(trial, cur), tran.zig:1138)
        if (s + 1 < ta.steps && status == kTranRunning) {
            for (uint32_t i = tid; i < n; i += kWave) {
                const double cu = x[i];
                x[i] = cu + (cu - xp[i]);
                xp[i] = cu;
            }
            wsync();
            limit_pass(dp, live, x, lim, tid);
            wsync();
        }
    }
    if (ta.adaptive) {
        // per-lane step control (tran.zig:728-1086, the trapezoid: orders
        // 1 and 2), `attempts` attempts per launch
        const AdaptArgs &ad = ta.ad;
        AdaptLane A = ad.st[lane];
        double *q2 = ad.q2 + (size_t)lane * n, *q3 = ad.q3 + (size_t)lane * n;
        double *xc = ad.xc + (size_t)lane * n;
        for (uint32_t att = 0; att < ad.attempts && status == kTranRunning; att++) {
            if (!(A.t < ad.t_stop && !almost_equal_ulps(A.t, ad.t_stop, 100) && A.steps < ad.max_steps)) {
                const bool reached = A.t >= ad.t_stop || almost_equal_ulps(A.t, ad.t_stop, 100);
                status = reached ? kConverged : kNoConvergence;
                break;
            }
            const int method = A.order == 1 ? 0 : 1;
            const double ag0 = method == 0 ? 1.0 / A.dt : 1.0 / A.dt / (1.0 - 0.5);
            const double ag1 = method == 0 ? 0.0 : 0.5 / (1.0 - 0.5);
            if (iter0 == 0) {
                // the first iterate: the accepted points extrapolated by
                // dt/dt_prev (tran.zig:1131-1137), one limit pass against
                // the accepted point (tran.zig:1138)
                const double xfact = A.dt / A.dt_old0;
                for (uint32_t i = tid; i < n; i += kWave) x[i] = xc[i];
                wsync();
                seed_lim_at(dp, live, x, lim, tid, true);
                wsync();
                for (uint32_t i = tid; i < n; i += kWave) x[i] = x[i] + xfact * (x[i] - xp[i]);
                wsync();
                limit_pass(dp, live, x, lim, tid);
                wsync();
            }
            uint32_t iter = iter0;
            const bool converged = newton(A.t + A.dt, ag0, ag1, iter);
            total_it += iter - iter0;
            iter0 = 0;
            if (status == kRepivot || status == kSingular) break;
            if (!converged) {
                // Newton failed: dt / 8 at order 1 (dctran.c:815, 823)
                A.rejects++;
                A.order = 1;
                A.dt = A.dt / 8.0;
                if (A.dt < ad.dt_min) status = kNoConvergence;
                continue;
            }
            double dt_next = A.steps == 0 ? A.dt : fmin(A.dt * 2.0, ad.dt_max);
            if (ad.any_charge && A.steps > 0) {
                const double del = step_bound(method, method, q, q1, q2, q3, ip, n, ad, a.reltol, a.abstol, A.dt,
                                              A.dt_old0, A.dt_old1, ag0, ag1, tid);
                if (!(del >= 0.9 * A.dt)) {
                    A.rejects++;
                    if (del != del) {
                        A.order = 1;
                        A.dt = A.dt / 8.0;
                    } else {
                        A.dt = del;
                    }
                    if (A.dt < ad.dt_min) status = kNoConvergence;
                    continue;
                }
                dt_next = fmin(fmin(fmax(del, ad.dt_min), 2.0 * A.dt), ad.dt_max);
            }
            // accepted: order promotion (dctran.c:901-913)
            if (ad.any_charge && A.steps > 0 && A.order == 1) {
                const double trial = step_bound(1, method, q, q1, q2, q3, ip, n, ad, a.reltol, a.abstol, A.dt,
                                                A.dt_old0, A.dt_old1, ag0, ag1, tid);
                if (trial == trial) {
                    const double nd2 = fmin(2.0 * A.dt, trial);
                    if (nd2 > 1.05 * A.dt) A.order = 2;
                    dt_next = fmin(fmax(nd2, ad.dt_min), ad.dt_max);
                }
            }
            // the dynamic current under the method used, the ring, the points
            wsync();
            for (uint32_t i = tid; i < n; i += kWave) {
                const double d = ag0 * (q[i] - q1[i]);
                ip[i] = d - ag1 * ip[i];
                q3[i] = q2[i];
                q2[i] = q1[i];
                q1[i] = q[i];
                xp[i] = xc[i];
                xc[i] = x[i];
            }
            A.dt_old1 = A.dt_old0;
            A.dt_old0 = A.dt;
            A.t = A.t + A.dt;
            if (tid == 0) ad.times[(size_t)lane * ad.max_steps + A.steps] = A.t;
            for (uint32_t k = tid; k < ta.n_probes; k += kWave)
                ta.probes[((size_t)lane * ad.max_steps + A.steps) * ta.n_probes + k] = x[ta.probe_ids[k]];
            A.steps++;
            // landed on a breakpoint: order 1, resume at a tenth of the gap
            // (spice3 dctran's rule); then clamp to the next one
            if (A.bp_next < __builtin_inf()) {
                const bool landed = A.bp_clamped ? fabs(A.t - A.bp_next) <= ad.min_break
                                                 : (A.bp_next - A.t <= ad.delmin ||
                                                    almost_equal_ulps(A.t, A.bp_next, 100));
                if (landed) {
                    A.order = 1;
                    double shrink = A.bp_save_dt;
                    const double nb = next_break(ad, A.t + ad.min_break);
                    if (nb < __builtin_inf()) shrink = fmin(shrink, nb - A.t);
                    dt_next = fmin(dt_next, 0.1 * shrink);
                }
            }
            A.bp_next = next_break(ad, A.t + ad.min_break);
            A.bp_clamped = 0;
            if (A.bp_next < __builtin_inf() && A.bp_next - A.t < dt_next) {
                A.bp_save_dt = dt_next;
                dt_next = A.bp_next - A.t;
                A.bp_clamped = 1;
            }
            A.dt = dt_next;
            if (A.t + A.dt > ad.t_stop) A.dt = ad.t_stop - A.t;
            wsync();
        }
        if (status == kTranRunning &&
            !(A.t < ad.t_stop && !almost_equal_ulps(A.t, ad.t_stop, 100) && A.steps < ad.max_steps))
            status = (A.t >= ad.t_stop || almost_equal_ulps(A.t, ad.t_stop, 100)) ? kConverged : kNoConvergence;
        if (tid == 0) ad.st[lane] = A;
        // a lane that ends publishes its last accepted point; one that
        // runs on resumes from it (the next attempt predicts from xc)
        if (status != kTranRunning) {
            wsync();
            for (uint32_t i = tid; i < n; i += kWave) x[i] = xc[i];
        }
        wsync();
    }
    if (!ta.adaptive && status == kTranRunning && s >= ta.steps) status = kConverged;
    if (ta.n_meas && status != kTranRunning && status != kRepivot) meas_finish(ta, lane, tid);
    for (uint32_t i = tid; i < n; i += kWave) a.x[(size_t)lane * n + i] = x[i];
    if (tid == 0) {
        if (status != kRepivot) {
            LaneResume o = rs;
            o.iter = 0;
            a.resume[lane] = o;
        }
        ta.step[lane] = s;
        a.status[lane] = status;
        a.iterations[lane] = total_it;
    }
}

__global__ __launch_bounds__(kWave, 2) void tran_steps(TranArgs ta)
{
    extern __shared__ double sm[];
    tran_lane<false>(ta, uni(ta.op.lane_ids[blockIdx.x]), sm, 1, 0);
}

__global__ __launch_bounds__(kWave, 2) void tran_steps_own(TranArgs ta)
{
    extern __shared__ double sm[];
    tran_lane<true>(ta, uni(ta.op.lane_ids[blockIdx.x]), sm, 1, ta.op.lane_lds);
}

// ---- the persistent batch ----

// A batch job, uniform across its wavefront.
__device__ inline BatchJob uni_job(const BatchJob &g)
{
    BatchJob r;
    r.op = uni(g.op);
    r.tran = uni(g.tran);
    r.lane = uni(g.lane);
    r.flags = uni(g.flags);
    return r;
}

// One job on the lane's leader (nw > 1: with a crew): the operating point,
// then, for a transient job whose point converged, the hand-over to the
// transient tape and the steps.
__device__ __forceinline__ void run_job(const BatchArgs &b, const BatchJob &jb, double *base, uint32_t nw,
                                        uint32_t avail)
{
    const uint32_t tid = threadIdx.x % kWave;
    if (jb.op == kNoArgs) {
        // a transient resumed at its saved step and iterate
        tran_lane<true>(b.trans[jb.tran], jb.lane, base, nw, avail);
        return;
    }
    const OpArgs &oa = b.ops[jb.op];
    const uint32_t st = op_lane<true>(oa, jb.lane, base, nw, avail);
    if (jb.tran == kNoArgs || st != kConverged) return;
    const TranArgs &ta = b.trans[jb.tran];
    if (tid == 0) {
        LaneResume r{};
        r.tape = ta.tran_tape;
        r.limit = 1e9;
        ta.op.resume[jb.lane] = r;
        ta.op.status[jb.lane] = kTranRunning;
        ta.op.iterations[jb.lane] = 0;
        ta.step[jb.lane] = 0;
    }
    wsync();
    tran_lane<true>(ta, jb.lane, base, nw, avail);
}

__device__ __forceinline__ void follow_job(const BatchArgs &b, const BatchJob &jb, double *base, uint32_t wv)
{
    const OpArgs &oa = jb.op == kNoArgs ? b.trans[jb.tran].op : b.ops[jb.op];
    crew_follow(oa.dp, oa.live + (size_t)jb.lane * oa.dp.n_live, oa.lim + (size_t)jb.lane * lim_width(oa.dp), base,
                wv, threadIdx.x % kWave);
}

// Jobs are claimed at run time, longest predicted first: the wavefronts
// first take the lead jobs (lane-per-wavefront jobs of items whose slowest
// point outlasts the rest of the batch), then a workgroup takes the next
// lane-per-block job while any is left, then each of its wavefronts takes
// the next lane-per-wavefront job, so a job that runs longer than
// predicted delays only the slot that holds it.
__global__ __launch_bounds__(kWave * kCrew) void batch_run(BatchArgs b)
{
    extern __shared__ double sm[];
    __shared__ uint32_t claim;
    const uint32_t wv = threadIdx.x / kWave, tid = threadIdx.x % kWave;
    double *mine = sm + (size_t)wv * b.wave_doubles;
    for (;;) {
        uint32_t j = 0;
        if (tid == 0) j = atomicAdd(b.next + 2, 1u);
        j = uni(j);
        if (j >= b.n_lead) break;
        const uint64_t t0 = wall_clock64();
        run_job(b, uni_job(b.wave_jobs[j]), mine, 1, b.wave_doubles);
        if (b.wave_ticks && tid == 0) b.wave_ticks[j] = wall_clock64() - t0;
    }
    __syncthreads();
    // lane per block: the whole workgroup on one lane at a time
    for (;;) {
        if (threadIdx.x == 0) claim = atomicAdd(b.next, 1u);
        __syncthreads();
        const uint32_t j = claim;
        __syncthreads();
        if (j >= b.n_block) break;
        const BatchJob jb = uni_job(b.block_jobs[j]);
        if (wv == 0) {
            const uint64_t t0 = wall_clock64();
            run_job(b, jb, sm, kCrew, kCrew * b.wave_doubles);
            crew_done(tid);
            if (b.block_ticks && tid == 0) b.block_ticks[j] = wall_clock64() - t0;
        } else {
            follow_job(b, jb, sm, wv);
        }
    }
    // lane per wavefront: each wavefront on its own lanes in its share of the LDS
    for (;;) {
        uint32_t j = 0;
        if (tid == 0) j = atomicAdd(b.next + 1, 1u);
        j = uni(j);
        if (j >= b.n_wave) break;
        const uint64_t t0 = wall_clock64();
        run_job(b, uni_job(b.wave_jobs[j]), mine, 1, b.wave_doubles);
        if (b.wave_ticks && tid == 0) b.wave_ticks[j] = wall_clock64() - t0;
    }
}

// ---- the LU boundary alone: replay one tape over lanes of A, solve -rhs ----


__global__ __launch_bounds__(kWave) void tape_replay(ReplayArgs r)
{
    extern __shared__ double sm[];
    const uint32_t tid = threadIdx.x, lane = blockIdx.x;
    const TapeProg &T = r.T;
    const uint32_t n = r.n;
    double *tv = sm;
    double *y = tv + T.n_tv;
    double *rhs = y + n;
    for (uint32_t i = tid; i < T.n_tv; i += kWave) tv[i] = 0.0;
    wsync();
    for (uint32_t p = tid; p < r.nnz; p += kWave) tv[T.amap[p]] = r.a[(size_t)lane * r.nnz + p];
    for (uint32_t i = tid; i < n; i += kWave) rhs[i] = r.rhs[(size_t)lane * n + i];
    wsync();
    const bool bad = refactor<double>(T, n, tv, tid);
    solve_neg(T, n, tv, rhs, y, tid);
    if (r.lx)
        for (uint32_t q = tid; q < T.n_l; q += kWave) r.lx[(size_t)lane * T.n_l + q] = tv[T.lx0 + q];
    if (r.ux)
        for (uint32_t p = tid; p < T.n_u; p += kWave) r.ux[(size_t)lane * T.n_u + p] = tv[p];
    if (r.udiag)
        for (uint32_t k = tid; k < n; k += kWave) r.udiag[(size_t)lane * n + k] = tv[T.n_u + k];
    if (r.dx)
        for (uint32_t i = tid; i < n; i += kWave) r.dx[(size_t)lane * n + i] = y[T.qinv[i]];
    if (tid == 0 && r.fail) r.fail[lane] = bad ? 1u : 0u;
}

// out[p][k] = x[p][ids[k]]: a solve's measures.
__global__ void gather_probes(const double *x, uint32_t n, const uint32_t *ids, uint32_t k, uint32_t count,
                              double *out)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)count * k) return;
    const size_t p = e / k, j = e % k;
    out[e] = x[p * n + ids[j]];
}

}  // namespace esl

extern "C" hipError_t esl_launch_gather(const double *x, uint32_t n, const uint32_t *ids, uint32_t k, uint32_t count,
                                        double *out, hipStream_t s)
{
    const size_t total = (size_t)count * k;
    hipLaunchKernelGGL(esl::gather_probes, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s, x, n, ids, k, count,
                       out);
    return hipGetLastError();
}

extern "C" hipError_t esl_launch_op_newton(const esl::OpArgs *a, uint32_t grid, uint32_t lds_bytes,
                                           hipStream_t s)
{
    if (a->dump)
        hipLaunchKernelGGL(esl::op_dump, dim3(grid), dim3(esl::kWave), lds_bytes, s, *a);
    else if (a->priv.cap) {
        esl::OpArgs b = *a;
        b.lane_lds = lds_bytes / 8;
        hipLaunchKernelGGL(esl::op_newton_own, dim3(grid), dim3(esl::kWave), lds_bytes, s, b);
    }
    else
        hipLaunchKernelGGL(esl::op_newton, dim3(grid), dim3(esl::kWave), lds_bytes, s, *a);
    return hipGetLastError();
}

extern "C" hipError_t esl_launch_tape_replay(const esl::ReplayArgs *r, uint32_t lanes, uint32_t lds_bytes,
                                             hipStream_t s)
{
    hipLaunchKernelGGL(esl::tape_replay, dim3(lanes), dim3(esl::kWave), lds_bytes, s, *r);
    return hipGetLastError();
}

extern "C" hipError_t esl_launch_tran_steps(const esl::TranArgs *ta, uint32_t grid, uint32_t lds_bytes, hipStream_t s)
{
    if (ta->op.priv.cap) {
        esl::TranArgs b = *ta;
        b.op.lane_lds = lds_bytes / 8;
        hipLaunchKernelGGL(esl::tran_steps_own, dim3(grid), dim3(esl::kWave), lds_bytes, s, b);
    }
    else
        hipLaunchKernelGGL(esl::tran_steps, dim3(grid), dim3(esl::kWave), lds_bytes, s, *ta);
    return hipGetLastError();
}

extern "C" hipError_t esl_launch_batch(const esl::BatchArgs *b, uint32_t grid, uint32_t lds_bytes, hipStream_t s)
{
    hipLaunchKernelGGL(esl::batch_run, dim3(grid), dim3(esl::kWave * esl::kCrew), lds_bytes, s, *b);
    return hipGetLastError();
}

extern "C" hipError_t esl_batch_occupancy(uint32_t lds_bytes, int *blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, esl::batch_run, esl::kWave * esl::kCrew,
                                                        lds_bytes);
}

// The largest static LDS (the crew's control block) any lane kernel
// declares: a workgroup's dynamic LDS is the device's per-workgroup LDS
// less this.
extern "C" hipError_t esl_kernel_static_lds(uint32_t *bytes)
{
    const void *fns[] = {(const void *)esl::op_newton,     (const void *)esl::op_dump,
                         (const void *)esl::tran_steps,    (const void *)esl::tran_steps_own,
                         (const void *)esl::op_newton_own, (const void *)esl::batch_run,
                         (const void *)esl::tape_replay};
    uint32_t m = 0;
    for (const void *f : fns) {
        hipFuncAttributes a{};
        const hipError_t e = hipFuncGetAttributes(&a, f);
        if (e != hipSuccess) return e;
        m = a.sharedSizeBytes > m ? (uint32_t)a.sharedSizeBytes : m;
    }
    *bytes = m;
    return hipSuccess;
}

extern "C" hipError_t esl_kernel_attrs(uint32_t lds_bytes)
{
    hipError_t e = hipFuncSetAttribute((const void *)esl::op_newton,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void *)esl::op_dump, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void *)esl::tran_steps, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void *)esl::tran_steps_own, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void *)esl::op_newton_own, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_bytes);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void *)esl::batch_run, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void *)esl::tape_replay, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_bytes);
}
