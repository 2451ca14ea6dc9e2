// Lane kernels for gfx950: one wavefront (64 threads) per sweep point.
//
// The reference runs one GPU thread per lane and walks the whole tape
// serially in every thread (src/solver/lane_device.zig:26-38, bodies
// lane_kernels.zig:49-143; eval per lane src/device/eval.zig:2998-3042),
// with the lane's factors streaming through global memory and the host
// syncing twice per Newton tick (src/analysis/gpu_lanes.zig:782-832). Here a
// lane is a wavefront: its factor slots, residual and vectors sit in LDS;
// devices evaluate 64 at a time (one per thread) and reduce in instance
// order; the refactor runs column levels with one thread per column and the
// solves run row levels with one thread per row, each thread repeating the
// serial operation order of its column or row; the Newton loop, its gates and
// the limiting stay on the device until the lane converges, fails, or needs a
// new pivot sequence.
//
// Compiled with -ffp-contract=off: no fused multiply-add anywhere, so every
// value is bitwise the serial CPU restatement's.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "esl_device.h"
#include "esl_internal.h"
#include "eslmath.h"

namespace esl {

__device__ inline uint64_t dbits(double d) { return __builtin_bit_cast(uint64_t, d); }
__device__ inline double bitsd(uint64_t b) { return __builtin_bit_cast(double, b); }

__device__ inline uint64_t wave_max_u64(uint64_t v)
{
    for (int o = 32; o > 0; o >>= 1) {
        uint64_t w = __shfl_xor(v, o, kWave);
        v = w > v ? w : v;
    }
    return v;
}

__device__ inline double wave_fmax(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ inline bool wave_any(bool p) { return __any((int)p) != 0; }

// Wrappers over the shared physics for the wavefront kernel: limiting state
// per lane at lim[j + k * stride] (vgs, vds, vbs, von), cells staged at
// st[e * 64 + t].
__device__ inline bool mos_limit(const Mos &m, double xd, double xg, double xs, double xb, double *lim,
                                 uint32_t stride)
{
    double lgs = lim[0], lds = lim[stride], lbs = lim[2 * stride];
    const bool moved = mos_limit_v(m, xd, xg, xs, xb, lgs, lds, lbs, lim[3 * stride]);
    lim[0] = lgs;
    lim[stride] = lds;
    lim[2 * stride] = lbs;
    return moved;
}

__device__ inline void mos_eval(const Mos &m, double xd, double xg, double xs, double xb, double *lim,
                                uint32_t stride, double *st, uint32_t t)
{
    double out[kStageEntries];
    lim[3 * stride] = mos_eval_v(m, xd, xg, xs, xb, lim[0], lim[stride], lim[2 * stride], out);
#pragma unroll
    for (int e = 0; e < kStageEntries; e++) st[e * kWave + t] = out[e];
}

// ---- device evaluation of one chunk (one instance per thread) into the
// staging cells st[e * 64 + t]
__device__ inline void eval_chunk(const DeckProg &dp, const Chunk &ch, const double *live, const double *x,
                                  double *lim, uint32_t lim_stride, double *st, uint32_t tid)
{
    if (tid >= ch.count) return;
    const uint32_t j = ch.first + tid;
    switch (ch.kind) {
    case kR: {
        const uint32_t p = dp.r_t[2 * j], q = dp.r_t[2 * j + 1];
        const double r = dp.r_bind[j] >= 0 ? live[dp.r_bind[j]] : dp.r_val[j];
        const double cur = (x[p] - x[q]) / r;
        const double g = 1.0 / r;
        st[0 * kWave + tid] = g;
        st[1 * kWave + tid] = -g;
        st[2 * kWave + tid] = -g;
        st[3 * kWave + tid] = g;
        st[4 * kWave + tid] = cur;
        st[5 * kWave + tid] = -cur;
        break;
    }
    case kV: {
        const uint32_t p = dp.v_t[3 * j], q = dp.v_t[3 * j + 1], br = dp.v_t[3 * j + 2];
        const double dc = dp.v_bind[j] >= 0 ? live[dp.v_bind[j]] : dp.v_val[j];
        const double ib = x[br];
        st[0 * kWave + tid] = 1.0;
        st[1 * kWave + tid] = -1.0;
        st[2 * kWave + tid] = 1.0;
        st[3 * kWave + tid] = -1.0;
        st[4 * kWave + tid] = ib;
        st[5 * kWave + tid] = -ib;
        st[6 * kWave + tid] = (x[p] - x[q]) - dc;
        break;
    }
    case kI: {
        const double dc = dp.i_bind[j] >= 0 ? live[dp.i_bind[j]] : dp.i_val[j];
        st[0 * kWave + tid] = dc;
        st[1 * kWave + tid] = -dc;
        break;
    }
    case kMos: {
        const Mos m = mos_load(dp, j, live);
        const uint32_t *t4 = dp.m_t + 4 * j;
        mos_eval(m, x[t4[0]], x[t4[1]], x[t4[2]], x[t4[3]], lim + j, lim_stride, st, tid);
        break;
    }
    default:
        break;
    }
}

// Assembly into pattern slots (the Jacobian dump for a host re-pivot): the
// deck's reduce tasks, one target per thread.
__device__ void assemble_pattern(const DeckProg &dp, const double *live, const double *x, double *lim,
                                 uint32_t lim_stride, double *a, double *rhs, double *st, uint32_t tid)
{
    const uint32_t n = dp.n;
    for (uint32_t i = tid; i < dp.nnz; i += kWave) a[i] = 0.0;
    for (uint32_t i = tid; i < n; i += kWave) rhs[i] = 0.0;
    __syncthreads();
    for (uint32_t c = 0; c < dp.n_chunks; c++) {
        const Chunk ch = dp.chunks[c];
        eval_chunk(dp, ch, live, x, lim, lim_stride, st, tid);
        __syncthreads();
        const uint32_t t1 = dp.chunks[c + 1].task_begin;
        for (uint32_t t = ch.task_begin + tid; t < t1; t += kWave) {
            const uint32_t tgt = dp.task_target[t];
            double *d = (tgt & kRhsBit) ? &rhs[tgt & ~kRhsBit] : &a[tgt];
            double acc = *d;
            const uint32_t c1 = dp.task_cell[t + 1];
            for (uint32_t cc = dp.task_cell[t]; cc < c1; cc++) acc = acc + st[dp.cells[cc]];
            *d = acc;
        }
        __syncthreads();
    }
    if (tid == 0) {
        a[dp.diag0_slot] = a[dp.diag0_slot] + 1.0;
        rhs[0] = rhs[0] + x[0];
    }
    __syncthreads();
}

struct RedAcc {
    int cur;
    bool rhs;
    double acc;
};

__device__ inline void red_flush(RedAcc &r, double *tv, double *rhs)
{
    if (r.cur >= 0) {
        if (r.rhs) rhs[r.cur] = r.acc;
        else tv[r.cur] = r.acc;
    }
}

__device__ inline void red_word(uint32_t w, RedAcc &r, double *tv, double *rhs, const double *st)
{
    if (w & kRedTarget) {
        red_flush(r, tv, rhs);
        r.cur = (int)(w & 0xfffffu);
        r.rhs = (w & kRedRhs) != 0;
        r.acc = r.rhs ? rhs[r.cur] : tv[r.cur];
    } else if (w < (uint32_t)(kWave * kStageEntries)) {
        r.acc = r.acc + st[w];
    }
}

// Assembly into the tape's factor slots: every chunk evaluates, then each
// thread runs its reduce words (coalesced rows, 16 rows in flight), adding
// each target's cells in instance order into a register.
__device__ void assemble(const DeckProg &dp, const TapeProg &T, const double *live, const double *x, double *lim,
                         uint32_t lim_stride, double *tv, double *rhs, double *st, uint32_t tid)
{
    const uint32_t n = dp.n;
    for (uint32_t i = tid; i < T.n_tv; i += kWave) tv[i] = 0.0;
    for (uint32_t i = tid; i < n; i += kWave) rhs[i] = 0.0;
    __syncthreads();
    for (uint32_t c = 0; c < dp.n_chunks; c++) {
        const Chunk ch = dp.chunks[c];
        eval_chunk(dp, ch, live, x, lim, lim_stride, st, tid);
        __syncthreads();
        const uint32_t *rs = T.red + T.red_off[c] + tid;
        const uint32_t len = T.red_len[c];
        RedAcc r{-1, false, 0.0};
        uint32_t q[kRing];
#pragma unroll
        for (int k = 0; k < kRing; k++) q[k] = rs[k * kWave];
        for (uint32_t w = 0; w < len; w += kRing) {
#pragma unroll
            for (int k = 0; k < kRing; k++) {
                red_word(q[k], r, tv, rhs, st);
                q[k] = rs[(w + kRing + k) * kWave];
            }
        }
        red_flush(r, tv, rhs);
        __syncthreads();
    }
    if (tid == 0) {
        const uint32_t s0 = T.diag_tv[0];
        tv[s0] = tv[s0] + 1.0;
        rhs[0] = rhs[0] + x[0];
    }
    __syncthreads();
}

// ---- chain streams (refactor and solve) ----
// Levels run back to back, 16 rows in flight; each group of four
// words first loads all its operands (a level reads only values final in
// earlier levels, and each slot belongs to one chain, so the loads may run
// ahead of the group's stores), then applies the words in order.
struct LuFlags {
    bool bad;
    uint64_t nonfinite, worst;
};

struct ChainAcc {
    double acc;
    uint32_t tgt;
};

// One word with its operands A = M[b] and B = M[c] already loaded (M holds
// the factor slots followed by the solve vector).
__device__ inline void chain_word(esl::uint2_t w, double A, double B, ChainAcc &s, double *M, double growth,
                                  LuFlags &f)
{
    const uint32_t kind = w.y >> 16;
    if (kind == W_FLOP || kind == W_FWD) {
        const double t = s.acc - A * B;
        s.acc = (kind == W_FLOP || B != 0.0) ? t : s.acc;
    } else if (kind == W_TSTART || kind == W_YSTART) {
        s.acc = A;
        s.tgt = w.x & 0xffffu;
    } else if (kind == W_TSTORE || kind == W_YSTORE) {
        M[s.tgt] = s.acc;
    } else if (kind == W_DIVU || kind == W_DIVS || kind == W_BDIV) {
        const double v = s.acc;
        const double q = v / A;
        M[s.tgt] = q;
        if (kind == W_DIVU) {
            if (growth > 0) f.bad |= !(fabs(A) >= growth * fabs(v));
            else f.nonfinite |= dbits(q - q);
            const uint64_t lb = dbits(fabs(q));
            f.worst = lb > f.worst ? lb : f.worst;
        } else if (kind == W_DIVS) {
            f.nonfinite |= dbits(q - q);
        }
    } else if (kind == W_VOID) {
        M[w.x & 0xffffu] = 1.0;
    }
}

#define ESL_GROUP(R0, R1, R2, R3, ROW)                                                     \
    do {                                                                                   \
        const double a0 = M[R0.x >> 16], b0 = M[R0.y & 0xffffu];                           \
        const double a1 = M[R1.x >> 16], b1 = M[R1.y & 0xffffu];                           \
        const double a2 = M[R2.x >> 16], b2 = M[R2.y & 0xffffu];                           \
        const double a3 = M[R3.x >> 16], b3 = M[R3.y & 0xffffu];                           \
        chain_word(R0, a0, b0, s, M, growth, f);                                           \
        chain_word(R1, a1, b1, s, M, growth, f);                                           \
        chain_word(R2, a2, b2, s, M, growth, f);                                           \
        chain_word(R3, a3, b3, s, M, growth, f);                                           \
        R0 = w[((ROW) + 16) * kWave];                                                      \
        R1 = w[((ROW) + 17) * kWave];                                                      \
        R2 = w[((ROW) + 18) * kWave];                                                      \
        R3 = w[((ROW) + 19) * kWave];                                                      \
        if (bar[(ROW) / kGroup]) __syncthreads();                                          \
    } while (0)

// Runs a chain stream: 16 rows in flight, a barrier after each level.
__device__ void run_chains(const esl::uint2_t *words, const uint32_t *bar, uint32_t rows, double *M, double growth,
                           LuFlags &f, uint32_t tid)
{
    const esl::uint2_t *w = words + tid;
    ChainAcc s{0.0, 0};
    esl::uint2_t r0 = w[0 * kWave], r1 = w[1 * kWave], r2 = w[2 * kWave], r3 = w[3 * kWave];
    esl::uint2_t r4 = w[4 * kWave], r5 = w[5 * kWave], r6 = w[6 * kWave], r7 = w[7 * kWave];
    esl::uint2_t r8 = w[8 * kWave], r9 = w[9 * kWave], r10 = w[10 * kWave], r11 = w[11 * kWave];
    esl::uint2_t r12 = w[12 * kWave], r13 = w[13 * kWave], r14 = w[14 * kWave], r15 = w[15 * kWave];
    for (uint32_t row = 0; row < rows; row += kRing) {
        ESL_GROUP(r0, r1, r2, r3, row);
        ESL_GROUP(r4, r5, r6, r7, row + 4);
        ESL_GROUP(r8, r9, r10, r11, row + 8);
        ESL_GROUP(r12, r13, r14, r15, row + 12);
    }
}

// ---- refactor: true when the lane must re-pivot (lane_kernels.zig:49-118).
// The growth test |d| >= growth * max(|d|, |v|) runs per L entry, the same
// predicate because multiplication by growth is monotone.
__device__ bool refactor(const TapeProg &T, uint32_t n, double *tv, uint32_t tid)
{
    LuFlags f{false, 0, 0};
    for (uint32_t v = tid; v < T.n_void; v += kWave) f.bad |= tv[T.amap[T.void_slots[v]]] != 0.0;
    run_chains(T.rw, T.rw_bar, T.rw_rows, tv, T.growth, f, tid);
    for (uint32_t p = tid; p < T.n_u; p += kWave) {
        const double u = tv[p];
        f.nonfinite |= dbits(u - u);
    }
    for (uint32_t k = tid; k < n; k += kWave) {
        if (T.flags[k] & kVoidBit) continue;
        const double d = tv[T.n_u + k];
        f.bad |= d == 0.0 || !(fabs(d) < __builtin_inf());
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
    __syncthreads();
    LuFlags f{false, 0, 0};
    run_chains(T.sw, T.sw_bar, T.sw_rows, tv, 0.0, f, tid);
}

__device__ inline void lane_init(const OpArgs &a, uint32_t lane, const double *live, double *x, double *lim,
                                 uint32_t tid)
{
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n, nm = dp.n_mos;
    const double *xg = a.x + (size_t)lane * n;
    if (a.mode == 0) {
        for (uint32_t i = tid; i < n; i += kWave) x[i] = 0.0;
    } else {
        for (uint32_t i = tid; i < n; i += kWave) x[i] = xg[i];
    }
    __syncthreads();
    if (a.mode == 2) return;
    for (uint32_t j = tid; j < nm; j += kWave) {
        const Mos m = mos_load(dp, j, live);
        if (a.mode == 0) {
            lim[j] = m.tp * m.vbit + m.gammax * sqrt(m.phit);
            lim[nm + j] = 0.0;
            lim[2 * nm + j] = -1.0;
        } else {
            const uint32_t *t4 = dp.m_t + 4 * j;
            const double xd = x[t4[0]], xgg = x[t4[1]], xs = x[t4[2]], xb = x[t4[3]];
            lim[j] = m.tp * (xgg - xs);
            lim[nm + j] = m.tp * (xd - xs);
            lim[2 * nm + j] = m.tp * (xb - xs);
        }
        lim[3 * nm + j] = 0.0;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kWave) void op_newton(OpArgs a)
{
    extern __shared__ double sm[];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = a.lane_ids[blockIdx.x];
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n, nm = dp.n_mos;
    double *tv = sm;   // the tape's factor slots, then y (solve vector) at tv + n_tv
    double *rhs = tv + a.tv_cap + n;
    double *x = rhs + n;
    double *xold = x + n;
    double *sdiag = xold + n;
    double *st = sdiag + n;
    const double *live = a.live + (size_t)lane * dp.n_live;
    double *lim = a.lim + (size_t)lane * 4 * nm;
    lane_init(a, lane, live, x, lim, tid);

    if (a.dump) {
        assemble_pattern(dp, live, x, lim, nm, tv, rhs, st, tid);
        for (uint32_t p = tid; p < dp.nnz; p += kWave) a.a_out[(size_t)blockIdx.x * dp.nnz + p] = tv[p];
        for (uint32_t i = tid; i < n; i += kWave) a.f_out[(size_t)blockIdx.x * n + i] = rhs[i];
        return;
    }

    LaneResume rs = a.resume[lane];
    const TapeProg T = a.tapes[rs.tape];
    double *y = tv + T.n_tv;
    uint32_t iter0 = 0;
    bool init_fix = a.mode == 0;
    double limit = 1e9;
    if (a.mode == 2) {
        iter0 = rs.iter;
        init_fix = rs.init_fix != 0;
        limit = rs.limit;
    }
    uint32_t status = kNoConvergence, iters = a.max_iter;
    unsigned long long pc[kProfPhases] = {0, 0, 0, 0};
    unsigned long long tc = clock64();
    auto lap = [&](int ph) {
        if (a.prof) {
            const unsigned long long now = clock64();
            pc[ph] += now - tc;
            tc = now;
        }
    };
    for (uint32_t iter = iter0; iter < a.max_iter; iter++) {
        lap(3);
        assemble(dp, T, live, x, lim, nm, tv, rhs, st, tid);
        if (iter == 0) {
            double gm = 1.0;
            uint64_t xm = 0, fm = 0;
            for (uint32_t i = tid; i < n; i += kWave) {
                gm = fmax(gm, fabs(tv[T.diag_tv[i]]));
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
        for (uint32_t i = tid; i < n; i += kWave) sdiag[i] = tv[T.diag_tv[i]];
        __syncthreads();
        lap(0);
        const bool must_repivot = refactor(T, n, tv, tid);
        lap(1);
        if (must_repivot) {
            for (uint32_t i = tid; i < n; i += kWave) a.x[(size_t)lane * n + i] = x[i];
            if (tid == 0) {
                LaneResume o;
                o.iter = iter;
                o.init_fix = init_fix ? 1u : 0u;
                o.tape = rs.tape;
                o.pad = 0;
                o.limit = limit;
                a.resume[lane] = o;
            }
            status = kRepivot;
            iters = iter;
            break;
        }
        solve_neg(T, n, tv, rhs, y, tid);
        lap(2);
        uint64_t worst = 0;
        for (uint32_t i = tid; i < n; i += kWave) {
            const double xo = x[i];
            const double d = y[T.qinv[i]];
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
        __syncthreads();
        bool limited = false;
        for (uint32_t j = tid; j < nm; j += kWave) {
            const Mos m = mos_load(dp, j, live);
            const uint32_t *t4 = dp.m_t + 4 * j;
            limited |= mos_limit(m, x[t4[0]], x[t4[1]], x[t4[2]], x[t4[3]], lim + j, nm);
        }
        limited = wave_any(limited);
        bool ok = !limited && iter != 0 && scaled < 1.0;
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
        __syncthreads();
    }
    lap(3);
    __syncthreads();
    if (status != kRepivot)
        for (uint32_t i = tid; i < n; i += kWave) a.x[(size_t)lane * n + i] = x[i];
    if (tid == 0) {
        a.status[lane] = status;
        a.iterations[lane] = iters;
        if (a.prof)
            for (int p = 0; p < kProfPhases; p++) a.prof[(size_t)lane * kProfPhases + p] += pc[p];
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
    __syncthreads();
    for (uint32_t p = tid; p < r.nnz; p += kWave) tv[T.amap[p]] = r.a[(size_t)lane * r.nnz + p];
    for (uint32_t i = tid; i < n; i += kWave) rhs[i] = r.rhs[(size_t)lane * n + i];
    __syncthreads();
    const bool bad = refactor(T, n, tv, tid);
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

}  // namespace esl

extern "C" hipError_t esl_launch_op_newton(const esl::OpArgs *a, uint32_t grid, uint32_t lds_bytes,
                                           hipStream_t s)
{
    hipLaunchKernelGGL(esl::op_newton, dim3(grid), dim3(esl::kWave), lds_bytes, s, *a);
    return hipGetLastError();
}

extern "C" hipError_t esl_launch_tape_replay(const esl::ReplayArgs *r, uint32_t lanes, uint32_t lds_bytes,
                                             hipStream_t s)
{
    hipLaunchKernelGGL(esl::tape_replay, dim3(lanes), dim3(esl::kWave), lds_bytes, s, *r);
    return hipGetLastError();
}

extern "C" hipError_t esl_kernel_attrs(uint32_t lds_bytes)
{
    hipError_t e = hipFuncSetAttribute((const void *)esl::op_newton,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void *)esl::tape_replay, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_bytes);
}
