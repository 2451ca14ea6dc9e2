// Lanes whose working set is beyond a workgroup's LDS (the 64 x 64 RC mesh:
// 4,099 unknowns and about 137,000 factor slots, 1.1 MB per lane): one
// wavefront per lane as in the LDS kernels, with the lane's factor slots and
// vectors in HBM (one contiguous region per lane) and LDS as the staging
// area of each phase:
// - assembly: the stamp cells of one chunk of 64 instances; each reduce task
//   adds its cells into its target in HBM;
// - refactor: the tape's replay cut into blocks of at most a few thousand
//   slots; a block gathers its slots into LDS, runs its rows (64 independent
//   updates per wavefront step, each slot's updates in the order of the
//   column replay, lane_kernels.zig:49-118) and scatters what it wrote;
// - solves: the solve vector in LDS, the L (forward) and U (backward)
//   entries staged a chunk of columns at a time (lane_kernels.zig:123-143).
// The reference's lanes keep these arrays in global memory too, one thread
// per lane walking the tape serially (lane_kernels.zig:7-8). Every value is
// bitwise the wavefront kernels' and the oracle's: the same sums in the same
// order, compiled without contraction.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "esl_device.h"
#include "esl_eval.h"
#include "esl_internal.h"
#include "eslmath.h"

namespace esl {

// Pointers with their address space explicit (gptr/lptr): global for the
// lane's HBM arrays and the tape, LDS for the staging area.
typedef __attribute__((address_space(1))) double gdbl;
typedef __attribute__((address_space(1))) const double gcdbl;
typedef __attribute__((address_space(1))) const uint32_t gcu32;
typedef __attribute__((address_space(3))) double ldbl;
typedef __attribute__((address_space(3))) uint32_t lu32;

// The lane's arrays in HBM and its LDS area (the kernel's dynamic LDS,
// passed down so that no function looks its address up).
struct BigLane {
    gdbl *tv, *rhs, *x, *xold, *sdiag, *q, *cq;
    ldbl *sm;
};

__device__ inline BigLane big_lane(const BigArgs &b, uint32_t g, ldbl *sm)
{
    const uint32_t n = b.ta.op.dp.n;
    BigLane L;
    L.tv = gptr(b.ws) + (size_t)g * b.ws_stride;
    L.rhs = L.tv + b.tv_cap;
    L.x = L.rhs + n;
    L.xold = L.x + n;
    L.sdiag = L.xold + n;
    L.q = L.sdiag + n;
    L.cq = L.q + n;
    L.sm = sm;
    return L;
}

// Assembly with the deck's reduce tasks, one target per thread: Jacobian
// entries into M[amap[slot]] (amap null: M[slot], pattern order). tran: the
// charges too (row charges into q, charge-Jacobian cells into cq), then the
// ground pin and the companion A = G + ag0 C, F += ag0 (q - q1) - ag1 i_prev
// (assemble_tran's arithmetic); otherwise capacitors and charge targets are
// skipped and only the ground pin follows (assemble_pattern's). The stamp
// cells of each chunk are staged in LDS.
__device__ void big_assemble(const DeckProg &dp, const uint32_t *amap_, uint32_t n_slots, const double *live,
                             gdbl *x, double *lim, gdbl *M, const BigLane &L, bool tran, double t, double ag0,
                             double ag1, gcdbl *q1, gcdbl *ip, uint32_t tid)
{
    const uint32_t n = dp.n;
    ldbl *st = L.sm;
    gcu32 *amap = amap_ ? gptr(amap_) : nullptr;
    gcu32 *task_target = gptr(dp.task_target), *task_cell = gptr(dp.task_cell);
    const __attribute__((address_space(1))) uint16_t *cells = gptr(dp.cells);
    for (uint32_t i = tid; i < n_slots; i += kWave) M[i] = 0.0;
    for (uint32_t i = tid; i < n; i += kWave) {
        L.rhs[i] = 0.0;
        if (tran) L.q[i] = 0.0;
    }
    if (tran)
        for (uint32_t k = tid; k < dp.n_cslot; k += kWave) L.cq[k] = 0.0;
    wsync();
    for (uint32_t c = 0; c < dp.n_chunks; c++) {
        const Chunk ch = dp.chunks[c];
        if (!tran && ch.kind == kC) continue;   // no static stamps
        eval_chunk(dp, ch, live, x, lim, dp.n_mos, st, tid, t, tran);
        wsync();
        const uint32_t t1 = dp.chunks[c + 1].task_begin;
        for (uint32_t k = ch.task_begin + tid; k < t1; k += kWave) {
            const uint32_t tgt = task_target[k];
            const uint32_t idx = tgt & kTargetMask;
            gdbl *d;
            if (tgt & kRhsBit) {
                d = &L.rhs[idx];
            } else if (tgt & (kCBit | kQBit)) {
                if (!tran) continue;
                d = (tgt & kQBit) ? &L.q[idx] : &L.cq[idx];
            } else {
                d = &M[amap ? amap[tgt] : tgt];
            }
            double acc = *d;
            const uint32_t cc1 = task_cell[k + 1];
            for (uint32_t cc = task_cell[k]; cc < cc1; cc++) acc = acc + st[cells[cc]];
            *d = acc;
        }
        wsync();
    }
    if (tid == 0) {
        const uint32_t d0 = amap ? amap[dp.diag0_slot] : dp.diag0_slot;
        M[d0] = M[d0] + 1.0;
        L.rhs[0] = L.rhs[0] + x[0];
    }
    wsync();
    if (!tran) return;
    gcu32 *cslot_pat = gptr(dp.cslot_pat);
    for (uint32_t k = tid; k < dp.n_cslot; k += kWave) {
        const uint32_t sl = amap ? amap[cslot_pat[k]] : cslot_pat[k];
        M[sl] = M[sl] + ag0 * L.cq[k];
    }
    for (uint32_t i = tid; i < n; i += kWave) {
        const double d = ag0 * (L.q[i] - q1[i]);
        const double base = L.rhs[i] + d;
        L.rhs[i] = base - ag1 * ip[i];
    }
    wsync();
}

// Batched moves between HBM and LDS: each thread issues kBatch independent
// loads before it uses any, so a wavefront has up to 16 x 64 loads in
// flight instead of waiting on each.
constexpr uint32_t kBatch = 16;

// dst[j] = src[idx[j]] for j < count (LDS from HBM through an index list).
// Pipelined: the next batch's indices load while this batch's values do, so
// a batch costs one memory latency, not two.
__device__ inline void gather(ldbl *dst, gcdbl *src, gcu32 *idx, uint32_t count, uint32_t tid)
{
    uint32_t id[kBatch];
#pragma unroll
    for (uint32_t u = 0; u < kBatch; u++) {
        const uint32_t j = u * kWave + tid;
        id[u] = idx[j < count ? j : 0];
    }
    for (uint32_t j0 = 0; j0 < count; j0 += kBatch * kWave) {
        double v[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; u++) v[u] = src[id[u]];
        const uint32_t j1 = j0 + kBatch * kWave;
#pragma unroll
        for (uint32_t u = 0; u < kBatch; u++) {
            const uint32_t j = j1 + u * kWave + tid;
            id[u] = idx[j < count ? j : 0];
        }
#pragma unroll
        for (uint32_t u = 0; u < kBatch; u++) {
            const uint32_t j = j0 + u * kWave + tid;
            if (j < count) dst[j] = v[u];
        }
    }
}

// dst[idx[j]] = src[j] for j < count (HBM from LDS through an index list).
__device__ inline void scatter(gdbl *dst, gcu32 *idx, const ldbl *src, uint32_t count, uint32_t tid)
{
    for (uint32_t j0 = 0; j0 < count; j0 += kBatch * kWave) {
        uint32_t id[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; u++) {
            const uint32_t j = j0 + u * kWave + tid;
            id[u] = idx[j < count ? j : 0];
        }
#pragma unroll
        for (uint32_t u = 0; u < kBatch; u++) {
            const uint32_t j = j0 + u * kWave + tid;
            if (j < count) dst[id[u]] = src[j];
        }
    }
}

// dst[j] = src[j] for j < count (LDS from a contiguous HBM range, doubles
// or 32-bit indices), with the same batching.
template <typename D, typename S>
__device__ inline void stage(D *dst, S *src, uint32_t count, uint32_t tid)
{
    for (uint32_t j0 = 0; j0 < count; j0 += kBatch * kWave) {
        decltype(+src[0]) v[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; u++) {
            const uint32_t j = j0 + u * kWave + tid;
            v[u] = src[j < count ? j : 0];
        }
#pragma unroll
        for (uint32_t u = 0; u < kBatch; u++) {
            const uint32_t j = j0 + u * kWave + tid;
            if (j < count) dst[j] = v[u];
        }
    }
}

// The replay on the lane's slots in HBM: void pivots set to 1, the blocks
// (gather into LDS, rows, scatter), then the peel tests of refactor(). True
// when the lane must re-pivot.
__device__ bool big_refactor(const TapeProg &T, uint32_t n, gdbl *tv, ldbl *M, uint32_t tid)
{
    gcu32 *flags = gptr(T.flags), *amap = gptr(T.amap), *void_slots = gptr(T.void_slots);
    LuFlags f{false, 0, 0};
    for (uint32_t v = tid; v < T.n_void; v += kWave) f.bad |= tv[amap[void_slots[v]]] != 0.0;
    for (uint32_t k = tid; k < n; k += kWave)
        if (flags[k] & kVoidBit) tv[T.n_u + k] = 1.0;
    wsync();
    const __attribute__((address_space(1))) BigBlock *blk = gptr(T.blk);
    for (uint32_t b = 0; b < T.n_blk; b++) {
        const uint32_t m = blk[b].m, w = blk[b].w, rows = blk[b].rows, row0 = blk[b].row0;
        gcu32 *sl = gptr(T.blk_slots) + blk[b].slot_off;
        gather(M, tv, sl, m, tid);
        wsync();
        run_rows<double>(gptr(T.blk_words) + (size_t)row0 * kWave, rows, M, T.growth, f, tid, 0u);
        wsync();
        scatter(tv, sl, M, w, tid);
        wsync();
    }
    for (uint32_t p = tid; p < T.n_u; p += kWave) {
        const double u = tv[p];
        f.nonfinite |= dbits(u - u);
    }
    for (uint32_t k = tid; k < n; k += kWave) {
        if (flags[k] & kVoidBit) continue;
        const double d = tv[T.n_u + k];
        f.bad |= d == 0.0 || !(fabs(d) < __builtin_inf());
    }
    const uint64_t worst = wave_max_u64(f.worst);
    const bool fail = f.bad || f.nonfinite != 0 || !(bitsd(worst) <= T.wide);
    return wave_any(fail);
}

// dx = -A^-1 rhs into y (LDS, pivot order; row i's dx is y[qinv[i]]): the
// forward sweep (L with zero skip) over chunks of columns whose L entries
// and rows are staged after y, then the backward sweep over chunks whose U
// entries, rows and pivots are staged (lane_kernels.zig:123-143). Within a
// column the updated rows are distinct; columns run in order (LDS accesses
// of one wavefront complete in issue order).
__device__ void big_solve(const TapeProg &T, uint32_t n, gcdbl *tv, gcdbl *rhs, ldbl *y, uint32_t tid)
{
    ldbl *sv = y + n;
    lu32 *si = (lu32 *)(sv + T.sol_cap);
    gcu32 *pinv = gptr(T.pinv), *lp = gptr(T.lp), *li = gptr(T.li), *up = gptr(T.up), *ui = gptr(T.ui);
    gcu32 *fch = gptr(T.fch), *bch = gptr(T.bch);
    for (uint32_t r = tid; r < n; r += kWave) y[pinv[r]] = -rhs[r];
    wsync();
    for (uint32_t c = 0; c < T.n_fch; c++) {
        const uint32_t k0 = fch[c], k1 = fch[c + 1];
        const uint32_t e0 = lp[k0], e1 = lp[k1];
        stage(sv, tv + T.lx0 + e0, e1 - e0, tid);
        stage(si, li + e0, e1 - e0, tid);
        wsync();
        for (uint32_t k = k0; k < k1; k++) {
            const double yk = y[k];
            const uint32_t q1 = lp[k + 1] - e0;
            for (uint32_t q = lp[k] - e0 + tid; q < q1; q += kWave) {
                const uint32_t r = si[q];
                const double old = y[r];
                const double m = sv[q] * yk;
                const double s = old - m;
                y[r] = yk != 0.0 ? s : old;
            }
        }
        wsync();
    }
    for (uint32_t c = T.n_bch; c > 0;) {
        c--;
        const uint32_t k0 = bch[c], k1 = bch[c + 1];
        const uint32_t e0 = up[k0], e1 = up[k1], ne = e1 - e0;
        stage(sv, tv + e0, ne, tid);
        stage(si, ui + e0, ne, tid);
        stage(sv + ne, tv + T.n_u + k0, k1 - k0, tid);
        wsync();
        for (uint32_t k = k1; k > k0;) {
            k--;
            const double zk = y[k] / sv[ne + (k - k0)];
            if (tid == 0) y[k] = zk;
            const uint32_t p1 = up[k + 1] - e0;
            for (uint32_t p = up[k] - e0 + tid; p < p1; p += kWave) {
                const uint32_t i = si[p];
                const double old = y[i];
                const double m = sv[p] * zk;
                const double s = old - m;
                y[i] = zk != 0.0 ? s : old;
            }
        }
        wsync();
    }
}

// The plausible limit of a Newton run's first iterate (converger.zig:377-384).
__device__ inline double big_limit_at(const TapeProg &T, uint32_t n, gcdbl *tv, gcdbl *x, gcdbl *rhs, uint32_t tid)
{
    gcu32 *diag_tv = gptr(T.diag_tv);
    double gm = 1.0;
    uint64_t xm = 0, fm = 0;
    for (uint32_t i = tid; i < n; i += kWave) {
        gm = fmax(gm, fabs(tv[diag_tv[i]]));
        const uint64_t xb = dbits(fabs(x[i])), fb = dbits(fabs(rhs[i]));
        xm = xb > xm ? xb : xm;
        fm = fb > fm ? fb : fm;
    }
    gm = wave_fmax(gm);
    xm = wave_max_u64(xm);
    fm = wave_max_u64(fm);
    const double scale = fmin(fmax(bitsd(xm), bitsd(fm) * gm), 1e18);
    return !(1e3 * scale > 1e9) ? 1e9 : 1e3 * scale;
}

// x += dx (dx = y[qinv[i]]), x_old kept; returns the largest scaled update
// (converger.zig:794-819).
__device__ inline double big_update(const OpArgs &a, const TapeProg &T, const BigLane &L, uint32_t tid)
{
    const DeckProg &dp = a.dp;
    const ldbl *y = L.sm;
    gcu32 *qinv = gptr(T.qinv);
    const __attribute__((address_space(1))) uint8_t *current_row = gptr(dp.current_row);
    uint64_t worst = 0;
    for (uint32_t i = tid; i < dp.n; i += kWave) {
        const double xo = L.x[i];
        const double d = y[qinv[i]];
        const double xn = xo + d;
        L.xold[i] = xo;
        L.x[i] = xn;
        const double tl = a.reltol * fmax(fabs(xn), fabs(xo)) + (current_row[i] ? a.abstol : a.vntol);
        const bool finite = fabs(xn) < __builtin_inf() && fabs(d) < __builtin_inf();
        const double qq = fabs(d) / tl;
        const double score = !finite ? __builtin_inf() : (qq == qq ? qq : 0.0);
        const uint64_t sb = dbits(score);
        worst = sb > worst ? sb : worst;
    }
    return bitsd(wave_max_u64(worst));
}

// Limiting after an update (MOS1, diodes, BJTs): true when any limited.
__device__ __forceinline__ bool big_limit(const DeckProg &dp, const double *live, gcdbl *x, double *lim, uint32_t tid)
{
    return wave_any(limit_pass(dp, live, x, lim, tid));
}

// One lane's operating point (op_lane's FP64 path: cold, given or resumed
// start; dump: the Jacobian and residual at the start or saved iterate into
// row g of a_out and f_out).
__device__ __forceinline__ void big_op_lane(const BigArgs &b, uint32_t g, uint32_t lane, ldbl *sm)
{
    const OpArgs &a = b.ta.op;
    const uint32_t tid = threadIdx.x % kWave;
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n;
    const BigLane L = big_lane(b, g, sm);
    const double *live = a.live + (size_t)lane * dp.n_live;
    double *lim = a.lim + (size_t)lane * lim_width(dp);
    gdbl *xg = gptr(a.x) + (size_t)lane * n;
    lane_init(a, lane, live, (double *)L.x, lim, tid);
    if (a.dump) {
        big_assemble(dp, nullptr, dp.nnz, live, L.x, lim, gptr(a.a_out) + (size_t)g * dp.nnz, L, false, 0.0, 0.0,
                     0.0, nullptr, nullptr, tid);
        gdbl *f = gptr(a.f_out) + (size_t)g * n;
        for (uint32_t i = tid; i < n; i += kWave) f[i] = L.rhs[i];
        return;
    }
    const LaneResume rs = a.resume[lane];
    const TapeProg T = a.tapes[rs.tape];
    uint32_t iter0 = 0, k0 = 0;
    bool init_fix = a.mode == 0;
    double limit = 1e9;
    if (a.mode == 2) {
        iter0 = rs.iter;
        init_fix = rs.init_fix != 0;
        limit = rs.limit;
        k0 = rs.k0;
    }
    uint32_t status = kNoConvergence, iters = a.max_iter;
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
    for (; iter - k0 < a.max_iter; iter++) {
        lap(3);
        big_assemble(dp, T.amap, T.n_tv, live, L.x, lim, L.tv, L, false, 0.0, 0.0, 0.0, nullptr, nullptr, tid);
        if (iter == k0) limit = big_limit_at(T, n, L.tv, L.x, L.rhs, tid);
        for (uint32_t i = tid; i < n; i += kWave) L.sdiag[i] = L.tv[T.diag_tv[i]];
        wsync();
        lap(0);
        const bool repivot = big_refactor(T, n, L.tv, L.sm, tid);
        lap(1);
        if (repivot) {
            for (uint32_t i = tid; i < n; i += kWave) xg[i] = L.x[i];
            if (tid == 0) {
                LaneResume o{};
                o.iter = iter;
                o.init_fix = init_fix ? 1u : 0u;
                o.tape = rs.tape;
                o.limit = limit;
                o.k0 = k0;
                a.resume[lane] = o;
            }
            if (a.a_out) {
                // the Jacobian at this iterate (pattern order) for the host's
                // re-pivot: the same sums as the failed assembly
                wsync();
                big_assemble(dp, nullptr, dp.nnz, live, L.x, lim, gptr(a.a_out) + (size_t)lane * dp.nnz, L, false,
                             0.0, 0.0, 0.0, nullptr, nullptr, tid);
            }
            status = kRepivot;
            iters = iter;
            break;
        }
        big_solve(T, n, L.tv, L.rhs, L.sm, tid);
        wsync();
        lap(2);
        const double scaled = big_update(a, T, L, tid);
        wsync();
        const bool limited = big_limit(dp, live, L.x, lim, tid);
        bool ok = !limited && iter != k0 && scaled < 1.0;
        if (ok) {
            bool miss = false;
            for (uint32_t i = tid; i < n; i += kWave) {
                const double xi = L.x[i];
                miss |= !(fabs(xi) <= limit);
                const double s0 = fabs(L.sdiag[i]);
                const double s = s0 == 0.0 ? 1.0 : s0;
                const double rt = fmax(a.residual_tol, 10.0 * s * (a.reltol * fabs(xi) + a.vntol));
                miss |= !(fabs(L.rhs[i]) <= rt);
            }
            ok = !wave_any(miss);
        }
        if (ok && init_fix) {
            init_fix = false;
        } else if (ok) {
            for (uint32_t i = tid; i < n; i += kWave) L.x[i] = L.xold[i];
            status = kConverged;
            iters = iter + 1;
            break;
        }
        wsync();
    }
    if (status == kNoConvergence) iters = iter;
    wsync();
    if (status != kRepivot)
        for (uint32_t i = tid; i < n; i += kWave) xg[i] = L.x[i];
    lap(3);
    if (tid == 0) {
        a.status[lane] = status;
        a.iterations[lane] = iters;
        if (a.prof)
            for (int p = 0; p < kProfPhases; p++) a.prof[(size_t)lane * kProfPhases + p] += pc[p];
    }
}

// Transient steps [s0, s1) of one lane (tran_lane's FP64 path: companion,
// Newton with the transient gates, acceptance, probes and predictor; a lane
// whose replay fails leaves its Jacobian in its row of a_out and resumes at
// the same iterate; dump: the first step's first Jacobian into that row).
__device__ __forceinline__ void big_tran_lane(const BigArgs &b, uint32_t g, uint32_t lane, ldbl *sm)
{
    const TranArgs &ta = b.ta;
    const OpArgs &a = ta.op;
    const uint32_t tid = threadIdx.x % kWave;
    if (a.status[lane] != kTranRunning) return;
    uint32_t s = ta.step[lane];
    if (s >= ta.s1) return;
    if (s < ta.s0) s = ta.s0;
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n;
    const BigLane L = big_lane(b, g, sm);
    const double *live = a.live + (size_t)lane * dp.n_live;
    double *lim = a.lim + (size_t)lane * lim_width(dp);
    gdbl *q1 = gptr(ta.q1) + (size_t)lane * n;
    gdbl *ip = gptr(ta.ip) + (size_t)lane * n;
    gdbl *xp = gptr(ta.xp) + (size_t)lane * n;
    gdbl *xg = gptr(a.x) + (size_t)lane * n;
    const LaneResume rs = a.resume[lane];
    const TapeProg T = a.tapes[rs.tape];
    gdbl *aout = gptr(a.a_out) + (size_t)lane * dp.nnz;
    gcu32 *amap = gptr(T.amap);
    for (uint32_t i = tid; i < n; i += kWave) L.x[i] = xg[i];
    wsync();
    if (ta.init) {
        // start at the operating point: limiting state at its own voltages,
        // charges from one evaluation there, no dynamic current
        seed_lim_at(dp, live, (const double *)L.x, lim, tid);
        for (uint32_t i = tid; i < n; i += kWave) {
            q1[i] = 0.0;
            ip[i] = 0.0;
            xp[i] = L.x[i];
        }
        wsync();
        big_assemble(dp, T.amap, T.n_tv, live, L.x, lim, L.tv, L, true, 0.0, 0.0, 0.0, q1, ip, tid);
        for (uint32_t i = tid; i < n; i += kWave) q1[i] = L.q[i];
        wsync();
        if (ta.dump) {
            big_assemble(dp, nullptr, dp.nnz, live, L.x, lim, aout, L, true, ta.dt, 1.0 / ta.dt, 0.0, q1, ip, tid);
            return;
        }
    }
    wsync();
    uint32_t iter0 = rs.iter;   // nonzero only when resuming inside a step
    double limit = rs.limit;
    uint32_t status = kTranRunning;
    uint32_t total_it = a.iterations[lane];
    unsigned long long pc[kProfPhases] = {};
    unsigned long long tc = clock64();
    auto lap = [&](int ph) {
        if (a.prof) {
            const unsigned long long now = clock64();
            pc[ph] += now - tc;
            tc = now;
        }
    };
    for (; s < ta.s1 && status == kTranRunning; s++) {
        const double t = (double)(s + 1) * ta.dt;
        const double ag0 = s == 0 ? 1.0 / ta.dt : 2.0 / ta.dt;
        const double ag1 = s == 0 ? 0.0 : 1.0;
        bool converged = false;
        uint32_t iter = iter0;
        for (; iter < a.max_iter; iter++) {
            lap(3);
            big_assemble(dp, T.amap, T.n_tv, live, L.x, lim, L.tv, L, true, t, ag0, ag1, q1, ip, tid);
            if (iter == 0) limit = big_limit_at(T, n, L.tv, L.x, L.rhs, tid);
            wsync();
            lap(0);
            const bool repivot = big_refactor(T, n, L.tv, L.sm, tid);
            lap(1);
            if (repivot) {
                // the Jacobian again, read out in pattern order for the
                // host's re-pivot; resume at this iterate
                wsync();
                big_assemble(dp, T.amap, T.n_tv, live, L.x, lim, L.tv, L, true, t, ag0, ag1, q1, ip, tid);
                for (uint32_t p = tid; p < dp.nnz; p += kWave) aout[p] = L.tv[amap[p]];
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
            big_solve(T, n, L.tv, L.rhs, L.sm, tid);
            wsync();
            lap(2);
            const double scaled = big_update(a, T, L, tid);
            wsync();
            const bool limited = big_limit(dp, live, L.x, lim, tid);
            bool ok = !limited && iter != 0 && scaled < 1.0;
            if (ok) {
                bool miss = false;
                for (uint32_t i = tid; i < n; i += kWave) {
                    miss |= !(fabs(L.x[i]) <= limit);
                    miss |= !(fabs(L.rhs[i]) <= __builtin_inf());
                }
                ok = !wave_any(miss);
            }
            wsync();
            if (ok) {
                for (uint32_t i = tid; i < n; i += kWave) L.x[i] = L.xold[i];
                converged = true;
                iter++;
                break;
            }
        }
        total_it += iter - iter0;
        if (status == kRepivot) break;
        iter0 = 0;
        if (!converged) {
            status = kNoConvergence;
            break;
        }
        // accept: q holds the charges of the converged point (the last
        // assembly ran at x_old)
        wsync();
        for (uint32_t i = tid; i < n; i += kWave) {
            const double d = ag0 * (L.q[i] - q1[i]);
            ip[i] = d - ag1 * ip[i];
            q1[i] = L.q[i];
        }
        for (uint32_t k = tid; k < ta.n_probes; k += kWave)
            ta.probes[((size_t)lane * ta.steps + s) * ta.n_probes + k] = L.x[ta.probe_ids[k]];
        wsync();
        seed_lim_at(dp, live, (const double *)L.x, lim, tid, true);
        wsync();
        // the next step's first iterate: the last two accepted points
        // extrapolated (tran.zig:1131-1137; lane_device.zig:141-153), then
        // one limit pass against the accepted point (tran.zig:1138)
        if (s + 1 < ta.steps) {
            for (uint32_t i = tid; i < n; i += kWave) {
                const double cu = L.x[i];
                L.x[i] = cu + (cu - xp[i]);
                xp[i] = cu;
            }
            wsync();
            limit_pass(dp, live, (gcdbl *)L.x, lim, tid);
            wsync();
        }
    }
    lap(3);
    if (status == kTranRunning && s >= ta.steps) status = kConverged;
    for (uint32_t i = tid; i < n; i += kWave) xg[i] = L.x[i];
    if (tid == 0) {
        if (status != kRepivot) {
            LaneResume o = rs;
            o.iter = 0;
            a.resume[lane] = o;
        }
        ta.step[lane] = s;
        a.status[lane] = status;
        a.iterations[lane] = total_it;
        if (a.prof)
            for (int p = 0; p < kProfPhases; p++) a.prof[(size_t)lane * kProfPhases + p] += pc[p];
    }
}

// Workgroups of b.waves wavefronts, each wavefront one lane in its own
// b.lds doubles of the LDS. The lanes of a workgroup start together and do
// the same work per Newton iteration (the same tape), so they walk the
// replay's row words nearly in step and share their fetches in the caches;
// they never wait on one another.
__global__ __launch_bounds__(kWave * kBigWaves) void big_lanes(BigArgs b)
{
    const uint32_t w = threadIdx.x / kWave;
    const uint32_t g = blockIdx.x * b.waves + w;
    if (g >= b.N) return;
    const uint32_t lane = b.ta.op.lane_ids[g];
    extern __shared__ double dyn[];
    ldbl *sm = lptr(dyn) + (size_t)w * b.lds;
    if (b.tran) big_tran_lane(b, g, lane, sm);
    else big_op_lane(b, g, lane, sm);
}

}  // namespace esl

// Sets the kernel's dynamic LDS limit (before launches, outside graph capture).
extern "C" hipError_t esl_big_attrs(uint32_t lds_bytes)
{
    return hipFuncSetAttribute((const void *)esl::big_lanes, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(lds_bytes > 65536 ? lds_bytes : 65536));
}

extern "C" hipError_t esl_launch_big(const esl::BigArgs *b, hipStream_t s)
{
    const uint32_t grid = (b->N + b->waves - 1) / b->waves;
    hipLaunchKernelGGL(esl::big_lanes, dim3(grid), dim3(esl::kWave * b->waves),
                       b->waves * b->lds * (uint32_t)sizeof(double), s, *b);
    return hipGetLastError();
}
