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

// The lane's arrays in HBM and its LDS area.
struct BigLane {
    double *tv, *rhs, *x, *xold, *sdiag, *q, *cq;
    double *lds;
};

__device__ inline BigLane big_lane(const BigArgs &b, uint32_t g, double *lds)
{
    const uint32_t n = b.ta.op.dp.n;
    BigLane L;
    L.tv = b.ws + (size_t)g * b.ws_stride;
    L.rhs = L.tv + b.tv_cap;
    L.x = L.rhs + n;
    L.xold = L.x + n;
    L.sdiag = L.xold + n;
    L.q = L.sdiag + n;
    L.cq = L.q + n;
    L.lds = lds;
    return L;
}

// Assembly with the deck's reduce tasks, one target per thread: Jacobian
// entries into M[amap[slot]] (amap null: M[slot], pattern order). tran: the
// charges too (row charges into q, charge-Jacobian cells into cq), then the
// ground pin and the companion A = G + ag0 C, F += ag0 (q - q1) - ag1 i_prev
// (assemble_tran's arithmetic); otherwise capacitors and charge targets are
// skipped and only the ground pin follows (assemble_pattern's).
__device__ void big_assemble(const DeckProg &dp, const uint32_t *amap, uint32_t n_slots, const double *live,
                             const double *x, double *lim, double *M, const BigLane &L, bool tran, double t,
                             double ag0, double ag1, const double *q1, const double *ip, uint32_t tid)
{
    const uint32_t n = dp.n;
    double *st = L.lds;
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
            const uint32_t tgt = dp.task_target[k];
            const uint32_t idx = tgt & kTargetMask;
            double *d;
            if (tgt & kRhsBit) {
                d = &L.rhs[idx];
            } else if (tgt & (kCBit | kQBit)) {
                if (!tran) continue;
                d = (tgt & kQBit) ? &L.q[idx] : &L.cq[idx];
            } else {
                d = &M[amap ? amap[tgt] : tgt];
            }
            double acc = *d;
            const uint32_t cc1 = dp.task_cell[k + 1];
            for (uint32_t cc = dp.task_cell[k]; cc < cc1; cc++) acc = acc + st[dp.cells[cc]];
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
    for (uint32_t k = tid; k < dp.n_cslot; k += kWave) {
        const uint32_t sl = amap ? amap[dp.cslot_pat[k]] : dp.cslot_pat[k];
        M[sl] = M[sl] + ag0 * L.cq[k];
    }
    for (uint32_t i = tid; i < n; i += kWave) {
        const double d = ag0 * (L.q[i] - q1[i]);
        const double base = L.rhs[i] + d;
        L.rhs[i] = base - ag1 * ip[i];
    }
    wsync();
}

// The replay on the lane's slots in HBM: void pivots set to 1, the blocks
// (gather, rows, scatter), then the peel tests of refactor(). True when the
// lane must re-pivot.
__device__ bool big_refactor(const TapeProg &T, uint32_t n, double *tv, double *M, uint32_t tid)
{
    LuFlags f{false, 0, 0};
    for (uint32_t v = tid; v < T.n_void; v += kWave) f.bad |= tv[T.amap[T.void_slots[v]]] != 0.0;
    for (uint32_t k = tid; k < n; k += kWave)
        if (T.flags[k] & kVoidBit) tv[T.n_u + k] = 1.0;
    wsync();
    for (uint32_t b = 0; b < T.n_blk; b++) {
        const BigBlock B = T.blk[b];
        const uint32_t *sl = T.blk_slots + B.slot_off;
        for (uint32_t j = tid; j < B.m; j += kWave) M[j] = tv[sl[j]];
        wsync();
        run_rows<double>(T.blk_words + (size_t)B.row0 * kWave, B.rows, M, T.growth, f, tid);
        wsync();
        for (uint32_t j = tid; j < B.w; j += kWave) tv[sl[j]] = M[j];
        wsync();
    }
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

// dx = -A^-1 rhs into y (LDS, pivot order; row i's dx is y[qinv[i]]): the
// forward sweep (L with zero skip) over chunks of columns whose L entries
// and rows are staged after y, then the backward sweep over chunks whose U
// entries, rows and pivots are staged (lane_kernels.zig:123-143). Within a
// column the updated rows are distinct; columns run in order (LDS accesses
// of one wavefront complete in issue order).
__device__ void big_solve(const TapeProg &T, uint32_t n, const double *tv, const double *rhs, double *y,
                          uint32_t tid)
{
    double *sv = y + n;
    uint32_t *si = reinterpret_cast<uint32_t *>(sv + T.sol_cap);
    for (uint32_t r = tid; r < n; r += kWave) y[T.pinv[r]] = -rhs[r];
    wsync();
    for (uint32_t c = 0; c < T.n_fch; c++) {
        const uint32_t k0 = T.fch[c], k1 = T.fch[c + 1];
        const uint32_t e0 = T.lp[k0], e1 = T.lp[k1];
        for (uint32_t e = tid; e < e1 - e0; e += kWave) {
            sv[e] = tv[T.lx0 + e0 + e];
            si[e] = T.li[e0 + e];
        }
        wsync();
        for (uint32_t k = k0; k < k1; k++) {
            const double yk = y[k];
            const uint32_t q1 = T.lp[k + 1] - e0;
            for (uint32_t q = T.lp[k] - e0 + tid; q < q1; q += kWave) {
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
        const uint32_t k0 = T.bch[c], k1 = T.bch[c + 1];
        const uint32_t e0 = T.up[k0], e1 = T.up[k1], ne = e1 - e0;
        for (uint32_t e = tid; e < ne; e += kWave) {
            sv[e] = tv[e0 + e];
            si[e] = T.ui[e0 + e];
        }
        for (uint32_t k = k0 + tid; k < k1; k += kWave) sv[ne + (k - k0)] = tv[T.n_u + k];
        wsync();
        for (uint32_t k = k1; k > k0;) {
            k--;
            const double zk = y[k] / sv[ne + (k - k0)];
            if (tid == 0) y[k] = zk;
            const uint32_t p1 = T.up[k + 1] - e0;
            for (uint32_t p = T.up[k] - e0 + tid; p < p1; p += kWave) {
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
__device__ inline double big_limit_at(const TapeProg &T, uint32_t n, const double *tv, const double *x,
                                      const double *rhs, uint32_t tid)
{
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
    return !(1e3 * scale > 1e9) ? 1e9 : 1e3 * scale;
}

// x += dx (dx = y[qinv[i]]), x_old kept; returns the largest scaled update
// (converger.zig:794-819).
__device__ inline double big_update(const OpArgs &a, const TapeProg &T, const BigLane &L, const double *y,
                                    uint32_t tid)
{
    const DeckProg &dp = a.dp;
    uint64_t worst = 0;
    for (uint32_t i = tid; i < dp.n; i += kWave) {
        const double xo = L.x[i];
        const double d = y[T.qinv[i]];
        const double xn = xo + d;
        L.xold[i] = xo;
        L.x[i] = xn;
        const double tl = a.reltol * fmax(fabs(xn), fabs(xo)) + (dp.current_row[i] ? a.abstol : a.vntol);
        const bool finite = fabs(xn) < __builtin_inf() && fabs(d) < __builtin_inf();
        const double qq = fabs(d) / tl;
        const double score = !finite ? __builtin_inf() : (qq == qq ? qq : 0.0);
        const uint64_t sb = dbits(score);
        worst = sb > worst ? sb : worst;
    }
    return bitsd(wave_max_u64(worst));
}

// Limiting after an update (MOS1, diodes, BJTs): true when any limited.
__device__ inline bool big_limit(const DeckProg &dp, const double *live, const double *x, double *lim, uint32_t tid)
{
    const uint32_t nm = dp.n_mos;
    bool limited = false;
    for (uint32_t j = tid; j < nm; j += kWave) {
        const Mos m = mos_load(dp, j, live);
        const uint32_t *t4 = dp.m_t + 4 * j;
        limited |= mos_limit(m, x[t4[0]], x[t4[1]], x[t4[2]], x[t4[3]], lim + j, nm);
    }
    for (uint32_t j = tid; j < dp.n_dio; j += kWave) {
        const Dio d = dio_load(dp, j, live);
        limited |= dio_limit_v(d, x[dp.d_t[2 * j]], x[dp.d_t[2 * j + 1]], lim[4 * nm + j]);
    }
    for (uint32_t j = tid; j < dp.n_bjt; j += kWave) {
        const Bjt b = bjt_load(dp, j, live);
        const uint32_t *t4 = dp.q_t + 4 * j;
        double *ql = lim + lim_bjt(dp) + 2 * j;
        limited |= bjt_limit_v(b, x[t4[0]], x[t4[1]], x[t4[2]], ql[0], ql[1]);
    }
    return wave_any(limited);
}

// One lane's operating point (op_lane's FP64 path: cold, given or resumed
// start; dump: the Jacobian and residual at the start or saved iterate into
// row g of a_out and f_out).
__device__ void big_op_lane(const BigArgs &b, uint32_t g, uint32_t lane, double *lds)
{
    const OpArgs &a = b.ta.op;
    const uint32_t tid = threadIdx.x;
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n;
    const BigLane L = big_lane(b, g, lds);
    const double *live = a.live + (size_t)lane * dp.n_live;
    double *lim = a.lim + (size_t)lane * lim_width(dp);
    lane_init(a, lane, live, L.x, lim, tid);
    if (a.dump) {
        big_assemble(dp, nullptr, dp.nnz, live, L.x, lim, a.a_out + (size_t)g * dp.nnz, L, false, 0.0, 0.0, 0.0,
                     nullptr, nullptr, tid);
        for (uint32_t i = tid; i < n; i += kWave) a.f_out[(size_t)g * n + i] = L.rhs[i];
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
    uint32_t iter = iter0;
    for (; iter - k0 < a.max_iter; iter++) {
        big_assemble(dp, T.amap, T.n_tv, live, L.x, lim, L.tv, L, false, 0.0, 0.0, 0.0, nullptr, nullptr, tid);
        if (iter == k0) limit = big_limit_at(T, n, L.tv, L.x, L.rhs, tid);
        for (uint32_t i = tid; i < n; i += kWave) L.sdiag[i] = L.tv[T.diag_tv[i]];
        wsync();
        if (big_refactor(T, n, L.tv, L.lds, tid)) {
            for (uint32_t i = tid; i < n; i += kWave) a.x[(size_t)lane * n + i] = L.x[i];
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
                big_assemble(dp, nullptr, dp.nnz, live, L.x, lim, a.a_out + (size_t)lane * dp.nnz, L, false, 0.0,
                             0.0, 0.0, nullptr, nullptr, tid);
            }
            status = kRepivot;
            iters = iter;
            break;
        }
        big_solve(T, n, L.tv, L.rhs, L.lds, tid);
        wsync();
        const double scaled = big_update(a, T, L, L.lds, tid);
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
        for (uint32_t i = tid; i < n; i += kWave) a.x[(size_t)lane * n + i] = L.x[i];
    if (tid == 0) {
        a.status[lane] = status;
        a.iterations[lane] = iters;
    }
}

// Transient steps [s0, s1) of one lane (tran_lane's FP64 path: companion,
// Newton with the transient gates, acceptance, probes and predictor; a lane
// whose replay fails leaves its Jacobian in its row of a_out and resumes at
// the same iterate; dump: the first step's first Jacobian into that row).
__device__ void big_tran_lane(const BigArgs &b, uint32_t g, uint32_t lane, double *lds)
{
    const TranArgs &ta = b.ta;
    const OpArgs &a = ta.op;
    const uint32_t tid = threadIdx.x;
    if (a.status[lane] != kTranRunning) return;
    uint32_t s = ta.step[lane];
    if (s >= ta.s1) return;
    if (s < ta.s0) s = ta.s0;
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n;
    const BigLane L = big_lane(b, g, lds);
    const double *live = a.live + (size_t)lane * dp.n_live;
    double *lim = a.lim + (size_t)lane * lim_width(dp);
    double *q1 = ta.q1 + (size_t)lane * n;
    double *ip = ta.ip + (size_t)lane * n;
    double *xp = ta.xp + (size_t)lane * n;
    const LaneResume rs = a.resume[lane];
    const TapeProg T = a.tapes[rs.tape];
    double *aout = a.a_out + (size_t)lane * dp.nnz;
    for (uint32_t i = tid; i < n; i += kWave) L.x[i] = a.x[(size_t)lane * n + i];
    wsync();
    if (ta.init) {
        // start at the operating point: limiting state at its own voltages,
        // charges from one evaluation there, no dynamic current
        seed_lim_at(dp, live, L.x, lim, tid);
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
    for (; s < ta.s1 && status == kTranRunning; s++) {
        const double t = (double)(s + 1) * ta.dt;
        const double ag0 = s == 0 ? 1.0 / ta.dt : 2.0 / ta.dt;
        const double ag1 = s == 0 ? 0.0 : 1.0;
        bool converged = false;
        uint32_t iter = iter0;
        for (; iter < a.max_iter; iter++) {
            big_assemble(dp, T.amap, T.n_tv, live, L.x, lim, L.tv, L, true, t, ag0, ag1, q1, ip, tid);
            if (iter == 0) limit = big_limit_at(T, n, L.tv, L.x, L.rhs, tid);
            wsync();
            if (big_refactor(T, n, L.tv, L.lds, tid)) {
                // the Jacobian again, read out in pattern order for the
                // host's re-pivot; resume at this iterate
                wsync();
                big_assemble(dp, T.amap, T.n_tv, live, L.x, lim, L.tv, L, true, t, ag0, ag1, q1, ip, tid);
                for (uint32_t p = tid; p < dp.nnz; p += kWave) aout[p] = L.tv[T.amap[p]];
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
            big_solve(T, n, L.tv, L.rhs, L.lds, tid);
            wsync();
            const double scaled = big_update(a, T, L, L.lds, tid);
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
        seed_lim_at(dp, live, L.x, lim, tid);
        wsync();
        // the next step's first iterate: the last two accepted points
        // extrapolated (tran.zig:1131-1137; lane_device.zig:141-153)
        if (s + 1 < ta.steps) {
            for (uint32_t i = tid; i < n; i += kWave) {
                const double cu = L.x[i];
                L.x[i] = cu + (cu - xp[i]);
                xp[i] = cu;
            }
            wsync();
        }
    }
    if (status == kTranRunning && s >= ta.steps) status = kConverged;
    for (uint32_t i = tid; i < n; i += kWave) a.x[(size_t)lane * n + i] = L.x[i];
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

__global__ __launch_bounds__(kWave) void big_lanes(BigArgs b)
{
    extern __shared__ double sm[];
    const uint32_t g = blockIdx.x;
    const uint32_t lane = b.ta.op.lane_ids[g];
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
    hipLaunchKernelGGL(esl::big_lanes, dim3(b->N), dim3(esl::kWave), b->lds * (uint32_t)sizeof(double), s, *b);
    return hipGetLastError();
}
