// Lanes too large for a workgroup's LDS (the 64 x 64 RC mesh: about 140,000
// factor slots per lane): one thread per lane, every lane-major array in a
// launch workspace in HBM (element i of a lane at ws[i * N + t], so a
// wavefront's 64 lanes touch one 512-byte row per access), the shared tape
// walked in its column order (lane_kernels.zig:49-143, the reference's own
// lane layout, lane_kernels.zig:7-8).
//
// Same arithmetic, in the same order, as the wavefront kernels and the
// oracle: device evaluation through the shared eval_chunk into staging
// cells, each target summed over its cells in instance order, the column
// replay and solves of the tape, the .op gate order and the transient step
// with its companion, predictor and acceptance.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "esl_device.h"
#include "esl_eval.h"
#include "esl_internal.h"
#include "eslmath.h"

namespace esl {

__device__ inline uint64_t bbits(double d) { return __builtin_bit_cast(uint64_t, d); }
__device__ inline double bbitsd(uint64_t b) { return __builtin_bit_cast(double, b); }

// The lane's lane-major arrays in the workspace.
struct BigLane {
    Strided tv, st, rhs, x, xo, y, sd, q, cq;
};

__device__ inline BigLane big_lane(const BigArgs &b, uint32_t t)
{
    const size_t N = b.N, n = b.ta.op.dp.n;
    double *w = b.ws + t;
    BigLane L;
    size_t off = 0;
    auto take = [&](size_t len) {
        Strided s{w + off * N, N};
        off += len;
        return s;
    };
    L.tv = take(b.tv_cap);
    L.st = take(kStageDoubles);
    L.rhs = take(n);
    L.x = take(n);
    L.xo = take(n);
    L.y = take(n);
    L.sd = take(n);
    L.q = take(n);
    L.cq = take(b.ta.op.dp.n_cslot);
    return L;
}

// Limiting state of a start at x (as seed_lim_at), one thread.
__device__ inline void big_seed_lim(const DeckProg &dp, const double *live, Strided x, double *lim)
{
    const uint32_t nm = dp.n_mos;
    for (uint32_t j = 0; j < nm; j++) {
        const Mos m = mos_load(dp, j, live);
        const uint32_t *t4 = dp.m_t + 4 * j;
        const double xd = x[t4[0]], xgg = x[t4[1]], xs = x[t4[2]], xb = x[t4[3]];
        lim[j] = m.tp * (xgg - xs);
        lim[nm + j] = m.tp * (xd - xs);
        lim[2 * nm + j] = m.tp * (xb - xs);
        lim[3 * nm + j] = 0.0;
    }
    for (uint32_t j = 0; j < dp.n_dio; j++) lim[4 * nm + j] = x[dp.d_t[2 * j]] - x[dp.d_t[2 * j + 1]];
    for (uint32_t j = 0; j < dp.n_bjt; j++) {
        const uint32_t *t4 = dp.q_t + 4 * j;
        const double tp = dp.q_par[QF_TYPE * dp.n_bjt + j];
        lim[lim_bjt(dp) + 2 * j] = tp * (x[t4[1]] - x[t4[2]]);
        lim[lim_bjt(dp) + 2 * j + 1] = tp * (x[t4[1]] - x[t4[0]]);
    }
}

// Assembly through the deck's reduce tasks: every chunk evaluated into the
// staging cells, each target the sum of its cells in instance order.
// tran: charges on (charge rows into q, charge-Jacobian cells into cq), then
// the ground pin and the companion; otherwise capacitors and charges are
// skipped. Jacobian entries go to M[amap[slot]] (amap null: M[slot]).
__device__ void big_assemble(const DeckProg &dp, const uint32_t *amap, uint32_t n_slots, const double *live,
                             Strided x, double *lim, Strided M, const BigLane &L, bool tran, double t, double ag0,
                             double ag1, const double *q1, const double *ip)
{
    const uint32_t n = dp.n, nm = dp.n_mos;
    for (uint32_t i = 0; i < n_slots; i++) M[i] = 0.0;
    for (uint32_t i = 0; i < n; i++) L.rhs[i] = 0.0;
    if (tran) {
        for (uint32_t i = 0; i < n; i++) L.q[i] = 0.0;
        for (uint32_t k = 0; k < dp.n_cslot; k++) L.cq[k] = 0.0;
    }
    for (uint32_t c = 0; c < dp.n_chunks; c++) {
        const Chunk ch = dp.chunks[c];
        if (!tran && ch.kind == kC) continue;
        for (uint32_t i = 0; i < ch.count; i++) eval_chunk(dp, ch, live, x, lim, nm, L.st, i, t, tran);
        const uint32_t t1 = dp.chunks[c + 1].task_begin;
        for (uint32_t k = ch.task_begin; k < t1; k++) {
            const uint32_t tgt = dp.task_target[k];
            const uint32_t idx = tgt & kTargetMask;
            double *d;
            if (tgt & kRhsBit) d = &L.rhs[idx];
            else if (tgt & (kCBit | kQBit)) {
                if (!tran) continue;
                d = (tgt & kQBit) ? &L.q[idx] : &L.cq[idx];
            } else {
                d = &M[amap ? amap[tgt] : tgt];
            }
            double acc = *d;
            const uint32_t c1 = dp.task_cell[k + 1];
            for (uint32_t cc = dp.task_cell[k]; cc < c1; cc++) acc = acc + L.st[dp.cells[cc]];
            *d = acc;
        }
    }
    const uint32_t d0 = amap ? amap[dp.diag0_slot] : dp.diag0_slot;
    M[d0] = M[d0] + 1.0;
    L.rhs[0] = L.rhs[0] + x[0];
    if (!tran) return;
    for (uint32_t k = 0; k < dp.n_cslot; k++) {
        const uint32_t sl = amap ? amap[dp.cslot_pat[k]] : dp.cslot_pat[k];
        M[sl] = M[sl] + ag0 * L.cq[k];
    }
    for (uint32_t i = 0; i < n; i++) {
        const double dd = ag0 * (L.q[i] - q1[i]);
        const double base = L.rhs[i] + dd;
        L.rhs[i] = base - ag1 * ip[i];
    }
}

// The column replay of the tape on tv (lane_kernels.zig:49-118): true when
// the lane must re-pivot.
__device__ bool big_refactor(const TapeProg &T, uint32_t n, Strided tv)
{
    bool bad = false;
    uint64_t nonfinite = 0, worst = 0;
    for (uint32_t v = 0; v < T.n_void; v++) bad |= tv[T.amap[T.void_slots[v]]] != 0.0;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t f = T.fptr[k];
        const uint32_t u1 = T.up[k + 1];
        for (uint32_t p = T.up[k]; p < u1; p++) {
            const uint32_t i = T.ui[p];
            const double u = tv[p];
            nonfinite |= bbits(u - u);
            const uint32_t q1 = T.lp[i + 1];
            for (uint32_t q = T.lp[i]; q < q1; q++, f++) {
                const uint32_t d = T.dst[f];
                const double m = tv[T.lx0 + q] * u;
                tv[d] = tv[d] - m;
            }
        }
        const uint32_t fl = T.flags[k];
        if (fl & kVoidBit) {
            tv[T.n_u + k] = 1.0;
            continue;
        }
        double dv = tv[T.n_u + k];
        const bool dead = dv == 0.0 || !(fabs(dv) < __builtin_inf());
        bad |= dead;
        if (dead) dv = 1.0;
        tv[T.n_u + k] = dv;
        const uint32_t l0 = T.lp[k], l1 = T.lp[k + 1];
        if (T.growth > 0 && !(fl & kScaledBit)) {
            uint64_t cb = bbits(fabs(dv));
            for (uint32_t q = l0; q < l1; q++) {
                const double v = tv[T.lx0 + q];
                const uint64_t vb = bbits(fabs(v));
                cb = vb > cb ? vb : cb;
                const double l = v / dv;
                tv[T.lx0 + q] = l;
                const uint64_t lb = bbits(fabs(l));
                worst = lb > worst ? lb : worst;
            }
            bad |= !(fabs(dv) >= T.growth * bbitsd(cb));
        } else {
            const bool wide = !(fl & kScaledBit);
            for (uint32_t q = l0; q < l1; q++) {
                const double l = tv[T.lx0 + q] / dv;
                nonfinite |= bbits(l - l);
                tv[T.lx0 + q] = l;
                if (wide) {
                    const uint64_t lb = bbits(fabs(l));
                    worst = lb > worst ? lb : worst;
                }
            }
        }
    }
    return bad || nonfinite != 0 || !(bbitsd(worst) <= T.wide);
}

// dx = -A^-1 rhs into y (pivot order) (lane_kernels.zig:123-143).
__device__ void big_solve(const TapeProg &T, uint32_t n, Strided tv, Strided rhs, Strided y)
{
    for (uint32_t r = 0; r < n; r++) y[T.pinv[r]] = -rhs[r];
    for (uint32_t k = 0; k < n; k++) {
        const double yk = y[k];
        const uint32_t q1 = T.lp[k + 1];
        for (uint32_t q = T.lp[k]; q < q1; q++) {
            const uint32_t r = T.li[q];
            const double old = y[r];
            const double m = tv[T.lx0 + q] * yk;
            const double s = old - m;
            y[r] = yk != 0.0 ? s : old;
        }
    }
    for (uint32_t k = n; k > 0;) {
        k--;
        const double zk = y[k] / tv[T.n_u + k];
        y[k] = zk;
        const uint32_t p1 = T.up[k + 1];
        for (uint32_t p = T.up[k]; p < p1; p++) {
            const uint32_t i = T.ui[p];
            const double old = y[i];
            const double m = tv[p] * zk;
            const double s = old - m;
            y[i] = zk != 0.0 ? s : old;
        }
    }
}

// Limiting after an update (MOS1, diodes, BJTs): true when any limited.
__device__ inline bool big_limit(const DeckProg &dp, const double *live, Strided x, double *lim)
{
    const uint32_t nm = dp.n_mos;
    bool limited = false;
    for (uint32_t j = 0; j < nm; j++) {
        const Mos m = mos_load(dp, j, live);
        const uint32_t *t4 = dp.m_t + 4 * j;
        limited |= mos_limit(m, x[t4[0]], x[t4[1]], x[t4[2]], x[t4[3]], lim + j, nm);
    }
    for (uint32_t j = 0; j < dp.n_dio; j++) {
        const Dio d = dio_load(dp, j, live);
        limited |= dio_limit_v(d, x[dp.d_t[2 * j]], x[dp.d_t[2 * j + 1]], lim[4 * nm + j]);
    }
    for (uint32_t j = 0; j < dp.n_bjt; j++) {
        const Bjt b = bjt_load(dp, j, live);
        const uint32_t *t4 = dp.q_t + 4 * j;
        double *ql = lim + lim_bjt(dp) + 2 * j;
        limited |= bjt_limit_v(b, x[t4[0]], x[t4[1]], x[t4[2]], ql[0], ql[1]);
    }
    return limited;
}

// The plausible limit of a Newton run's first iterate (converger.zig:377-384).
__device__ inline double big_limit_at(uint32_t n, const TapeProg *T, const uint32_t *diag_slot, Strided M, Strided x,
                                      Strided rhs)
{
    double gm = 1.0;
    uint64_t xm = 0, fm = 0;
    for (uint32_t i = 0; i < n; i++) {
        gm = fmax(gm, fabs(M[T ? T->diag_tv[i] : diag_slot[i]]));
        const uint64_t xb = bbits(fabs(x[i])), fb = bbits(fabs(rhs[i]));
        xm = xb > xm ? xb : xm;
        fm = fb > fm ? fb : fm;
    }
    const double scale = fmin(fmax(bbitsd(xm), bbitsd(fm) * gm), 1e18);
    return !(1e3 * scale > 1e9) ? 1e9 : 1e3 * scale;
}

// The operating point of one lane (op_lane for FP64 lanes, modes 0 cold,
// 1 given x, 2 resume; dump: the Jacobian and residual at the start or
// saved iterate into row t of a_out and f_out).
__device__ void big_op(const BigArgs &b, uint32_t t, uint32_t lane)
{
    const OpArgs &a = b.ta.op;
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n, nm = dp.n_mos;
    const BigLane L = big_lane(b, t);
    const double *live = a.live + (size_t)lane * dp.n_live;
    double *lim = a.lim + (size_t)lane * lim_width(dp);
    const double *xg = a.x + (size_t)lane * n;
    for (uint32_t i = 0; i < n; i++) L.x[i] = a.mode == 0 ? 0.0 : xg[i];
    if (a.mode == 0) {
        for (uint32_t j = 0; j < nm; j++) {
            const Mos m = mos_load(dp, j, live);
            lim[j] = m.tp * m.vbit + m.gammax * sqrt(m.phit);
            lim[nm + j] = 0.0;
            lim[2 * nm + j] = -1.0;
            lim[3 * nm + j] = 0.0;
        }
        for (uint32_t j = 0; j < dp.n_dio; j++) lim[4 * nm + j] = dio_load(dp, j, live).vcrit;
        for (uint32_t j = 0; j < dp.n_bjt; j++) {
            lim[lim_bjt(dp) + 2 * j] = bjt_load(dp, j, live).vcrit;
            lim[lim_bjt(dp) + 2 * j + 1] = 0.0;
        }
    } else if (a.mode == 1) {
        big_seed_lim(dp, live, L.x, lim);
    }
    if (a.dump) {
        Strided A{a.a_out + (size_t)t * dp.nnz, 1};
        big_assemble(dp, nullptr, dp.nnz, live, L.x, lim, A, L, false, 0.0, 0.0, 0.0, nullptr, nullptr);
        for (uint32_t i = 0; i < n; i++) a.f_out[(size_t)t * n + i] = L.rhs[i];
        return;
    }
    LaneResume rs = a.resume[lane];
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
        big_assemble(dp, T.amap, T.n_tv, live, L.x, lim, L.tv, L, false, 0.0, 0.0, 0.0, nullptr, nullptr);
        if (iter == k0) limit = big_limit_at(n, &T, nullptr, L.tv, L.x, L.rhs);
        for (uint32_t i = 0; i < n; i++) L.sd[i] = L.tv[T.diag_tv[i]];
        if (big_refactor(T, n, L.tv)) {
            for (uint32_t i = 0; i < n; i++) a.x[(size_t)lane * n + i] = L.x[i];
            LaneResume o{};
            o.iter = iter;
            o.init_fix = init_fix ? 1u : 0u;
            o.tape = rs.tape;
            o.limit = limit;
            o.k0 = k0;
            a.resume[lane] = o;
            if (a.a_out) {
                // the Jacobian at this iterate, pattern order, for the host's
                // re-pivot (the same sums as the failed assembly)
                Strided A{a.a_out + (size_t)lane * dp.nnz, 1};
                big_assemble(dp, nullptr, dp.nnz, live, L.x, lim, A, L, false, 0.0, 0.0, 0.0, nullptr, nullptr);
            }
            status = kRepivot;
            iters = iter;
            break;
        }
        big_solve(T, n, L.tv, L.rhs, L.y);
        uint64_t worst = 0;
        for (uint32_t i = 0; i < n; i++) {
            const double xv = L.x[i];
            const double d = L.y[T.qinv[i]];
            const double xn = xv + d;
            L.xo[i] = xv;
            L.x[i] = xn;
            const double tl = a.reltol * fmax(fabs(xn), fabs(xv)) + (dp.current_row[i] ? a.abstol : a.vntol);
            const bool finite = fabs(xn) < __builtin_inf() && fabs(d) < __builtin_inf();
            const double qq = fabs(d) / tl;
            const double score = !finite ? __builtin_inf() : (qq == qq ? qq : 0.0);
            const uint64_t sb = bbits(score);
            worst = sb > worst ? sb : worst;
        }
        const double scaled = bbitsd(worst);
        const bool limited = big_limit(dp, live, L.x, lim);
        bool ok = !limited && iter != k0 && scaled < 1.0;
        if (ok) {
            for (uint32_t i = 0; i < n; i++) {
                const double xi = L.x[i];
                if (!(fabs(xi) <= limit)) ok = false;
                const double s0 = fabs(L.sd[i]);
                const double s = s0 == 0.0 ? 1.0 : s0;
                const double rt = fmax(a.residual_tol, 10.0 * s * (a.reltol * fabs(xi) + a.vntol));
                if (!(fabs(L.rhs[i]) <= rt)) ok = false;
            }
        }
        if (ok && init_fix) {
            init_fix = false;
        } else if (ok) {
            for (uint32_t i = 0; i < n; i++) L.x[i] = L.xo[i];
            status = kConverged;
            iters = iter + 1;
            break;
        }
    }
    if (status == kNoConvergence) iters = iter;
    if (status != kRepivot)
        for (uint32_t i = 0; i < n; i++) a.x[(size_t)lane * n + i] = L.x[i];
    a.status[lane] = status;
    a.iterations[lane] = iters;
}

// Transient steps [s0, s1) of one lane (tran_lane's steps, acceptance,
// predictor and re-pivot exit; dump: the first step's first Jacobian into
// the lane's row of a_out).
__device__ void big_tran(const BigArgs &b, uint32_t t, uint32_t lane)
{
    const TranArgs &ta = b.ta;
    const OpArgs &a = ta.op;
    if (a.status[lane] != kTranRunning) return;
    uint32_t s = ta.step[lane];
    if (s >= ta.s1) return;
    if (s < ta.s0) s = ta.s0;
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n;
    const BigLane L = big_lane(b, t);
    const double *live = a.live + (size_t)lane * dp.n_live;
    double *lim = a.lim + (size_t)lane * lim_width(dp);
    double *q1 = ta.q1 + (size_t)lane * n;
    double *ip = ta.ip + (size_t)lane * n;
    double *xp = ta.xp + (size_t)lane * n;
    LaneResume rs = a.resume[lane];
    const TapeProg T = a.tapes[rs.tape];
    for (uint32_t i = 0; i < n; i++) L.x[i] = a.x[(size_t)lane * n + i];
    if (ta.init) {
        big_seed_lim(dp, live, L.x, lim);
        for (uint32_t i = 0; i < n; i++) {
            q1[i] = 0.0;
            ip[i] = 0.0;
            xp[i] = L.x[i];
        }
        big_assemble(dp, T.amap, T.n_tv, live, L.x, lim, L.tv, L, true, 0.0, 0.0, 0.0, q1, ip);
        for (uint32_t i = 0; i < n; i++) q1[i] = L.q[i];
        if (ta.dump) {
            Strided A{a.a_out + (size_t)lane * dp.nnz, 1};
            big_assemble(dp, nullptr, dp.nnz, live, L.x, lim, A, L, true, ta.dt, 1.0 / ta.dt, 0.0, q1, ip);
            return;
        }
    }
    uint32_t iter0 = rs.iter;
    double limit = rs.limit;
    uint32_t status = kTranRunning;
    uint32_t total_it = a.iterations[lane];
    for (; s < ta.s1 && status == kTranRunning; s++) {
        const double tt = (double)(s + 1) * ta.dt;
        const double ag0 = s == 0 ? 1.0 / ta.dt : 2.0 / ta.dt;
        const double ag1 = s == 0 ? 0.0 : 1.0;
        bool converged = false;
        uint32_t iter = iter0;
        for (; iter < a.max_iter; iter++) {
            big_assemble(dp, T.amap, T.n_tv, live, L.x, lim, L.tv, L, true, tt, ag0, ag1, q1, ip);
            if (iter == 0) limit = big_limit_at(n, &T, nullptr, L.tv, L.x, L.rhs);
            if (big_refactor(T, n, L.tv)) {
                // the Jacobian again, in the tape's slots, read out in
                // pattern order for the host's re-pivot
                big_assemble(dp, T.amap, T.n_tv, live, L.x, lim, L.tv, L, true, tt, ag0, ag1, q1, ip);
                for (uint32_t p = 0; p < dp.nnz; p++) a.a_out[(size_t)lane * dp.nnz + p] = L.tv[T.amap[p]];
                LaneResume o = rs;
                o.iter = iter;
                o.limit = limit;
                o.fresh = 0;
                a.resume[lane] = o;
                status = kRepivot;
                break;
            }
            big_solve(T, n, L.tv, L.rhs, L.y);
            uint64_t worst = 0;
            for (uint32_t i = 0; i < n; i++) {
                const double xv = L.x[i];
                const double d = L.y[T.qinv[i]];
                const double xn = xv + d;
                L.xo[i] = xv;
                L.x[i] = xn;
                const double tl = a.reltol * fmax(fabs(xn), fabs(xv)) + (dp.current_row[i] ? a.abstol : a.vntol);
                const bool finite = fabs(xn) < __builtin_inf() && fabs(d) < __builtin_inf();
                const double qq = fabs(d) / tl;
                const double score = !finite ? __builtin_inf() : (qq == qq ? qq : 0.0);
                const uint64_t sb = bbits(score);
                worst = sb > worst ? sb : worst;
            }
            const double scaled = bbitsd(worst);
            const bool limited = big_limit(dp, live, L.x, lim);
            bool ok = !limited && iter != 0 && scaled < 1.0;
            if (ok)
                for (uint32_t i = 0; i < n; i++)
                    if (!(fabs(L.x[i]) <= limit) || !(fabs(L.rhs[i]) <= __builtin_inf())) ok = false;
            if (ok) {
                for (uint32_t i = 0; i < n; i++) L.x[i] = L.xo[i];
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
        for (uint32_t i = 0; i < n; i++) {
            const double d = ag0 * (L.q[i] - q1[i]);
            ip[i] = d - ag1 * ip[i];
            q1[i] = L.q[i];
        }
        for (uint32_t k = 0; k < ta.n_probes; k++)
            ta.probes[((size_t)lane * ta.steps + s) * ta.n_probes + k] = L.x[ta.probe_ids[k]];
        big_seed_lim(dp, live, L.x, lim);
        if (s + 1 < ta.steps)
            for (uint32_t i = 0; i < n; i++) {
                const double cu = L.x[i];
                L.x[i] = cu + (cu - xp[i]);
                xp[i] = cu;
            }
    }
    if (status == kTranRunning && s >= ta.steps) status = kConverged;
    for (uint32_t i = 0; i < n; i++) a.x[(size_t)lane * n + i] = L.x[i];
    if (status != kRepivot) {
        LaneResume o = rs;
        o.iter = 0;
        a.resume[lane] = o;
    }
    ta.step[lane] = s;
    a.status[lane] = status;
    a.iterations[lane] = total_it;
}

__global__ __launch_bounds__(64) void big_lanes(BigArgs b)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.N) return;
    const uint32_t lane = b.ta.op.lane_ids[t];
    if (b.tran) big_tran(b, t, lane);
    else big_op(b, t, lane);
}

}  // namespace esl

extern "C" hipError_t esl_launch_big(const esl::BigArgs *b, hipStream_t s)
{
    hipLaunchKernelGGL(esl::big_lanes, dim3((b->N + 63) / 64), dim3(64), 0, s, *b);
    return hipGetLastError();
}
