// Lane-per-thread Newton kernel for gfx950: one GPU thread per sweep point,
// the tiling of the reference's lane kernels (src/solver/lane_device.zig:14,
// 26-139; per-lane eval src/device/eval.zig:2998-3042), run as one
// persistent Newton loop per lane instead of 7+ launches and 2 host syncs per
// tick (src/analysis/gpu_lanes.zig:782-832).
//
// Every thread of a wavefront walks the same deck and the same pivot tape, so
// the instance data and the tape indices are wavefront-uniform (scalar
// loads); each lane's state is lane-major in HBM (element i of thread t at
// i * N + t), so every per-lane access of a wavefront is one coalesced 512-byte
// transaction. The refactor and solves follow the column order of the
// serial replay (lane_kernels.zig:49-143), stamps follow instance order, and
// the code is compiled without contraction: each lane is bitwise the oracle.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "esl_device.h"
#include "esl_internal.h"
#include "eslmath.h"

namespace esl {

__device__ inline uint64_t lbits(double d) { return __builtin_bit_cast(uint64_t, d); }
__device__ inline double lbitsd(uint64_t b) { return __builtin_bit_cast(double, b); }

__global__ __launch_bounds__(256) void op_newton_lpt(LptArgs a)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.N) return;
    const DeckProg &dp = a.dp;
    const TapeProg &T = a.T;
    const size_t N = a.N;
    const uint32_t n = dp.n, nm = dp.n_mos;
    const uint32_t lane = a.lane_ids[t];
    const double *live = a.live_t + t;
    double *tv = a.tvg + t, *rhs = a.rhsg + t, *x = a.xg + t, *xo = a.xold + t, *y = a.yg + t;
    double *sd = a.sdiag + t, *lim = a.limg + t;
#define V(p, i) (p)[(size_t)(i) * N]

    // start point and limiting state (dc/op.zig:24-27, eval.zig:1841-1853)
    for (uint32_t i = 0; i < n; i++) V(x, i) = a.mode == 0 ? 0.0 : a.x[(size_t)lane * n + i];
    for (uint32_t j = 0; j < nm; j++) {
        const Mos m = mos_load(dp, j, live, a.N);
        const uint32_t *t4 = dp.m_t + 4 * j;
        if (a.mode == 0) {
            V(lim, j) = m.tp * m.vbit + m.gammax * sqrt(m.phit);
            V(lim, nm + j) = 0.0;
            V(lim, 2 * nm + j) = -1.0;
        } else {
            const double xd = V(x, t4[0]), xg = V(x, t4[1]), xs = V(x, t4[2]), xb = V(x, t4[3]);
            V(lim, j) = m.tp * (xg - xs);
            V(lim, nm + j) = m.tp * (xd - xs);
            V(lim, 2 * nm + j) = m.tp * (xb - xs);
        }
        V(lim, 3 * nm + j) = 0.0;
    }
    const uint32_t nd = dp.n_dio, nq = dp.n_bjt, lb = 4 * nm + nd;   // diode, BJT limiting state
    for (uint32_t j = 0; j < nd; j++) {
        const uint32_t pa = dp.d_t[2 * j], pk = dp.d_t[2 * j + 1];
        V(lim, 4 * nm + j) = a.mode == 0 ? dio_load(dp, j, live, a.N).vcrit : V(x, pa) - V(x, pk);
    }
    for (uint32_t j = 0; j < nq; j++) {
        const uint32_t *t4 = dp.q_t + 4 * j;
        if (a.mode == 0) {
            V(lim, lb + 2 * j) = bjt_load(dp, j, live, a.N).vcrit;
            V(lim, lb + 2 * j + 1) = 0.0;
        } else {
            const double tp = dp.q_par[QF_TYPE * nq + j];
            V(lim, lb + 2 * j) = tp * (V(x, t4[1]) - V(x, t4[2]));
            V(lim, lb + 2 * j + 1) = tp * (V(x, t4[1]) - V(x, t4[0]));
        }
    }

    bool init_fix = a.mode == 0;
    double limit = 1e9;
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
    for (uint32_t iter = 0; iter < a.max_iter; iter++) {
        lap(3);
        // ---- assembly: kinds in stamp order, instances in deck order ----
        for (uint32_t s = 0; s < T.n_tv; s++) V(tv, s) = 0.0;
        for (uint32_t i = 0; i < n; i++) V(rhs, i) = 0.0;
        const uint32_t nr = dp.kind_first[kR + 1] - dp.kind_first[kR];
        for (uint32_t j = 0; j < nr; j++) {
            const uint32_t p = dp.r_t[2 * j], q = dp.r_t[2 * j + 1];
            const double r = dp.r_bind[j] >= 0 ? V(live, dp.r_bind[j]) : dp.r_val[j];
            const double cur = (V(x, p) - V(x, q)) / r;
            const double g = 1.0 / r;
            if (p != 0) V(rhs, p) += cur;
            if (q != 0) V(rhs, q) += -cur;
            const double gv[4] = {g, -g, -g, g};
            for (int e = 0; e < 4; e++) {
                const uint32_t sl = dp.r_slot[4 * j + e];
                if (sl != kNoSlot) V(tv, T.amap[sl]) += gv[e];
            }
        }
        const uint32_t nv = dp.kind_first[kV + 1] - dp.kind_first[kV];
        for (uint32_t j = 0; j < nv; j++) {
            const uint32_t p = dp.v_t[3 * j], q = dp.v_t[3 * j + 1], br = dp.v_t[3 * j + 2];
            const double dc = dp.v_bind[j] >= 0 ? V(live, dp.v_bind[j]) : dp.v_val[j];
            const double ib = V(x, br);
            if (p != 0) V(rhs, p) += ib;
            if (q != 0) V(rhs, q) += -ib;
            V(rhs, br) += (V(x, p) - V(x, q)) - dc;
            const double gv[4] = {1.0, -1.0, 1.0, -1.0};
            for (int e = 0; e < 4; e++) {
                const uint32_t sl = dp.v_slot[4 * j + e];
                if (sl != kNoSlot) V(tv, T.amap[sl]) += gv[e];
            }
        }
        const uint32_t ni = dp.kind_first[kI + 1] - dp.kind_first[kI];
        for (uint32_t j = 0; j < ni; j++) {
            const uint32_t p = dp.i_t[2 * j], q = dp.i_t[2 * j + 1];
            const double dc = dp.i_bind[j] >= 0 ? V(live, dp.i_bind[j]) : dp.i_val[j];
            if (p != 0) V(rhs, p) += dc;
            if (q != 0) V(rhs, q) += -dc;
        }
        for (uint32_t j = 0; j < nd; j++) {
            const Dio d = dio_load(dp, j, live, a.N);
            const uint32_t pa = dp.d_t[2 * j], pk = dp.d_t[2 * j + 1];
            double out[12];
            dio_eval_v(d, V(x, pa), V(x, pk), V(lim, 4 * nm + j), false, out);
            for (int e = 0; e < 4; e++) {
                const uint32_t sl = dp.d_slot[4 * j + e];
                if (sl != kNoSlot) V(tv, T.amap[sl]) += out[e];
            }
            if (pa != 0) V(rhs, pa) += out[4];
            if (pk != 0) V(rhs, pk) += out[5];
        }
        for (uint32_t j = 0; j < nm; j++) {
            const Mos m = mos_load(dp, j, live, a.N);
            const uint32_t *t4 = dp.m_t + 4 * j;
            double out[kStageEntries];
            V(lim, 3 * nm + j) = mos_eval_v(m, V(x, t4[0]), V(x, t4[1]), V(x, t4[2]), V(x, t4[3]), V(lim, j),
                                            V(lim, nm + j), V(lim, 2 * nm + j), out);
            if (t4[0] != 0) V(rhs, t4[0]) += out[11];
            if (t4[2] != 0) V(rhs, t4[2]) += out[12];
            if (t4[3] != 0) V(rhs, t4[3]) += out[13];
            for (int e = 0; e < 11; e++) {
                const uint32_t sl = dp.m_slot[11 * j + e];
                if (sl != kNoSlot) V(tv, T.amap[sl]) += out[e];
            }
        }
        for (uint32_t j = 0; j < nq; j++) {
            const Bjt b = bjt_load(dp, j, live, a.N);
            const uint32_t *t4 = dp.q_t + 4 * j;
            double out[16];
            bjt_eval_v(b, V(x, t4[0]), V(x, t4[1]), V(x, t4[2]), V(x, t4[3]), V(lim, lb + 2 * j),
                       V(lim, lb + 2 * j + 1), out);
            for (int e = 0; e < 12; e++) {
                const uint32_t sl = dp.q_slot[12 * j + e];
                if (sl != kNoSlot) V(tv, T.amap[sl]) += out[e];
            }
            for (int r = 0; r < 4; r++)
                if (t4[r] != 0) V(rhs, t4[r]) += out[12 + r];
        }
        V(tv, T.diag_tv[0]) += 1.0;
        V(rhs, 0) += V(x, 0);
        if (iter == 0) {
            double gm = 1.0;
            uint64_t xm = 0, fm = 0;
            for (uint32_t i = 0; i < n; i++) {
                gm = fmax(gm, fabs(V(tv, T.diag_tv[i])));
                const uint64_t xb = lbits(fabs(V(x, i))), fb = lbits(fabs(V(rhs, i)));
                xm = xb > xm ? xb : xm;
                fm = fb > fm ? fb : fm;
            }
            const double scale = fmin(fmax(lbitsd(xm), lbitsd(fm) * gm), 1e18);
            limit = !(1e3 * scale > 1e9) ? 1e9 : 1e3 * scale;
        }
        for (uint32_t i = 0; i < n; i++) V(sd, i) = V(tv, T.diag_tv[i]);
        lap(0);

        // ---- refactor (lane_kernels.zig:49-118 over the tape's slots) ----
        bool bad = false;
        uint64_t nonfinite = 0, worst = 0;
        for (uint32_t v = 0; v < T.n_void; v++) bad |= V(tv, T.amap[T.void_slots[v]]) != 0.0;
        for (uint32_t k = 0; k < n; k++) {
            uint32_t f = T.fptr[k];
            const uint32_t u1 = T.up[k + 1];
            for (uint32_t p = T.up[k]; p < u1; p++) {
                const uint32_t i = T.ui[p];
                const double u = V(tv, p);
                nonfinite |= lbits(u - u);
                const uint32_t q1 = T.lp[i + 1];
                for (uint32_t q = T.lp[i]; q < q1; q++, f++) {
                    const uint32_t d = T.dst[f];
                    V(tv, d) = V(tv, d) - V(tv, T.lx0 + q) * u;
                }
            }
            const uint32_t fl = T.flags[k];
            if (fl & kVoidBit) {
                V(tv, T.n_u + k) = 1.0;
                continue;
            }
            double dv = V(tv, T.n_u + k);
            const bool dead = dv == 0.0 || !(fabs(dv) < __builtin_inf());
            bad |= dead;
            if (dead) dv = 1.0;
            V(tv, T.n_u + k) = dv;
            const uint32_t l0 = T.lp[k], l1 = T.lp[k + 1];
            if (T.growth > 0 && !(fl & kScaledBit)) {
                uint64_t cb = lbits(fabs(dv));
                for (uint32_t q = l0; q < l1; q++) {
                    const double v = V(tv, T.lx0 + q);
                    const uint64_t vb = lbits(fabs(v));
                    cb = vb > cb ? vb : cb;
                    const double l = v / dv;
                    V(tv, T.lx0 + q) = l;
                    const uint64_t lb = lbits(fabs(l));
                    worst = lb > worst ? lb : worst;
                }
                bad |= !(fabs(dv) >= T.growth * lbitsd(cb));
            } else {
                const bool wide = !(fl & kScaledBit);
                for (uint32_t q = l0; q < l1; q++) {
                    const double l = V(tv, T.lx0 + q) / dv;
                    nonfinite |= lbits(l - l);
                    V(tv, T.lx0 + q) = l;
                    if (wide) {
                        const uint64_t lb = lbits(fabs(l));
                        worst = lb > worst ? lb : worst;
                    }
                }
            }
        }
        lap(1);
        if (bad || nonfinite != 0 || !(lbitsd(worst) <= T.wide)) {
            // hand the lane to a host re-pivot, in the wavefront kernel's layout
            for (uint32_t i = 0; i < n; i++) a.x[(size_t)lane * n + i] = V(x, i);
            const uint32_t lw = lb + 2 * nq;   // the wavefront kernel's limiting-state width
            for (uint32_t k = 0; k < lw; k++) a.lim[(size_t)lane * lw + k] = V(lim, k);
            LaneResume o;
            o.iter = iter;
            o.init_fix = init_fix ? 1u : 0u;
            o.tape = 0;
            o.fresh = 0;
            o.limit = limit;
            o.k0 = 0;
            o.warm = 0;
            o.f0 = 0.0;
            a.resume[lane] = o;
            status = kRepivot;
            iters = iter;
            break;
        }

        // ---- dx = -A^-1 F (lane_kernels.zig:123-143) ----
        for (uint32_t r = 0; r < n; r++) V(y, T.pinv[r]) = -V(rhs, r);
        for (uint32_t k = 0; k < n; k++) {
            const double yk = V(y, k);
            const uint32_t q1 = T.lp[k + 1];
            for (uint32_t q = T.lp[k]; q < q1; q++) {
                const uint32_t r = T.li[q];
                const double old = V(y, r);
                const double s = old - V(tv, T.lx0 + q) * yk;
                V(y, r) = yk != 0.0 ? s : old;
            }
        }
        for (uint32_t k = n; k > 0;) {
            k--;
            const double zk = V(y, k) / V(tv, T.n_u + k);
            V(y, k) = zk;
            const uint32_t p1 = T.up[k + 1];
            for (uint32_t p = T.up[k]; p < p1; p++) {
                const uint32_t i = T.ui[p];
                const double old = V(y, i);
                const double s = old - V(tv, p) * zk;
                V(y, i) = zk != 0.0 ? s : old;
            }
        }
        lap(2);

        // ---- update, limiting, gates (converger.zig:454-522, 794-819) ----
        uint64_t wsc = 0;
        for (uint32_t i = 0; i < n; i++) {
            const double xv = V(x, i);
            const double d = V(y, T.qinv[i]);
            const double xn = xv + d;
            V(xo, i) = xv;
            V(x, i) = xn;
            const double tl = a.reltol * fmax(fabs(xn), fabs(xv)) + (dp.current_row[i] ? a.abstol : a.vntol);
            const bool finite = fabs(xn) < __builtin_inf() && fabs(d) < __builtin_inf();
            const double qq = fabs(d) / tl;
            const double score = !finite ? __builtin_inf() : (qq == qq ? qq : 0.0);
            const uint64_t sb = lbits(score);
            wsc = sb > wsc ? sb : wsc;
        }
        const double scaled = lbitsd(wsc);
        bool limited = false;
        for (uint32_t j = 0; j < nm; j++) {
            const Mos m = mos_load(dp, j, live, a.N);
            const uint32_t *t4 = dp.m_t + 4 * j;
            double lgs = V(lim, j), lds = V(lim, nm + j), lbs = V(lim, 2 * nm + j);
            limited |= mos_limit_v(m, V(x, t4[0]), V(x, t4[1]), V(x, t4[2]), V(x, t4[3]), lgs, lds, lbs,
                                   V(lim, 3 * nm + j));
            V(lim, j) = lgs;
            V(lim, nm + j) = lds;
            V(lim, 2 * nm + j) = lbs;
        }
        for (uint32_t j = 0; j < nd; j++) {
            const Dio d = dio_load(dp, j, live, a.N);
            double lv = V(lim, 4 * nm + j);
            limited |= dio_limit_v(d, V(x, dp.d_t[2 * j]), V(x, dp.d_t[2 * j + 1]), lv);
            V(lim, 4 * nm + j) = lv;
        }
        for (uint32_t j = 0; j < nq; j++) {
            const Bjt b = bjt_load(dp, j, live, a.N);
            const uint32_t *t4 = dp.q_t + 4 * j;
            double lbe = V(lim, lb + 2 * j), lbc = V(lim, lb + 2 * j + 1);
            limited |= bjt_limit_v(b, V(x, t4[0]), V(x, t4[1]), V(x, t4[2]), lbe, lbc);
            V(lim, lb + 2 * j) = lbe;
            V(lim, lb + 2 * j + 1) = lbc;
        }
        bool ok = !limited && iter != 0 && scaled < 1.0;
        if (ok) {
            for (uint32_t i = 0; i < n; i++) {
                const double xi = V(x, i);
                if (!(fabs(xi) <= limit)) ok = false;
                const double s0 = fabs(V(sd, i));
                const double s = s0 == 0.0 ? 1.0 : s0;
                const double rt = fmax(a.residual_tol, 10.0 * s * (a.reltol * fabs(xi) + a.vntol));
                if (!(fabs(V(rhs, i)) <= rt)) ok = false;
            }
        }
        if (ok && init_fix) {
            init_fix = false;
        } else if (ok) {
            for (uint32_t i = 0; i < n; i++) V(x, i) = V(xo, i);
            status = kConverged;
            iters = iter + 1;
            break;
        }
    }
    lap(3);
    if (status != kRepivot)
        for (uint32_t i = 0; i < n; i++) a.x[(size_t)lane * n + i] = V(x, i);
    a.status[lane] = status;
    a.iterations[lane] = iters;
    if (a.prof)
        for (int p = 0; p < kProfPhases; p++) a.prof[(size_t)lane * kProfPhases + p] += pc[p];
#undef V
}

// lane-major copy of point-major live values: out[k * N + l] = in[l * P + k]
__global__ __launch_bounds__(256) void transpose_live(const double *in, double *out, uint32_t N, uint32_t P)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * P) return;
    const uint32_t l = i % N, k = i / N;
    out[i] = in[(size_t)l * P + k];
}

}  // namespace esl

extern "C" hipError_t esl_launch_op_newton_lpt(const esl::LptArgs *a, hipStream_t s)
{
    const uint32_t block = 256;
    const uint32_t grid = (a->N + block - 1) / block;
    hipLaunchKernelGGL(esl::op_newton_lpt, dim3(grid), dim3(block), 0, s, *a);
    return hipGetLastError();
}

extern "C" hipError_t esl_launch_transpose_live(const double *in, double *out, uint32_t N, uint32_t P, hipStream_t s)
{
    const size_t total = (size_t)N * P;
    if (total == 0) return hipSuccess;
    const uint32_t block = 256;
    hipLaunchKernelGGL(esl::transpose_live, dim3((uint32_t)((total + block - 1) / block)), dim3(block), 0, s, in, out,
                       N, P);
    return hipGetLastError();
}
