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

// ---- ngspice limiters (devsup.c, as quoted in the reference's
// dev/solvers/newton-raphson-convergence.md:42-69,151-162) ----
__device__ inline double pnjlim(double vnew, double vold, double vt, double vcrit)
{
    if (vnew > vcrit && fabs(vnew - vold) > vt + vt) {
        if (vold > 0) {
            const double arg = (vnew - vold) / vt;
            if (arg > 0) vnew = vold + vt * (2.0 + esl::log(arg - 2.0));
            else vnew = vold - vt * (2.0 + esl::log(2.0 - arg));
        } else {
            vnew = vt * esl::log(vnew / vt);
        }
    } else if (vnew < 0) {
        const double arg = vold > 0 ? -1.0 * vold - 1.0 : 2.0 * vold - 1.0;
        if (vnew < arg) vnew = arg;
    }
    return vnew;
}

__device__ inline double fetlim(double vnew, double vold, double vto)
{
    const double vtsthi = fabs(2.0 * (vold - vto)) + 2.0;
    const double vtstlo = fabs(vold - vto) + 1.0;
    const double vtox = vto + 3.5;
    const double delv = vnew - vold;
    if (vold >= vto) {
        if (vold >= vtox) {
            if (delv <= 0) {
                if (vnew >= vtox) {
                    if (-delv > vtstlo) vnew = vold - vtstlo;
                } else {
                    vnew = fmax(vnew, vto + 2.0);
                }
            } else if (delv >= vtsthi) {
                vnew = vold + vtsthi;
            }
        } else {
            if (delv <= 0) vnew = fmax(vnew, vto - 0.5);
            else vnew = fmin(vnew, vto + 4.0);
        }
    } else {
        if (delv <= 0) {
            if (-delv > vtsthi) vnew = vold - vtsthi;
        } else {
            const double vtemp = vto + 0.5;
            if (vnew <= vtemp) {
                if (delv > vtstlo) vnew = vold + vtstlo;
            } else {
                vnew = vtemp;
            }
        }
    }
    return vnew;
}

__device__ inline double limvds(double vnew, double vold)
{
    if (vold >= 3.5) {
        if (vnew > vold) vnew = fmin(vnew, 3.0 * vold + 2.0);
        else if (vnew < 3.5) vnew = fmax(vnew, 2.0);
    } else {
        if (vnew > vold) vnew = fmin(vnew, 4.0);
        else vnew = fmax(vnew, -0.5);
    }
    return vnew;
}

// ---- MOS1 (models/mos1.va) ----
struct Mos {
    double tp, vbit, gammax, phit, beta, lambda, vt, isats, isatd, vcrits, vcritd, gmin;
};

__device__ inline Mos mos_load(const DeckProg &dp, uint32_t j, const double *live)
{
    const uint32_t nm = dp.n_mos;
    const double *f = dp.m_par;
    Mos m;
    m.tp = f[MF_TYPE * nm + j];
    const int bv = dp.m_bind_vto[j], bk = dp.m_bind_kp[j];
    const double vtox = bv >= 0 ? live[bv] : f[MF_VTO * nm + j];
    const double kpx = bk >= 0 ? live[bk] : f[MF_KP * nm + j];
    m.vbit = ((vtox - f[MF_VBIT_A * nm + j]) + f[MF_VBIT_B * nm + j]) + f[MF_VBIT_C * nm + j];
    const double kpt = kpx / f[MF_RATIO4 * nm + j];
    m.beta = kpt * f[MF_M * nm + j] * f[MF_W * nm + j] / f[MF_LEFF * nm + j];
    m.gammax = f[MF_GAMMA * nm + j];
    m.phit = f[MF_PHIT * nm + j];
    m.lambda = f[MF_LAMBDA * nm + j];
    m.vt = f[MF_VT * nm + j];
    m.isats = f[MF_ISATS * nm + j];
    m.isatd = f[MF_ISATD * nm + j];
    m.vcrits = f[MF_VCRITS * nm + j];
    m.vcritd = f[MF_VCRITD * nm + j];
    m.gmin = f[MF_GMIN * nm + j];
    return m;
}

__device__ inline void jcur(double v, double isat, double vtx, double gm, double &i, double &g)
{
    if (v > -3.0 * vtx) {
        const double u = fmin(v / vtx, 80.0);
        const double e = esl::exp(u);
        i = isat * (e * (1.0 + (v / vtx - u)) - 1.0) + gm * v;
        g = isat * e / vtx + gm;
    } else {
        i = gm * v - isat;
        g = gm;
    }
}

// Limit step after an update: returns true when any branch voltage moved.
__device__ inline bool mos_limit(const Mos &m, double xd, double xg, double xs, double xb, double *lim,
                                 uint32_t stride)
{
    const double tp = m.tp;
    const double cgs = tp * (xg - xs), cds = tp * (xd - xs), cbs = tp * (xb - xs);
    const double ogs = lim[0], ods = lim[stride], obs = lim[2 * stride], von = lim[3 * stride];
    double vgs = cgs, vds = cds, vbs = cbs;
    double vgd = vgs - vds;
    const double vgdo = ogs - ods;
    if (ods >= 0) {
        vgs = fetlim(vgs, ogs, von);
        vds = vgs - vgd;
        vds = limvds(vds, ods);
    } else {
        vgd = fetlim(vgd, vgdo, von);
        vds = vgs - vgd;
        vds = -limvds(-vds, -ods);
        vgs = vgd + vds;
    }
    if (vds >= 0) {
        vbs = pnjlim(vbs, obs, m.vt, m.vcrits);
    } else {
        const double vbd = pnjlim(vbs - vds, obs - ods, m.vt, m.vcritd);
        vbs = vbd + vds;
    }
    const double scale = fmax(fmax(fabs(xd), fabs(xg)), fmax(fabs(xs), fabs(xb)));
    const double tol = 16.0 * 2.220446049250313e-16 * scale;
    if (fabs(vgs - cgs) <= tol) vgs = cgs;
    if (fabs(vds - cds) <= tol) vds = cds;
    if (fabs(vbs - cbs) <= tol) vbs = cbs;
    lim[0] = vgs;
    lim[stride] = vds;
    lim[2 * stride] = vbs;
    return vgs != cgs || vds != cds || vbs != cbs;
}

// DC evaluation at the limited voltages plus the residual correction; stages
// 11 Jacobian cells (e 0..10) and 3 residual cells (d, s, b at e 11..13).
__device__ inline void mos_eval(const Mos &m, double xd, double xg, double xs, double xb, double *lim,
                                uint32_t stride, double *st, uint32_t t)
{
    const double tp = m.tp;
    const double vgs = lim[0], vds = lim[stride], vbs = lim[2 * stride];
    const double dgs = tp * (xg - xs) - vgs;
    const double dds = tp * (xd - xs) - vds;
    const double dbs = tp * (xb - xs) - vbs;
    const double vbd = vbs - vds, vgd = vgs - vds;
    double ibs, gbs, ibd, gbd;
    jcur(vbs, m.isats, m.vt, m.gmin, ibs, gbs);
    jcur(vbd, m.isatd, m.vt, m.gmin, ibd, gbd);
    double mode, vbx, vgsx;
    if (vds >= 0) {
        mode = 1.0; vbx = vbs; vgsx = vgs;
    } else {
        mode = -1.0; vbx = vbd; vgsx = vgd;
    }
    const double vdsx = mode * vds;
    double sarg;
    if (vbx <= 0) sarg = sqrt(m.phit - vbx);
    else sarg = fmax(sqrt(m.phit) - vbx / (2.0 * sqrt(m.phit)), 0.0);
    const double dsarg = sarg > 0 ? -0.5 / sarg : 0.0;
    const double von = m.vbit * tp + m.gammax * sarg;
    const double vgst = vgsx - von;
    const double bp = m.beta * (1.0 + m.lambda * vdsx);
    double id, gm, gds;
    if (vgst <= 0) {
        id = 0; gm = 0; gds = 0;
    } else if (vgst <= vdsx) {
        id = 0.5 * bp * vgst * vgst;
        gm = bp * vgst;
        gds = 0.5 * m.beta * m.lambda * vgst * vgst;
    } else {
        id = bp * vdsx * (vgst - 0.5 * vdsx);
        gm = bp * vdsx;
        gds = m.beta * m.lambda * vdsx * (vgst - 0.5 * vdsx) + bp * (vgst - vdsx);
    }
    const double gmbs = -gm * (m.gammax * dsarg);
    lim[3 * stride] = von;
    double ags, ads, abs_;
    if (mode > 0) {
        ags = gm; ads = gds; abs_ = gmbs;
    } else {
        ags = gm; ads = -(gm + gds + gmbs); abs_ = gmbs;
    }
    double ich = tp * mode * id;
    double ib_d = tp * ibd, ib_s = tp * ibs;
    if (dgs != 0 || dds != 0 || dbs != 0) {
        ich = ich + tp * mode * ((ags * dgs + ads * dds) + abs_ * dbs);
        ib_d = ib_d + tp * (gbd * (dbs - dds));
        ib_s = ib_s + tp * (gbs * dbs);
    }
    const double cg = mode * ags, cd = mode * ads, cb = mode * abs_;
    const double cs = -((cg + cd) + cb);
    st[0 * kWave + t] = cd + gbd;
    st[1 * kWave + t] = cg;
    st[2 * kWave + t] = cs;
    st[3 * kWave + t] = cb - gbd;
    st[4 * kWave + t] = -cd;
    st[5 * kWave + t] = -cg;
    st[6 * kWave + t] = -cs + gbs;
    st[7 * kWave + t] = -cb - gbs;
    st[8 * kWave + t] = -gbd;
    st[9 * kWave + t] = -gbs;
    st[10 * kWave + t] = gbd + gbs;
    st[11 * kWave + t] = ich - ib_d;
    st[12 * kWave + t] = -ich - ib_s;
    st[13 * kWave + t] = ib_d + ib_s;
}

// ---- assembly: every device, chunk by chunk, into tv (through amap) and rhs.
// amap == nullptr stamps pattern slots directly (the dump of A).
__device__ void assemble(const DeckProg &dp, const uint32_t *amap, const double *live, const double *x,
                         double *lim, uint32_t lim_stride, double *tv, uint32_t n_tv, double *rhs,
                         double *st, uint32_t tid)
{
    const uint32_t n = dp.n;
    for (uint32_t i = tid; i < n_tv; i += kWave) tv[i] = 0.0;
    for (uint32_t i = tid; i < n; i += kWave) rhs[i] = 0.0;
    __syncthreads();
    for (uint32_t c = 0; c < dp.n_chunks; c++) {
        const Chunk ch = dp.chunks[c];
        if (tid < ch.count) {
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
                const uint32_t jm = j;
                const Mos m = mos_load(dp, jm, live);
                const uint32_t *t4 = dp.m_t + 4 * jm;
                mos_eval(m, x[t4[0]], x[t4[1]], x[t4[2]], x[t4[3]], lim + jm, lim_stride, st, tid);
                break;
            }
            default:
                break;
            }
        }
        __syncthreads();
        const uint32_t t1 = dp.chunks[c + 1].task_begin;
        for (uint32_t t = ch.task_begin + tid; t < t1; t += kWave) {
            const uint32_t tgt = dp.task_target[t];
            double *d = (tgt & kRhsBit) ? &rhs[tgt & ~kRhsBit] : &tv[amap ? amap[tgt] : tgt];
            double acc = *d;
            const uint32_t c1 = dp.task_cell[t + 1];
            for (uint32_t cc = dp.task_cell[t]; cc < c1; cc++) acc = acc + st[dp.cells[cc]];
            *d = acc;
        }
        __syncthreads();
    }
    if (tid == 0) {
        const uint32_t s0 = amap ? amap[dp.diag0_slot] : dp.diag0_slot;
        tv[s0] = tv[s0] + 1.0;
        rhs[0] = rhs[0] + x[0];
    }
    __syncthreads();
}

// ---- refactor: true when the lane must re-pivot (lane_kernels.zig:49-118) ----
__device__ bool refactor(const TapeProg &T, double *tv, uint32_t tid)
{
    bool bad = false;
    uint64_t nonfinite = 0, worst = 0;
    for (uint32_t v = tid; v < T.n_void; v += kWave) bad |= tv[T.amap[T.void_slots[v]]] != 0.0;
    for (uint32_t lev = 0; lev < T.n_levels; lev++) {
        const uint32_t c1 = T.lev_ptr[lev + 1];
        for (uint32_t c = T.lev_ptr[lev] + tid; c < c1; c += kWave) {
            const uint32_t k = T.lev_cols[c];
            uint32_t f = T.fptr[k];
            const uint32_t u1 = T.up[k + 1];
            for (uint32_t p = T.up[k]; p < u1; p++) {
                const uint32_t i = T.ui[p];
                const double u = tv[p];
                nonfinite |= dbits(u - u);
                const uint32_t q1 = T.lp[i + 1];
                for (uint32_t q = T.lp[i]; q < q1; q++, f++) {
                    const uint32_t d = T.dst[f];
                    tv[d] = tv[d] - tv[T.lx0 + q] * u;
                }
            }
            const uint32_t fl = T.flags[k];
            if (fl & kVoidBit) {
                tv[T.n_u + k] = 1.0;
                continue;
            }
            double dp = tv[T.n_u + k];
            const bool dead = dp == 0.0 || !(fabs(dp) < __builtin_inf());
            bad |= dead;
            if (dead) dp = 1.0;
            tv[T.n_u + k] = dp;
            const uint32_t l0 = T.lp[k], l1 = T.lp[k + 1];
            if (T.growth > 0 && !(fl & kScaledBit)) {
                uint64_t cb = dbits(fabs(dp));
                for (uint32_t q = l0; q < l1; q++) {
                    const double v = tv[T.lx0 + q];
                    const uint64_t vb = dbits(fabs(v));
                    cb = vb > cb ? vb : cb;
                    const double l = v / dp;
                    tv[T.lx0 + q] = l;
                    const uint64_t lb = dbits(fabs(l));
                    worst = lb > worst ? lb : worst;
                }
                bad |= !(fabs(dp) >= T.growth * bitsd(cb));
            } else {
                const bool wide = !(fl & kScaledBit);
                for (uint32_t q = l0; q < l1; q++) {
                    const double l = tv[T.lx0 + q] / dp;
                    nonfinite |= dbits(l - l);
                    tv[T.lx0 + q] = l;
                    if (wide) {
                        const uint64_t lb = dbits(fabs(l));
                        worst = lb > worst ? lb : worst;
                    }
                }
            }
        }
        __syncthreads();
    }
    worst = wave_max_u64(worst);
    const bool fail = bad || nonfinite != 0 || !(bitsd(worst) <= T.wide);
    return wave_any(fail);
}

// ---- dx = -A^-1 rhs into y (pivot order); row i's dx is y[qinv[i]] ----
__device__ void solve_neg(const TapeProg &T, uint32_t n, const double *tv, const double *rhs, double *y,
                          uint32_t tid)
{
    for (uint32_t r = tid; r < n; r += kWave) y[T.pinv[r]] = -rhs[r];
    __syncthreads();
    for (uint32_t lev = 0; lev < T.n_flev; lev++) {
        const uint32_t e1 = T.flev_ptr[lev + 1];
        for (uint32_t e = T.flev_ptr[lev] + tid; e < e1; e += kWave) {
            const uint32_t k = T.flev_rows[e];
            double acc = y[k];
            const uint32_t p1 = T.lrow_ptr[k + 1];
            for (uint32_t p = T.lrow_ptr[k]; p < p1; p++) {
                const double yj = y[T.lrow_k[p]];
                const double s = acc - tv[T.lrow_slot[p]] * yj;
                acc = yj != 0.0 ? s : acc;
            }
            y[k] = acc;
        }
        __syncthreads();
    }
    for (uint32_t lev = 0; lev < T.n_blev; lev++) {
        const uint32_t e1 = T.blev_ptr[lev + 1];
        for (uint32_t e = T.blev_ptr[lev] + tid; e < e1; e += kWave) {
            const uint32_t i = T.blev_rows[e];
            double acc = y[i];
            const uint32_t p1 = T.urow_ptr[i + 1];
            for (uint32_t p = T.urow_ptr[i]; p < p1; p++) {
                const double zk = y[T.urow_k[p]];
                const double s = acc - tv[T.urow_slot[p]] * zk;
                acc = zk != 0.0 ? s : acc;
            }
            y[i] = acc / tv[T.n_u + i];
        }
        __syncthreads();
    }
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
    double *tv = sm;
    double *rhs = tv + a.tv_cap;
    double *x = rhs + n;
    double *y = x + n;
    double *xold = y + n;
    double *sdiag = xold + n;
    double *st = sdiag + n;
    const double *live = a.live + (size_t)lane * dp.n_live;
    double *lim = a.lim + (size_t)lane * 4 * nm;
    lane_init(a, lane, live, x, lim, tid);

    if (a.dump) {
        assemble(dp, nullptr, live, x, lim, nm, tv, dp.nnz, rhs, st, tid);
        for (uint32_t p = tid; p < dp.nnz; p += kWave) a.a_out[(size_t)blockIdx.x * dp.nnz + p] = tv[p];
        for (uint32_t i = tid; i < n; i += kWave) a.f_out[(size_t)blockIdx.x * n + i] = rhs[i];
        return;
    }

    LaneResume rs = a.resume[lane];
    const TapeProg T = a.tapes[rs.tape];
    uint32_t iter0 = 0;
    bool init_fix = a.mode == 0;
    double limit = 1e9;
    if (a.mode == 2) {
        iter0 = rs.iter;
        init_fix = rs.init_fix != 0;
        limit = rs.limit;
    }
    uint32_t status = kNoConvergence, iters = a.max_iter;
    for (uint32_t iter = iter0; iter < a.max_iter; iter++) {
        assemble(dp, T.amap, live, x, lim, nm, tv, T.n_tv, rhs, st, tid);
        if (iter == 0) {
            double gm = 1.0;
            uint64_t xm = 0, fm = 0;
            for (uint32_t i = tid; i < n; i += kWave) {
                gm = fmax(gm, fabs(tv[T.amap[dp.diag_slot[i]]]));
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
        for (uint32_t i = tid; i < n; i += kWave) sdiag[i] = tv[T.amap[dp.diag_slot[i]]];
        __syncthreads();
        if (refactor(T, tv, tid)) {
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
    __syncthreads();
    if (status != kRepivot)
        for (uint32_t i = tid; i < n; i += kWave) a.x[(size_t)lane * n + i] = x[i];
    if (tid == 0) {
        a.status[lane] = status;
        a.iterations[lane] = iters;
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
    double *rhs = tv + T.n_tv;
    double *y = rhs + n;
    for (uint32_t i = tid; i < T.n_tv; i += kWave) tv[i] = 0.0;
    __syncthreads();
    for (uint32_t p = tid; p < r.nnz; p += kWave) tv[T.amap[p]] = r.a[(size_t)lane * r.nnz + p];
    for (uint32_t i = tid; i < n; i += kWave) rhs[i] = r.rhs[(size_t)lane * n + i];
    __syncthreads();
    const bool bad = refactor(T, tv, tid);
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
