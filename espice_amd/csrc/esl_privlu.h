// A lane's own LU on the device: the re-pivot without a trip to the host.
//
// When a lane's replay of its shared tape fails the peel tests, the lane
// factors its Jacobian itself with threshold partial pivoting (left-looking
// Gilbert-Peierls, sparse_lu.zig:260-622; the host's gp_factor and the
// oracle's orc_lu_factor state the same steps), keeps the factors and pivot
// sequence in its own slice of memory (the pattern in LDS), and from then on replays
// them column by column (lane_kernels.zig:49-118 with the wide-multiplier
// test), factoring again whenever the replay fails. Every operation runs in
// the oracle's order on one thread, so the lane's numbers are bitwise those
// of the host re-pivot path (a host factor, then replays of that tape).
#pragma once
#include <stdint.h>

#include "esl_eval.h"
#include "esl_internal.h"

namespace esl {

// One lane's view: the pattern of its factors (lp, up, li, ui) and the
// dense column w live in the lane's LDS (the factor-slot area, unused while
// the lane is on its own LU), with copies in global memory that carry them
// from one launch to the next; values, pivots and work arrays stay global.
struct PrivLane {
    uint32_t n, nnz, cap, lcap;   // cap: L and U entries in global memory; lcap: in LDS
    uint32_t *hdr, *pinv, *flags, *flag, *topo, *stack, *pstack, *want, *want_of, *prow, *void_slots;
    uint32_t *lp, *up, *li, *ui;        // LDS
    uint32_t *glp, *gup, *gli, *gui;    // global copies
    double *lx, *ux, *udiag, *w, *rscale, *ap;
};

// lds: the lane's factor-slot area, lds_doubles long.
__device__ inline PrivLane priv_lane(const PrivLU &P, uint32_t lane, double *lds, uint32_t lds_doubles)
{
    const size_t n = P.n, n1 = P.n + 1, z = P.nnz, c = P.cap, l = lane;
    PrivLane L;
    L.n = P.n;
    L.nnz = P.nnz;
    L.cap = P.cap;
    L.hdr = P.hdr + l * 4;
    L.pinv = P.pinv + l * n;
    L.flags = P.flags + l * n;
    L.flag = P.flag + l * n;
    L.topo = P.topo + l * n;
    L.stack = P.stack + l * n;
    L.pstack = P.pstack + l * n;
    L.want = P.want + l * n;
    L.want_of = P.want_of + l * n;
    L.glp = P.lp + l * n1;
    L.gup = P.up + l * n1;
    L.prow = P.prow + l * z;
    L.void_slots = P.void_slots + l * z;
    L.gli = P.li + l * c;
    L.gui = P.ui + l * c;
    L.lx = P.lx + l * c;
    L.ux = P.ux + l * c;
    L.udiag = P.udiag + l * n;
    L.rscale = P.rscale + l * n;
    L.ap = P.ap + l * z;
    // LDS: [w (n doubles) | lp (n + 1) | up (n + 1) | li (lcap) | ui (lcap)] in u32 after w
    L.w = lds;
    const int64_t words = 2 * ((int64_t)lds_doubles - (int64_t)n) - 2 * (int64_t)n1;
    L.lcap = words > 0 ? (uint32_t)(words / 2) : 0u;
    L.lp = reinterpret_cast<uint32_t *>(lds + n);
    L.up = L.lp + n1;
    L.li = L.up + n1;
    L.ui = L.li + L.lcap;
    return L;
}

__device__ inline uint64_t pl_bits(double d) { return __builtin_bit_cast(uint64_t, d); }
__device__ inline double pl_double(uint64_t b) { return __builtin_bit_cast(double, b); }

constexpr uint32_t kPlNone = 0xffffffffu;
constexpr int kPlOk = 0, kPlSingular = -1, kPlFull = -2;

// voidUnknown (sparse_lu.zig:847-871): column c and row c are structurally
// present but hold only zeros, and c has no pivot yet.
__device__ inline bool pl_void_unknown(const PrivLane &L, const uint32_t *col_ptr, const uint32_t *row_idx,
                                       const double *vals, uint32_t c)
{
    if (L.pinv[c] != kPlNone) return false;
    for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++)
        if (vals[p] != 0) return false;
    for (uint32_t j = 0; j < L.n; j++) {
        if (j == c) continue;
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++)
            if (row_idx[p] == c && vals[p] != 0) return false;
    }
    return true;
}

// The full factor of L.ap with threshold `pivot_tol` (one thread). kPlFull:
// the factors outgrew the lane's capacity (w is left all zero).
__device__ int priv_factor(const PrivLane &L, const uint32_t *col_ptr, const uint32_t *row_idx, const uint32_t *q,
                           double pivot_tol)
{
    const uint32_t n = L.n, nnz = L.nnz;
    const double *vals = L.ap;
    double *w = L.w;
    for (uint32_t i = 0; i < n; i++) {
        L.pinv[i] = kPlNone;
        L.flags[i] = 0;
        L.flag[i] = 0;
        L.rscale[i] = 0.0;
    }
    uint32_t nnz_l = 0, nnz_u = 0, n_void = 0;
    bool has_void = false;
    uint64_t amaxb = 0;
    for (uint32_t p = 0; p < nnz; p++) {
        const uint64_t bb = pl_bits(fabs(vals[p]));
        amaxb = bb > amaxb ? bb : amaxb;
    }
    if (!(pl_double(amaxb) <= 1.7976931348623157e308)) return kPlSingular;
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++)
            L.rscale[row_idx[p]] = fmax(L.rscale[row_idx[p]], fabs(vals[p]));
    for (uint32_t i = 0; i < n; i++) {
        const double r = L.rscale[i];
        L.rscale[i] = (r > 0 && r < __builtin_inf()) ? 1.0 / r : 1.0;
    }
    uint64_t nonfinite = 0;
    for (uint32_t i = 0; i < n; i++) {
        L.want[i] = i;
        L.want_of[i] = i;
    }
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t c = q[k];
        L.lp[k] = nnz_l;
        L.up[k] = nnz_u;
        const uint32_t mark = k + 1;
        // reach of A[:, c] through the graph of L (rows still original)
        uint32_t nt = 0;
        for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) {
            uint32_t r = row_idx[p];
            if (L.flag[r] == mark) continue;
            uint32_t sp = 0;
            L.stack[0] = r;
            L.pstack[0] = L.pinv[r] == kPlNone ? 0 : L.lp[L.pinv[r]];
            L.flag[r] = mark;
            for (;;) {
                r = L.stack[sp];
                const uint32_t kc = L.pinv[r];
                const uint32_t end = kc == kPlNone ? 0 : L.lp[kc + 1];
                uint32_t pos = L.pstack[sp];
                bool descended = false;
                while (pos < end) {
                    const uint32_t child = L.li[pos];
                    pos++;
                    if (L.flag[child] != mark) {
                        L.pstack[sp] = pos;
                        L.flag[child] = mark;
                        sp++;
                        L.stack[sp] = child;
                        L.pstack[sp] = L.pinv[child] == kPlNone ? 0 : L.lp[L.pinv[child]];
                        descended = true;
                        break;
                    }
                }
                if (descended) continue;
                L.topo[nt++] = r;
                if (sp == 0) break;
                sp--;
            }
        }
        const uint32_t room = L.cap < L.lcap ? L.cap : L.lcap;
        if (nnz_u + nt > room || nnz_l + nt > room) {
            for (uint32_t i = 0; i < n; i++) w[i] = 0.0;
            return kPlFull;
        }
        for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) w[row_idx[p]] = vals[p];
        // triangular solve in reverse finish order
        for (uint32_t idx = nt; idx > 0;) {
            idx--;
            const uint32_t r = L.topo[idx], kc = L.pinv[r];
            if (kc == kPlNone) continue;
            const double ukr = w[r];
            L.ui[nnz_u] = kc;
            L.ux[nnz_u] = ukr;
            nnz_u++;
            for (uint32_t pl = L.lp[kc]; pl < L.lp[kc + 1]; pl++) {
                const double t = L.lx[pl] * ukr;
                w[L.li[pl]] = w[L.li[pl]] - t;
            }
        }
        double amax = 0, smax = 0;
        uint32_t piv = kPlNone;
        for (uint32_t s = 0; s < nt; s++) {
            const uint32_t r = L.topo[s];
            if (L.pinv[r] != kPlNone) continue;
            const double a = fabs(w[r]);
            if (a > amax) {
                amax = a;
                piv = r;
            }
            smax = fmax(smax, a * L.rscale[r]);
        }
        if (piv == kPlNone || amax == 0 || !(amax < __builtin_inf())) {
            const bool is_void = pl_void_unknown(L, col_ptr, row_idx, vals, c);
            for (uint32_t s = 0; s < nt; s++) w[L.topo[s]] = 0;
            if (!is_void) return kPlSingular;
            L.flags[k] |= kVoidBit;
            has_void = true;
            L.udiag[k] = 1;
            L.pinv[c] = k;
            continue;
        }
        const uint32_t pr = L.want[c];
        if (L.pinv[pr] == kPlNone) {
            const double dmag = fabs(w[pr]);
            if (dmag > 0 && dmag >= pivot_tol * amax) {
                piv = pr;
            } else if (dmag > 0 && dmag * L.rscale[pr] >= pivot_tol * smax) {
                piv = pr;
                L.flags[k] |= kScaledBit;
            }
        }
        if (piv != pr) {
            const uint32_t c2 = L.want_of[piv];
            L.want[c2] = pr;
            L.want_of[pr] = c2;
        }
        const double d = w[piv];
        L.udiag[k] = d;
        L.pinv[piv] = k;
        for (uint32_t s = 0; s < nt; s++) {
            const uint32_t r = L.topo[s];
            const double v = w[r];
            nonfinite |= pl_bits(v - v);
            if (L.pinv[r] == kPlNone) {
                const double l = v / d;
                nonfinite |= pl_bits(l - l);
                L.li[nnz_l] = r;
                L.lx[nnz_l] = l;
                nnz_l++;
            }
            w[r] = 0;
        }
    }
    L.lp[n] = nnz_l;
    L.up[n] = nnz_u;
    if (nonfinite != 0) return kPlSingular;
    for (uint32_t p = 0; p < nnz_l; p++) L.li[p] = L.pinv[L.li[p]];
    for (uint32_t p = 0; p < nnz; p++) L.prow[p] = L.pinv[row_idx[p]];
    if (has_void)
        for (uint32_t j = 0; j < n; j++) {
            const uint32_t c = q[j];
            for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++)
                if ((L.flags[j] & kVoidBit) || (L.flags[L.prow[p]] & kVoidBit)) L.void_slots[n_void++] = p;
        }
    L.hdr[0] = nnz_l;
    L.hdr[1] = nnz_u;
    L.hdr[2] = n_void;
    // the pattern's global copy, for the lane's later launches
    for (uint32_t k = 0; k <= n; k++) {
        L.glp[k] = L.lp[k];
        L.gup[k] = L.up[k];
    }
    for (uint32_t p = 0; p < nnz_l; p++) L.gli[p] = L.li[p];
    for (uint32_t p = 0; p < nnz_u; p++) L.gui[p] = L.ui[p];
    return kPlOk;
}

// A launch's start on the lane's own LU: the pattern back into LDS (all
// threads of the wavefront). False when it does not fit this launch's LDS.
__device__ inline bool priv_stage(const PrivLane &L, uint32_t tid, uint32_t wave)
{
    const uint32_t nl = L.hdr[0], nu = L.hdr[1];
    if (nl > L.lcap || nu > L.lcap) return false;
    for (uint32_t k = tid; k <= L.n; k += wave) {
        L.lp[k] = L.glp[k];
        L.up[k] = L.gup[k];
    }
    for (uint32_t p = tid; p < nl; p += wave) L.li[p] = L.gli[p];
    for (uint32_t p = tid; p < nu; p += wave) L.ui[p] = L.gui[p];
    return true;
}

// The lane's factor values (lx, ux, udiag) live in the LDS past the
// lane's working set when nnz_l + nnz_u + n doubles fit there (`tail`,
// tail_n doubles), and in global memory otherwise. Uniform.
__device__ inline bool priv_resident(PrivLane &L, double *tail, uint32_t tail_n)
{
    const uint32_t nl = uni(L.hdr[0]), nu = uni(L.hdr[1]);
    if ((uint64_t)nl + nu + L.n > tail_n) return false;
    L.lx = tail;
    L.ux = tail + nl;
    L.udiag = tail + nl + nu;
    return true;
}

template <bool kRes>
__device__ inline auto pl_vals(double *p)
{
    if constexpr (kRes) return lptr(p);
    else return gptr(p);
}

// The lane replay of its own factors on L.ap (lane_kernels.zig:49-118 with
// the wide-multiplier test), spread over the wavefront: column k's scatter,
// each U entry's update of the dense column (the targets of one L column
// are distinct rows, so its terms go to different threads while the U
// entries stay in their order), and the division of column k by its pivot
// run one term per thread; every slot takes the serial replay's operations
// in its order, so the factors are bitwise the serial replay's. LDS
// operations of one wavefront complete in issue order, which carries each
// step's writes to the next step's reads. True when a peel test fails (the
// lane then factors again).
template <bool kRes>
__device__ bool priv_replay(const PrivLane &L, const uint32_t *col_ptr_, const uint32_t *q_, double growth_limit,
                            double wide, uint32_t tid)
{
    const uint32_t n = L.n;
    const auto vals = gptr(L.ap);
    const auto prow = gptr(L.prow);
    const auto flags = gptr(L.flags);
    const auto col_ptr = gptr(col_ptr_);
    const auto q = gptr(q_);
    const auto w = lptr(L.w);
    const auto lp = lptr(L.lp), up = lptr(L.up), li = lptr(L.li), ui = lptr(L.ui);
    const auto lx = pl_vals<kRes>(L.lx), ux = pl_vals<kRes>(L.ux), ud = pl_vals<kRes>(L.udiag);
    const uint32_t n_void = uni(L.hdr[2]);
    bool nz = false;
    for (uint32_t v = tid; v < n_void; v += kWave) nz = nz || vals[gptr(L.void_slots)[v]] != 0;
    if (wave_any(nz)) return true;
    uint64_t nonfinite = 0, worst = 0;
    bool bad = false;
    uint32_t u0 = uni(up[0]);
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t c = uni(q[k]);
        const uint32_t p0 = uni(col_ptr[c]), p1 = uni(col_ptr[c + 1]);
        for (uint32_t p = p0 + tid; p < p1; p += kWave) w[prow[p]] = vals[p];
        const uint32_t u1 = uni(up[k + 1]);
        for (uint32_t p = u0; p < u1; p++) {
            const uint32_t i = uni(ui[p]);
            const double uki = uni(w[i]);
            const uint32_t l0 = uni(lp[i]), l1 = uni(lp[i + 1]);
            nonfinite |= dbits(uki - uki);
            for (uint32_t pl = l0 + tid; pl < l1; pl += kWave) {
                const uint32_t r = li[pl];
                const double t = lx[pl] * uki;
                w[r] = w[r] - t;
            }
            // row i is no target of its own L column
            if (tid == 0) {
                w[i] = 0;
                ux[p] = uki;
            }
        }
        const uint32_t fk = uni(flags[k]);
        const uint32_t l0 = uni(lp[k]), l1 = uni(lp[k + 1]);
        if (fk & kVoidBit) {
            if (tid == 0) {
                ud[k] = 1;
                for (uint32_t p = l0; p < l1; p++) {
                    lx[p] = 0;
                    w[li[p]] = 0;
                }
                w[k] = 0;
                for (uint32_t p = u0; p < u1; p++) {
                    const uint32_t i = ui[p];
                    for (uint32_t pl = lp[i]; pl < lp[i + 1]; pl++) w[li[pl]] = 0;
                }
            }
            u0 = u1;
            continue;
        }
        u0 = u1;
        double d = uni(w[k]);
        const bool dead = d == 0 || !(fabs(d) < __builtin_inf());
        bad = bad || dead;
        if (dead) d = 1;
        if (tid == 0) {
            w[k] = 0;
            ud[k] = d;
        }
        const bool scaled = (fk & kScaledBit) != 0;
        if (growth_limit > 0 && !scaled) {
            uint64_t cb = 0;
            for (uint32_t pl = l0 + tid; pl < l1; pl += kWave) {
                const uint32_t r = li[pl];
                const double v = w[r];
                w[r] = 0;
                const uint64_t vb = dbits(fabs(v));
                cb = vb > cb ? vb : cb;
                const double lv = v / d;
                lx[pl] = lv;
                const uint64_t lb = dbits(fabs(lv));
                worst = lb > worst ? lb : worst;
            }
            cb = wave_max_u64(cb);
            const uint64_t db = dbits(fabs(d));
            cb = db > cb ? db : cb;
            bad = bad || !(fabs(d) >= growth_limit * bitsd(cb));
        } else {
            for (uint32_t pl = l0 + tid; pl < l1; pl += kWave) {
                const uint32_t r = li[pl];
                const double lv = w[r] / d;
                nonfinite |= dbits(lv - lv);
                lx[pl] = lv;
                w[r] = 0;
                const uint64_t lb = dbits(fabs(lv));
                if (!scaled) worst = lb > worst ? lb : worst;
            }
        }
    }
    if (wave_any(nonfinite != 0)) return true;
    worst = wave_max_u64(worst);
    return bad || !(bitsd(worst) <= wide);
}

// dx = -A^-1 rhs from the lane's factors, spread over the wavefront as the
// replay (the terms of one column's update go to different threads, the
// columns stay in their order); dx in unknown order. rhs and dx are in LDS;
// w serves as the work vector and is all zero again on return.
template <bool kRes>
__device__ void priv_solve(const PrivLane &L, const uint32_t *q_, const double *rhs_, double *dx_, uint32_t tid)
{
    const uint32_t n = L.n;
    const auto y = lptr(L.w);
    const auto rhs = lptr(rhs_);
    const auto dx = lptr(dx_);
    const auto pinv = gptr(L.pinv);
    const auto q = gptr(q_);
    const auto lp = lptr(L.lp), up = lptr(L.up), li = lptr(L.li), ui = lptr(L.ui);
    const auto lx = pl_vals<kRes>(L.lx), ux = pl_vals<kRes>(L.ux), ud = pl_vals<kRes>(L.udiag);
    for (uint32_t r = tid; r < n; r += kWave) y[pinv[r]] = -rhs[r];
    uint32_t l0 = uni(lp[0]);
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t l1 = uni(lp[k + 1]);
        const double yk = uni(y[k]);
        if (yk != 0)
            for (uint32_t p = l0 + tid; p < l1; p += kWave) {
                const uint32_t r = li[p];
                const double t = lx[p] * yk;
                y[r] = y[r] - t;
            }
        l0 = l1;
    }
    uint32_t u1 = uni(up[n]);
    for (uint32_t k = n; k > 0;) {
        k--;
        const uint32_t u0 = uni(up[k]);
        const double zk = uni(y[k]) / uni(ud[k]);
        if (tid == 0) y[k] = zk;
        if (zk != 0)
            for (uint32_t p = u0 + tid; p < u1; p += kWave) {
                const uint32_t r = ui[p];
                const double t = ux[p] * zk;
                y[r] = y[r] - t;
            }
        u1 = u0;
    }
    for (uint32_t j = tid; j < n; j += kWave) {
        dx[q[j]] = y[j];
        y[j] = 0.0;
    }
}

// A fresh factor's values into the LDS tail when they fit there (all
// threads; the factor wrote them to global memory).
__device__ inline void priv_load(const PrivLane &G, double *tail, uint32_t tail_n, uint32_t tid)
{
    PrivLane L = G;
    if (!priv_resident(L, tail, tail_n)) return;
    const uint32_t nl = uni(G.hdr[0]), nu = uni(G.hdr[1]);
    for (uint32_t p = tid; p < nl; p += kWave) L.lx[p] = G.lx[p];
    for (uint32_t p = tid; p < nu; p += kWave) L.ux[p] = G.ux[p];
    for (uint32_t k = tid; k < G.n; k += kWave) L.udiag[k] = G.udiag[k];
}

// Entry points for the lane kernels, called by every thread of the lane's
// wavefront. lds: the lane's factor-slot area (lds_doubles long); its w is
// all zero on entry and on return. tail: LDS past the lane's working set
// (tail_n doubles) for the factor values.

// The lane's full factor of its Jacobian (PrivLU::ap of the lane) on one
// thread: kPlOk, kPlSingular or kPlFull (uniform).
__device__ __forceinline__ int priv_factor_at(const PrivLU &P, uint32_t lane, const uint32_t *col_ptr,
                                              const uint32_t *row_idx, const uint32_t *q, double *lds,
                                              uint32_t lds_doubles, double *tail, uint32_t tail_n, uint32_t tid)
{
    const PrivLane L = priv_lane(P, lane, lds, lds_doubles);
    int rc = kPlOk;
    if (tid == 0) rc = priv_factor(L, col_ptr, row_idx, q, 1e-3);
    rc = (int)uni((uint32_t)rc);
    wsync();
    if (rc == kPlOk) priv_load(L, tail, tail_n, tid);
    return rc;
}

// The replay of the lane's own factors, and on a failed peel test its full
// factor: kPlOk, kPlSingular or kPlFull (uniform). fresh: the last factor
// met a singular matrix and left no usable LU.
__device__ __forceinline__ int priv_update_at(const PrivLU &P, uint32_t lane, const uint32_t *col_ptr,
                                              const uint32_t *row_idx, const uint32_t *q, double *lds,
                                              uint32_t lds_doubles, double *tail, uint32_t tail_n, uint32_t tid,
                                              bool fresh = false)
{
    if (!fresh) {
        PrivLane L = priv_lane(P, lane, lds, lds_doubles);
        const bool failed = priv_resident(L, tail, tail_n) ? priv_replay<true>(L, col_ptr, q, 1e-12, 1.0 / 1e-3, tid)
                                                           : priv_replay<false>(L, col_ptr, q, 1e-12, 1.0 / 1e-3, tid);
        if (!failed) return kPlOk;
        wsync();
    }
    return priv_factor_at(P, lane, col_ptr, row_idx, q, lds, lds_doubles, tail, tail_n, tid);
}

// dx = -A^-1 rhs in unknown order (rhs and dx in LDS).
__device__ __forceinline__ void priv_solve_at(const PrivLU &P, uint32_t lane, const uint32_t *q, const double *rhs,
                                              double *dx, double *lds, uint32_t lds_doubles, double *tail,
                                              uint32_t tail_n, uint32_t tid)
{
    PrivLane L = priv_lane(P, lane, lds, lds_doubles);
    if (priv_resident(L, tail, tail_n)) priv_solve<true>(L, q, rhs, dx, tid);
    else priv_solve<false>(L, q, rhs, dx, tid);
}

// A launch's start for a lane already on its own LU: w cleared and the
// pattern staged (all threads); false when it does not fit. The values are
// the next replay's (a replay always precedes the next solve).
__device__ __forceinline__ bool priv_enter(const PrivLU &P, uint32_t lane, double *lds, uint32_t lds_doubles,
                                           uint32_t tid, uint32_t wave)
{
    const PrivLane L = priv_lane(P, lane, lds, lds_doubles);
    for (uint32_t i = tid; i < L.n; i += wave) L.w[i] = 0.0;
    return priv_stage(L, tid, wave);
}

}  // namespace esl
