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

// The lane replay of its own factors on L.ap (one thread); true when a
// peel test fails (the lane then factors again).
__device__ bool priv_refactor(const PrivLane &L, const uint32_t *col_ptr, const uint32_t *q, double growth_limit,
                              double wide)
{
    const uint32_t n = L.n;
    const double *vals = L.ap;
    double *w = L.w;
    bool bad = false;
    uint64_t nonfinite = 0;
    const uint32_t n_void = L.hdr[2];
    for (uint32_t v = 0; v < n_void; v++)
        if (vals[L.void_slots[v]] != 0) return true;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t c = q[k];
        for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) w[L.prow[p]] = vals[p];
        for (uint32_t p = L.up[k]; p < L.up[k + 1]; p++) {
            const uint32_t i = L.ui[p];
            const double uki = w[i];
            w[i] = 0;
            L.ux[p] = uki;
            nonfinite |= pl_bits(uki - uki);
            for (uint32_t pl = L.lp[i]; pl < L.lp[i + 1]; pl++) {
                const double t = L.lx[pl] * uki;
                w[L.li[pl]] = w[L.li[pl]] - t;
            }
        }
        if (L.flags[k] & kVoidBit) {
            L.udiag[k] = 1;
            for (uint32_t p = L.lp[k]; p < L.lp[k + 1]; p++) {
                L.lx[p] = 0;
                w[L.li[p]] = 0;
            }
            w[k] = 0;
            for (uint32_t p = L.up[k]; p < L.up[k + 1]; p++) {
                const uint32_t i = L.ui[p];
                for (uint32_t pl = L.lp[i]; pl < L.lp[i + 1]; pl++) w[L.li[pl]] = 0;
            }
            continue;
        }
        double d = w[k];
        w[k] = 0;
        const bool dead = d == 0 || !(fabs(d) < __builtin_inf());
        bad = bad || dead;
        if (dead) d = 1;
        L.udiag[k] = d;
        if (growth_limit > 0 && !(L.flags[k] & kScaledBit)) {
            int64_t cbits = (int64_t)pl_bits(fabs(d));
            for (uint32_t p = L.lp[k]; p < L.lp[k + 1]; p++) {
                const uint32_t r = L.li[p];
                const double v = w[r];
                w[r] = 0;
                const int64_t vb = (int64_t)pl_bits(fabs(v));
                if (vb > cbits) cbits = vb;
                L.lx[p] = v / d;
            }
            const double cmax = pl_double((uint64_t)cbits);
            bad = bad || !(fabs(d) >= growth_limit * cmax);
        } else {
            for (uint32_t p = L.lp[k]; p < L.lp[k + 1]; p++) {
                const uint32_t r = L.li[p];
                const double lv = w[r] / d;
                nonfinite |= pl_bits(lv - lv);
                L.lx[p] = lv;
                w[r] = 0;
            }
        }
    }
    if (nonfinite != 0) return true;
    uint64_t worst = 0;
    for (uint32_t k = 0; k < n; k++) {
        if (L.flags[k] & kScaledBit) continue;
        for (uint32_t p = L.lp[k]; p < L.lp[k + 1]; p++) {
            const uint64_t b = pl_bits(fabs(L.lx[p]));
            worst = b > worst ? b : worst;
        }
    }
    return bad || !(pl_double(worst) <= wide);
}

// dx = -A^-1 rhs from the lane's factors (one thread); dx in unknown
// order. y: n doubles of scratch.
__device__ void priv_solve_neg(const PrivLane &L, const uint32_t *q, const double *rhs, double *dx, double *y)
{
    const uint32_t n = L.n;
    for (uint32_t r = 0; r < n; r++) y[L.pinv[r]] = -rhs[r];
    for (uint32_t k = 0; k < n; k++) {
        const double yk = y[k];
        if (yk != 0)
            for (uint32_t p = L.lp[k]; p < L.lp[k + 1]; p++) {
                const double t = L.lx[p] * yk;
                y[L.li[p]] = y[L.li[p]] - t;
            }
    }
    for (uint32_t k = n; k > 0;) {
        k--;
        const double zk = y[k] / L.udiag[k];
        y[k] = zk;
        if (zk != 0)
            for (uint32_t p = L.up[k]; p < L.up[k + 1]; p++) {
                const double t = L.ux[p] * zk;
                y[L.ui[p]] = y[L.ui[p]] - t;
            }
    }
    for (uint32_t j = 0; j < n; j++) dx[q[j]] = y[j];
}

// Entry points for the lane kernels (one thread). lds: the lane's factor-
// slot area (lds_doubles long); its w is all zero on entry and on return.

// The lane's full factor of its Jacobian (PrivLU::ap of the lane).
__device__ __forceinline__ int priv_factor_at(const PrivLU &P, uint32_t lane, const uint32_t *col_ptr,
                                              const uint32_t *row_idx, const uint32_t *q, double *lds,
                                              uint32_t lds_doubles)
{
    return priv_factor(priv_lane(P, lane, lds, lds_doubles), col_ptr, row_idx, q, 1e-3);
}

// The replay of the lane's own factors, and on a failed peel test its full
// factor: kPlOk, kPlSingular or kPlFull.
__device__ __forceinline__ int priv_update_at(const PrivLU &P, uint32_t lane, const uint32_t *col_ptr,
                                              const uint32_t *row_idx, const uint32_t *q, double *lds,
                                              uint32_t lds_doubles, bool fresh = false)
{
    // fresh: the last factor met a singular matrix and left no usable LU
    const PrivLane L = priv_lane(P, lane, lds, lds_doubles);
    if (!fresh && !priv_refactor(L, col_ptr, q, 1e-12, 1.0 / 1e-3)) return kPlOk;
    return priv_factor(L, col_ptr, row_idx, q, 1e-3);
}

// dx = -A^-1 rhs in unknown order; w serves as the solve's scratch and is
// cleared again.
__device__ __forceinline__ void priv_solve_at(const PrivLU &P, uint32_t lane, const uint32_t *q, const double *rhs,
                                              double *dx, double *lds, uint32_t lds_doubles)
{
    const PrivLane L = priv_lane(P, lane, lds, lds_doubles);
    priv_solve_neg(L, q, rhs, dx, L.w);
    for (uint32_t i = 0; i < L.n; i++) L.w[i] = 0.0;
}

// A launch's start for a lane already on its own LU: w cleared and the
// pattern staged (all threads); false when it does not fit.
__device__ __forceinline__ bool priv_enter(const PrivLU &P, uint32_t lane, double *lds, uint32_t lds_doubles,
                                           uint32_t tid, uint32_t wave)
{
    const PrivLane L = priv_lane(P, lane, lds, lds_doubles);
    for (uint32_t i = tid; i < L.n; i += wave) L.w[i] = 0.0;
    return priv_stage(L, tid, wave);
}

}  // namespace esl
