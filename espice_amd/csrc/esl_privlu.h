// A lane's own LU on the device: the re-pivot without a trip to the host.
//
// When a lane's replay of its shared tape fails the peel tests, the lane
// factors its Jacobian itself with threshold partial pivoting (left-looking
// Gilbert-Peierls, sparse_lu.zig:260-622; the host's gp_factor and the
// oracle's orc_lu_factor state the same steps), keeps the factors and pivot
// sequence in its own slice of memory (the pattern in LDS, the values in LDS
// where they fit), and from then on replays them column by column
// (lane_kernels.zig:49-118 with the wide-multiplier test), factoring again
// whenever the replay fails. The factor's reach runs on one thread; its
// numeric steps, the replay and the solves spread each column's terms over
// the wavefront. Every slot takes the oracle's operations in the oracle's
// order, so the lane's numbers are bitwise those of the host re-pivot path
// (a host factor, then replays of that tape).
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
    uint32_t *cr;                       // LDS, values resident: column records [n + 1][4]
    uint32_t *ue;                       // LDS, values resident: U-entry records [nnz_u][2]
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
// present but hold only zeros, and c has no pivot yet (pinv: the factor's
// row pivots so far).
__device__ inline bool pl_void_unknown_p(const PrivLane &L, const uint32_t *pinv, const uint32_t *col_ptr,
                                         const uint32_t *row_idx, const double *vals, uint32_t c)
{
    if (pinv[c] != kPlNone) return false;
    for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++)
        if (vals[p] != 0) return false;
    for (uint32_t j = 0; j < L.n; j++) {
        if (j == c) continue;
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++)
            if (row_idx[p] == c && vals[p] != 0) return false;
    }
    return true;
}

// The full factor of L.ap with threshold `pivot_tol` (sparse_lu.zig:260-622,
// the steps the host's gp_factor and the oracle's orc_lu_factor state),
// spread over the wavefront:
// the reach of each column through the graph of L runs on thread 0, the
// triangular solve takes the reached columns in reverse finish order with
// each column's L terms on different threads, the pivot search reduces over
// the wavefront (largest magnitude, the first in finish order on ties, as
// the serial scan keeps it), and the new L column is written in finish
// order. ws: 5n words of LDS for the reach (pinv, marks, stacks, finish
// order), or null to use the lane's global arrays. kPlFull: the factors
// outgrew the lane's capacity (w is left all zero).
__device__ int priv_factor_w(const PrivLane &L, const uint32_t *col_ptr, const uint32_t *row_idx, const uint32_t *q,
                             double pivot_tol, uint32_t *ws, uint32_t tid)
{
    const uint32_t n = L.n, nnz = L.nnz;
    const double *vals = L.ap;
    const auto w = lptr(L.w);
    uint32_t *pinv = ws ? ws : L.pinv;
    uint32_t *flag = ws ? ws + n : L.flag;
    uint32_t *stack = ws ? ws + 2 * n : L.stack;
    uint32_t *pstack = ws ? ws + 3 * n : L.pstack;
    uint32_t *topo = ws ? ws + 4 * n : L.topo;
    uint64_t *rsb = reinterpret_cast<uint64_t *>(L.rscale);
    const auto lp = lptr(L.lp), up = lptr(L.up), li = lptr(L.li), ui = lptr(L.ui);
    for (uint32_t i = tid; i < n; i += kWave) {
        pinv[i] = kPlNone;
        L.flags[i] = 0;
        flag[i] = 0;
        rsb[i] = 0;
        L.want[i] = i;
        L.want_of[i] = i;
    }
    uint64_t amaxb = 0;
    for (uint32_t p = tid; p < nnz; p += kWave) {
        const uint64_t bb = pl_bits(fabs(vals[p]));
        amaxb = bb > amaxb ? bb : amaxb;
    }
    amaxb = wave_max_u64(amaxb);
    if (!(pl_double(amaxb) <= 1.7976931348623157e308)) return kPlSingular;
    wsync();
    // row scales: every entry finite here, so the largest magnitude is the
    // largest bit pattern
    for (uint32_t p = tid; p < nnz; p += kWave)
        atomicMax(reinterpret_cast<unsigned long long *>(rsb + row_idx[p]), (unsigned long long)pl_bits(fabs(vals[p])));
    wsync();
    for (uint32_t i = tid; i < n; i += kWave) {
        const double r = pl_double(rsb[i]);
        L.rscale[i] = (r > 0 && r < __builtin_inf()) ? 1.0 / r : 1.0;
    }
    wsync();
    uint32_t nnz_l = 0, nnz_u = 0;
    bool has_void = false;
    uint64_t nonfinite = 0;
    const uint32_t room = L.cap < L.lcap ? L.cap : L.lcap;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t c = uni(q[k]);
        const uint32_t mark = k + 1;
        uint32_t nt = 0;
        if (tid == 0) {
            lp[k] = nnz_l;
            up[k] = nnz_u;
            // reach of A[:, c] through the graph of L (rows still original)
            for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) {
                uint32_t r = row_idx[p];
                if (flag[r] == mark) continue;
                uint32_t sp = 0;
                stack[0] = r;
                pstack[0] = pinv[r] == kPlNone ? 0 : lp[pinv[r]];
                flag[r] = mark;
                for (;;) {
                    r = stack[sp];
                    const uint32_t kc = pinv[r];
                    const uint32_t end = kc == kPlNone ? 0 : lp[kc + 1];
                    uint32_t pos = pstack[sp];
                    bool descended = false;
                    while (pos < end) {
                        const uint32_t child = li[pos];
                        pos++;
                        if (flag[child] != mark) {
                            pstack[sp] = pos;
                            flag[child] = mark;
                            sp++;
                            stack[sp] = child;
                            pstack[sp] = pinv[child] == kPlNone ? 0 : lp[pinv[child]];
                            descended = true;
                            break;
                        }
                    }
                    if (descended) continue;
                    topo[nt++] = r;
                    if (sp == 0) break;
                    sp--;
                }
            }
        }
        nt = uni(nt);
        wsync();
        if (nnz_u + nt > room || nnz_l + nt > room) {
            for (uint32_t i = tid; i < n; i += kWave) w[i] = 0.0;
            return kPlFull;
        }
        {
            const uint32_t p0 = uni(col_ptr[c]), p1 = uni(col_ptr[c + 1]);
            for (uint32_t p = p0 + tid; p < p1; p += kWave) w[row_idx[p]] = vals[p];
        }
        // triangular solve in reverse finish order
        for (uint32_t idx = nt; idx > 0;) {
            idx--;
            const uint32_t r = uni(topo[idx]), kc = uni(pinv[r]);
            if (kc == kPlNone) continue;
            const double ukr = uni(w[r]);
            if (tid == 0) {
                ui[nnz_u] = kc;
                L.ux[nnz_u] = ukr;
            }
            nnz_u++;
            const uint32_t l0 = uni(lp[kc]), l1 = uni(lp[kc + 1]);
            for (uint32_t pl = l0 + tid; pl < l1; pl += kWave) {
                const uint32_t t_r = li[pl];
                const double t = L.lx[pl] * ukr;
                w[t_r] = w[t_r] - t;
            }
        }
        // pivot search: the largest magnitude, the first in finish order on
        // ties (a strictly larger one replaces it in the serial scan)
        double amax = 0, smax = 0;
        uint32_t best_s = kPlNone;
        for (uint32_t s0 = tid; s0 < nt; s0 += kWave) {
            const uint32_t r = topo[s0];
            if (pinv[r] != kPlNone) continue;
            const double a = fabs(w[r]);
            if (a > amax) {
                amax = a;
                best_s = s0;
            }
            smax = fmax(smax, a * L.rscale[r]);
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double oa = __shfl_xor(amax, o, kWave);
            const uint32_t os = (uint32_t)__shfl_xor((int)best_s, o, kWave);
            if (oa > amax || (oa == amax && os < best_s)) {
                amax = oa;
                best_s = os;
            }
            smax = fmax(smax, __shfl_xor(smax, o, kWave));
        }
        amax = uni(amax);
        smax = uni(smax);
        best_s = uni(best_s);
        uint32_t piv = (best_s == kPlNone || !(amax > 0)) ? kPlNone : uni(topo[best_s]);
        if (piv == kPlNone || amax == 0 || !(amax < __builtin_inf())) {
            int is_void = 0;
            if (tid == 0) is_void = pl_void_unknown_p(L, pinv, col_ptr, row_idx, vals, c) ? 1 : 0;
            is_void = (int)uni((uint32_t)is_void);
            for (uint32_t s0 = tid; s0 < nt; s0 += kWave) w[topo[s0]] = 0;
            if (!is_void) return kPlSingular;
            has_void = true;
            if (tid == 0) {
                L.flags[k] |= kVoidBit;
                L.udiag[k] = 1;
                pinv[c] = k;
            }
            wsync();
            continue;
        }
        if (tid == 0) {
            const uint32_t pr = L.want[c];
            if (pinv[pr] == kPlNone) {
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
            L.udiag[k] = w[piv];
            pinv[piv] = k;
        }
        piv = uni(piv);
        const double d = uni(w[piv]);
        wsync();
        // the L column in finish order
        for (uint32_t b = 0; b < nt; b += kWave) {
            const uint32_t s0 = b + tid;
            bool in_l = false;
            double v = 0;
            uint32_t r = 0;
            if (s0 < nt) {
                r = topo[s0];
                v = w[r];
                nonfinite |= pl_bits(v - v);
                in_l = pinv[r] == kPlNone;
            }
            const uint64_t mask = __ballot(in_l);
            const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << tid) - 1ull));
            if (in_l) {
                const double l = v / d;
                nonfinite |= pl_bits(l - l);
                li[nnz_l + rank] = r;
                L.lx[nnz_l + rank] = l;
            }
            if (s0 < nt) w[r] = 0;
            nnz_l += (uint32_t)__popcll(mask);
        }
        wsync();
    }
    if (tid == 0) {
        lp[n] = nnz_l;
        up[n] = nnz_u;
    }
    if (wave_any(nonfinite != 0)) return kPlSingular;
    wsync();
    for (uint32_t p = tid; p < nnz_l; p += kWave) li[p] = pinv[li[p]];
    for (uint32_t p = tid; p < nnz; p += kWave) L.prow[p] = pinv[row_idx[p]];
    if (ws)
        for (uint32_t i = tid; i < n; i += kWave) L.pinv[i] = pinv[i];
    wsync();
    uint32_t n_void = 0;
    if (has_void && tid == 0)
        for (uint32_t j = 0; j < n; j++) {
            const uint32_t cj = q[j];
            for (uint32_t p = col_ptr[cj]; p < col_ptr[cj + 1]; p++)
                if ((L.flags[j] & kVoidBit) || (L.flags[L.prow[p]] & kVoidBit)) L.void_slots[n_void++] = p;
        }
    if (tid == 0) {
        L.hdr[0] = nnz_l;
        L.hdr[1] = nnz_u;
        L.hdr[2] = n_void;
    }
    // the pattern's global copy, for the lane's later launches
    for (uint32_t k = tid; k <= n; k += kWave) {
        L.glp[k] = lp[k];
        L.gup[k] = up[k];
    }
    for (uint32_t p = tid; p < nnz_l; p += kWave) L.gli[p] = li[p];
    for (uint32_t p = tid; p < nnz_u; p += kWave) L.gui[p] = ui[p];
    wsync();
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
// lane's working set (`tail`, tail_n doubles) when they fit there with two
// record arrays the replay reads ahead of its chain: per column k (and a
// sentinel at n) {up[k], lp[k], A's column span p0, p1 | flags << 30}, and
// per U entry {i | (L column i's length) << 16, lp[i]}; otherwise in
// global memory. Uniform.
__device__ inline bool priv_resident(PrivLane &L, double *tail, uint32_t tail_n)
{
    const uint32_t nl = uni(L.hdr[0]), nu = uni(L.hdr[1]), n = L.n;
    if (n >= 65536 || L.nnz >= (1u << 30)) return false;
    uint64_t o = (uint64_t)nl + nu + n;
    if (reinterpret_cast<uintptr_t>(tail + o) & 15) o++;
    if (o + 2 * ((uint64_t)n + 1) + nu > tail_n) return false;
    L.lx = tail;
    L.ux = tail + nl;
    L.udiag = tail + nl + nu;
    L.cr = reinterpret_cast<uint32_t *>(tail + o);
    L.ue = reinterpret_cast<uint32_t *>(tail + o + 2 * ((uint64_t)n + 1));
    return true;
}

// The record arrays of a resident lane (all threads; the pattern is staged).
__device__ inline void priv_stage_static(const PrivLane &L, const uint32_t *col_ptr, const uint32_t *q, uint32_t tid)
{
    const uint32_t n = L.n, nu = uni(L.hdr[1]);
    for (uint32_t k = tid; k <= n; k += kWave) {
        uint32_t *r = L.cr + 4 * (size_t)k;
        r[0] = L.up[k];
        r[1] = L.lp[k];
        if (k < n) {
            const uint32_t c = q[k];
            r[2] = col_ptr[c];
            r[3] = col_ptr[c + 1] | (L.flags[k] << 30);
        } else {
            r[2] = r[3] = 0;
        }
    }
    for (uint32_t p = tid; p < nu; p += kWave) {
        const uint32_t i = L.ui[p], a0 = L.lp[i];
        L.ue[2 * (size_t)p] = i | ((L.lp[i + 1] - a0) << 16);
        L.ue[2 * (size_t)p + 1] = a0;
    }
}

// The replay of a resident lane (priv_replay's steps in its order): the
// records of the next column, its U entries and its first L rows are read
// while this column runs, A's terms two columns ahead, and each U entry's first
// terms while the entry before it runs, so an entry costs one LDS round
// trip (its pivot-row value and its targets together). True when a peel
// test fails.
__device__ bool priv_replay_res(const PrivLane &L, double growth_limit, double wide, uint32_t tid)
{
    const uint32_t n = L.n, nl = uni(L.hdr[0]), nu = uni(L.hdr[1]);
    const auto vals = gptr(L.ap);
    const auto prow = gptr(L.prow);
    const auto w = lptr(L.w);
    const auto li = lptr(L.li);
    const auto lx = lptr(L.lx), ux = lptr(L.ux), ud = lptr(L.udiag);
    const auto cr = lptr(L.cr);
    const auto ue = lptr(L.ue);
    const uint32_t n_void = uni(L.hdr[2]);
    bool nz = false;
    for (uint32_t v = tid; v < n_void; v += kWave) nz = nz || vals[gptr(L.void_slots)[v]] != 0;
    if (wave_any(nz)) return true;
    uint64_t nonfinite = 0, worst = 0;
    bool bad = false, growth_bad = false;
    struct Rec {
        uint32_t x, y, z, w;
    };
    struct Ent {
        uint32_t x, y;
    };
    auto rec = [&](uint32_t k) { return Rec{cr[4 * k], cr[4 * k + 1], cr[4 * k + 2], cr[4 * k + 3]}; };
    auto ent = [&](uint32_t p) { return Ent{ue[2 * p], ue[2 * p + 1]}; };
    // column 0 and 1 records, column 0's U entries, first L rows and terms of A
    Rec c = rec(0), cn = rec(n >= 1 ? 1 : 0);
    Ent e{0, 0};
    if (c.x + tid < nu) e = ent(c.x + tid);
    uint32_t kr = c.y + tid < nl ? li[c.y + tid] : 0u;
    // this thread's first term of A in columns k and k + 1 (read two columns
    // ahead, so a column's global read has two columns' time to arrive)
    double sv = 0, sv1 = 0;
    uint32_t sr = 0, sr1 = 0;
    {
        const uint32_t p0 = uni(c.z), p1 = uni(c.w) & 0x3fffffffu;
        if (p0 + tid < p1) {
            sv = vals[p0 + tid];
            sr = prow[p0 + tid];
        }
        if (n >= 2) {
            const uint32_t q0 = uni(cn.z), q1 = uni(cn.w) & 0x3fffffffu;
            if (q0 + tid < q1) {
                sv1 = vals[q0 + tid];
                sr1 = prow[q0 + tid];
            }
        }
    }
    // entry j of the column: (i, L column start, length)
    auto entry = [&](uint32_t u0, uint32_t j, uint32_t &i, uint32_t &a0, uint32_t &len) {
        uint32_t x, y;
        if (j < kWave) {
            x = (uint32_t)__builtin_amdgcn_readlane((int)e.x, (int)j);
            y = (uint32_t)__builtin_amdgcn_readlane((int)e.y, (int)j);
        } else {
            const Ent r = ent(u0 + j);
            x = uni(r.x);
            y = uni(r.y);
        }
        i = x & 0xffffu;
        len = x >> 16;
        a0 = y;
    };
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t u0 = uni(c.x), l0 = uni(c.y), p0 = uni(c.z), pw = uni(c.w);
        const uint32_t p1 = pw & 0x3fffffffu, fk = pw >> 30;
        const uint32_t u1 = uni(cn.x), l1 = uni(cn.y);
        const uint32_t m = u1 - u0;
        // ahead: the record after next, the next column's entries, L rows and terms of A
        const Rec cnn = k + 2 <= n ? rec(k + 2) : cn;
        Ent en{0, 0};
        if (u1 + tid < nu) en = ent(u1 + tid);
        const uint32_t krn = l1 + tid < nl ? li[l1 + tid] : 0u;
        double sv2 = 0;
        uint32_t sr2 = 0;
        if (k + 2 < n) {
            const uint32_t q0 = uni(cnn.z), q1 = uni(cnn.w) & 0x3fffffffu;
            if (q0 + tid < q1) {
                sv2 = vals[q0 + tid];
                sr2 = prow[q0 + tid];
            }
        }
        // this column's terms of A
        if (p0 + tid < p1) w[sr] = sv;
        for (uint32_t p = p0 + kWave + tid; p < p1; p += kWave) w[prow[p]] = vals[p];
        // the U entries in order, each entry's first terms read one entry ahead
        uint32_t r0 = 0;
        double m0 = 0;
        if (m > 0) {
            uint32_t i, a0, len;
            entry(u0, 0, i, a0, len);
            if (tid < len) {
                r0 = li[a0 + tid];
                m0 = lx[a0 + tid];
            }
        }
        for (uint32_t j = 0; j < m; j++) {
            uint32_t i, a0, len;
            entry(u0, j, i, a0, len);
            uint32_t r1 = 0;
            double m1 = 0;
            if (j + 1 < m) {
                uint32_t i1, b0, len1;
                entry(u0, j + 1, i1, b0, len1);
                if (tid < len1) {
                    r1 = li[b0 + tid];
                    m1 = lx[b0 + tid];
                }
            }
            const bool on = tid < len;
            const double wr = on ? w[r0] : 0.0;
            const double uki = uni(w[i]);
            nonfinite |= dbits(uki - uki);
            if (on) {
                const double t = m0 * uki;
                w[r0] = wr - t;
            }
            for (uint32_t pl = a0 + kWave + tid; pl < a0 + len; pl += kWave) {
                const uint32_t r = li[pl];
                const double t = lx[pl] * uki;
                w[r] = w[r] - t;
            }
            // row i is no target of its own L column
            if (tid == 0) {
                w[i] = 0;
                ux[u0 + j] = uki;
            }
            r0 = r1;
            m0 = m1;
        }
        const bool own = l0 + tid < l1;
        if (fk & kVoidBit) {
            if (tid == 0) {
                ud[k] = 1;
                for (uint32_t p = l0; p < l1; p++) {
                    lx[p] = 0;
                    w[li[p]] = 0;
                }
                w[k] = 0;
                for (uint32_t j = 0; j < m; j++) {
                    const Ent r = ent(u0 + j);
                    const uint32_t a0 = r.y, a1 = r.y + (r.x >> 16);
                    for (uint32_t pl = a0; pl < a1; pl++) w[li[pl]] = 0;
                }
            }
        } else {
            const double v0 = own ? w[kr] : 0.0;
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
                // each thread tests growth on its own terms (|d| >= g m holds
                // for the largest m exactly when it holds for every term's)
                const uint64_t db = dbits(fabs(d));
                uint64_t cb = db;
                if (own) {
                    w[kr] = 0;
                    const uint64_t vb = dbits(fabs(v0));
                    cb = vb > cb ? vb : cb;
                    const double lv = v0 / d;
                    lx[l0 + tid] = lv;
                    const uint64_t lb = dbits(fabs(lv));
                    worst = lb > worst ? lb : worst;
                }
                for (uint32_t pl = l0 + kWave + tid; pl < l1; pl += kWave) {
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
                growth_bad = growth_bad || !(fabs(d) >= growth_limit * bitsd(cb));
            } else {
                if (own) {
                    const double lv = v0 / d;
                    nonfinite |= dbits(lv - lv);
                    lx[l0 + tid] = lv;
                    w[kr] = 0;
                    const uint64_t lb = dbits(fabs(lv));
                    if (!scaled) worst = lb > worst ? lb : worst;
                }
                for (uint32_t pl = l0 + kWave + tid; pl < l1; pl += kWave) {
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
        c = cn;
        cn = cnn;
        e = en;
        kr = krn;
        sv = sv1;
        sr = sr1;
        sv1 = sv2;
        sr1 = sr2;
    }
    if (wave_any(nonfinite != 0)) return true;
    worst = wave_max_u64(worst);
    return bad || wave_any(growth_bad) || !(bitsd(worst) <= wide);
}

template <bool kRes>
__device__ inline auto pl_vals(double *p)
{
    if constexpr (kRes) return lptr(p);
    else return gptr(p);
}

// The lane replay of its own factors on L.ap (lane_kernels.zig:49-118 with
// the wide-multiplier test) with the values in global memory, spread over
// the wavefront: column k's scatter, each U entry's update of the dense
// column (the targets of one L column are distinct rows, so its terms go
// to different threads while the U entries stay in their order), and the
// division of column k by its pivot run one term per thread; every slot
// takes the serial replay's operations in its order, so the factors are
// bitwise the serial replay's. LDS operations of one wavefront complete in
// issue order, which carries each step's writes to the next step's reads.
// True when a peel test fails (the lane then factors again).
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
    const auto lx = gptr(L.lx), ux = gptr(L.ux), ud = gptr(L.udiag);
    const uint32_t n_void = uni(L.hdr[2]);
    bool nz = false;
    for (uint32_t v = tid; v < n_void; v += kWave) nz = nz || vals[gptr(L.void_slots)[v]] != 0;
    if (wave_any(nz)) return true;
    uint64_t nonfinite = 0, worst = 0;
    bool bad = false, growth_bad = false;
    uint32_t u0 = uni(up[0]);
    for (uint32_t k = 0; k < n; k++) {
        uint32_t p0, p1, fk;
        // this column's U entries (thread t: the t-th, its L column's span)
        // and its own L rows, read together before any of them is needed
        const uint32_t u1 = uni(up[k + 1]);
        const uint32_t l0 = uni(lp[k]), l1 = uni(lp[k + 1]);
        const uint32_t m = u1 - u0;
        uint32_t ei = 0, ea = 0, eb = 0;
        if (tid < m) {
            ei = ui[u0 + tid];
            ea = lp[ei];
            eb = lp[ei + 1];
        }
        const bool own = l0 + tid < l1;
        const uint32_t kr = own ? li[l0 + tid] : 0u;
        {
            const uint32_t c = uni(q[k]);
            p0 = uni(col_ptr[c]);
            p1 = uni(col_ptr[c + 1]);
            for (uint32_t p = p0 + tid; p < p1; p += kWave) w[prow[p]] = vals[p];
            fk = uni(flags[k]);
        }
        for (uint32_t j = 0; j < m; j++) {
            uint32_t i, a0, a1;
            if (j < kWave) {
                i = (uint32_t)__builtin_amdgcn_readlane((int)ei, (int)j);
                a0 = (uint32_t)__builtin_amdgcn_readlane((int)ea, (int)j);
                a1 = (uint32_t)__builtin_amdgcn_readlane((int)eb, (int)j);
            } else {
                i = uni(ui[u0 + j]);
                a0 = uni(lp[i]);
                a1 = uni(lp[i + 1]);
            }
            const bool on = a0 + tid < a1;
            const uint32_t r0 = on ? li[a0 + tid] : 0u;
            const double m0 = on ? lx[a0 + tid] : 0.0;
            const double uki = uni(w[i]);
            nonfinite |= dbits(uki - uki);
            if (on) {
                const double t = m0 * uki;
                w[r0] = w[r0] - t;
            }
            for (uint32_t pl = a0 + kWave + tid; pl < a1; pl += kWave) {
                const uint32_t r = li[pl];
                const double t = lx[pl] * uki;
                w[r] = w[r] - t;
            }
            // row i is no target of its own L column
            if (tid == 0) {
                w[i] = 0;
                ux[u0 + j] = uki;
            }
        }
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
        const double v0 = own ? w[kr] : 0.0;
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
            // the growth test on the column's largest magnitude: each thread
            // tests its own terms (|d| >= g m holds for the largest m exactly
            // when it holds for every term's)
            const uint64_t db = dbits(fabs(d));
            uint64_t cb = db;
            if (own) {
                w[kr] = 0;
                const uint64_t vb = dbits(fabs(v0));
                cb = vb > cb ? vb : cb;
                const double lv = v0 / d;
                lx[l0 + tid] = lv;
                const uint64_t lb = dbits(fabs(lv));
                worst = lb > worst ? lb : worst;
            }
            for (uint32_t pl = l0 + kWave + tid; pl < l1; pl += kWave) {
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
            growth_bad = growth_bad || !(fabs(d) >= growth_limit * bitsd(cb));
        } else {
            if (own) {
                const double lv = v0 / d;
                nonfinite |= dbits(lv - lv);
                lx[l0 + tid] = lv;
                w[kr] = 0;
                const uint64_t lb = dbits(fabs(lv));
                if (!scaled) worst = lb > worst ? lb : worst;
            }
            for (uint32_t pl = l0 + kWave + tid; pl < l1; pl += kWave) {
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
    return bad || wave_any(growth_bad) || !(bitsd(worst) <= wide);
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
    // each column's first wavefront of terms is read before its pivot entry
    // is known, so the two reads overlap
    uint32_t sa = 0, sb = 0;   // thread t: lp[k0 + t], lp[k0 + t + 1] of the current 64 columns
    for (uint32_t k = 0; k < n; k++) {
        if ((k & (kWave - 1)) == 0 && k + tid < n) {
            sa = lp[k + tid];
            sb = lp[k + tid + 1];
        }
        const uint32_t l0 = (uint32_t)__builtin_amdgcn_readlane((int)sa, (int)(k & (kWave - 1)));
        const uint32_t l1 = (uint32_t)__builtin_amdgcn_readlane((int)sb, (int)(k & (kWave - 1)));
        const bool on = l0 + tid < l1;
        const uint32_t r0 = on ? li[l0 + tid] : 0u;
        const double m0 = on ? lx[l0 + tid] : 0.0;
        const double yk = uni(y[k]);
        if (yk != 0) {
            if (on) {
                const double t = m0 * yk;
                y[r0] = y[r0] - t;
            }
            for (uint32_t p = l0 + kWave + tid; p < l1; p += kWave) {
                const uint32_t r = li[p];
                const double t = lx[p] * yk;
                y[r] = y[r] - t;
            }
        }
    }
    for (uint32_t k = n; k > 0;) {
        k--;
        // thread t: up[k0 + t], up[k0 + t + 1] of the 64 columns ending at k
        const uint32_t k0 = k & ~(kWave - 1);
        if (k == n - 1 || (k & (kWave - 1)) == kWave - 1) {
            if (k0 + tid <= k) {
                sa = up[k0 + tid];
                sb = up[k0 + tid + 1];
            }
        }
        const uint32_t u0 = (uint32_t)__builtin_amdgcn_readlane((int)sa, (int)(k - k0));
        const uint32_t u1 = (uint32_t)__builtin_amdgcn_readlane((int)sb, (int)(k - k0));
        const bool on = u0 + tid < u1;
        const uint32_t r0 = on ? ui[u0 + tid] : 0u;
        const double m0 = on ? ux[u0 + tid] : 0.0;
        const double zk = uni(y[k]) / uni(ud[k]);
        if (tid == 0) y[k] = zk;
        if (zk != 0) {
            if (on) {
                const double t = m0 * zk;
                y[r0] = y[r0] - t;
            }
            for (uint32_t p = u0 + kWave + tid; p < u1; p += kWave) {
                const uint32_t r = ui[p];
                const double t = ux[p] * zk;
                y[r] = y[r] - t;
            }
        }
    }
    for (uint32_t j = tid; j < n; j += kWave) {
        dx[q[j]] = y[j];
        y[j] = 0.0;
    }
}

// A fresh factor's values into the LDS tail when they fit there (all
// threads; the factor wrote them to global memory).
__device__ inline void priv_load(const PrivLane &G, const uint32_t *col_ptr, const uint32_t *q, double *tail,
                                 uint32_t tail_n, uint32_t tid)
{
    PrivLane L = G;
    if (!priv_resident(L, tail, tail_n)) return;
    priv_stage_static(L, col_ptr, q, tid);
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
    // the reach's work arrays in the LDS tail when they fit (its values are
    // loaded after the factor)
    uint32_t *ws = 5 * (uint64_t)L.n <= 2 * (uint64_t)tail_n ? reinterpret_cast<uint32_t *>(tail) : nullptr;
    const int rc = priv_factor_w(L, col_ptr, row_idx, q, 1e-3, ws, tid);
    wsync();
    if (rc == kPlOk) priv_load(L, col_ptr, q, tail, tail_n, tid);
    return rc;
}

// The replay of the lane's own factors, and on a failed peel test its full
// factor: kPlOk, kPlSingular or kPlFull (uniform). fresh: the last factor
// met a singular matrix and left no usable LU.
// The replay alone: true when a peel test fails (uniform); resident: the
// values were in LDS.
__device__ __forceinline__ bool priv_replay_at(const PrivLU &P, uint32_t lane, const uint32_t *col_ptr,
                                               const uint32_t *q, double *lds, uint32_t lds_doubles, double *tail,
                                               uint32_t tail_n, uint32_t tid, bool *resident = nullptr)
{
    PrivLane L = priv_lane(P, lane, lds, lds_doubles);
    const bool res = priv_resident(L, tail, tail_n);
    if (resident) *resident = res;
    const bool failed = res ? priv_replay_res(L, 1e-12, 1.0 / 1e-3, tid)
                            : priv_replay(L, col_ptr, q, 1e-12, 1.0 / 1e-3, tid);
    wsync();
    return failed;
}

__device__ __forceinline__ int priv_update_at(const PrivLU &P, uint32_t lane, const uint32_t *col_ptr,
                                              const uint32_t *row_idx, const uint32_t *q, double *lds,
                                              uint32_t lds_doubles, double *tail, uint32_t tail_n, uint32_t tid,
                                              bool fresh = false)
{
    if (!fresh && !priv_replay_at(P, lane, col_ptr, q, lds, lds_doubles, tail, tail_n, tid)) return kPlOk;
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

// A launch's start for a lane already on its own LU: w cleared, the
// pattern staged, and for a resident lane its column ranges and flags (all
// threads); false when it does not fit. The values are the next replay's
// (a replay always precedes the next solve).
__device__ __forceinline__ bool priv_enter(const PrivLU &P, uint32_t lane, const uint32_t *col_ptr,
                                           const uint32_t *q, double *lds, uint32_t lds_doubles, double *tail,
                                           uint32_t tail_n, uint32_t tid)
{
    PrivLane L = priv_lane(P, lane, lds, lds_doubles);
    for (uint32_t i = tid; i < L.n; i += kWave) L.w[i] = 0.0;
    if (!priv_stage(L, tid, kWave)) return false;
    if (priv_resident(L, tail, tail_n)) priv_stage_static(L, col_ptr, q, tid);
    return true;
}

}  // namespace esl
