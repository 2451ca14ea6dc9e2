// Device code shared by the lane kernels: wavefront reductions and the
// lane's own barrier, device evaluation of one instance of a chunk into the
// staging cells st[e * 64 + t] (x and st any indexable: a pointer, or a
// lane-major view), the limiting-state seed and the row programs of the
// refactor and solves.
#pragma once
#include <stdint.h>

#include "esl_device.h"
#include "esl_internal.h"

namespace esl {

__device__ inline uint64_t dbits(double d) { return __builtin_bit_cast(uint64_t, d); }
__device__ inline double bitsd(uint64_t b) { return __builtin_bit_cast(double, b); }

// Uniform values the compiler cannot see as uniform (a lane reduction, a
// load from a lane-indexed array) pass through readfirstlane, so branches
// on them are scalar branches and they live in scalar registers.
__device__ inline uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ inline uint64_t uni(uint64_t v)
{
    return ((uint64_t)uni((uint32_t)(v >> 32)) << 32) | uni((uint32_t)v);
}
__device__ inline double uni(double v) { return __builtin_bit_cast(double, uni(__builtin_bit_cast(uint64_t, v))); }

__device__ inline uint64_t wave_max_u64(uint64_t v)
{
    for (int o = 32; o > 0; o >>= 1) {
        uint64_t w = __shfl_xor(v, o, kWave);
        v = w > v ? w : v;
    }
    return uni(v);
}

__device__ inline double wave_fmax(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return uni(v);
}

__device__ inline bool wave_any(bool p) { return __any((int)p) != 0; }

// A lane's own barrier: orders the accesses of its wavefront's threads to
// LDS and memory (waits for outstanding accesses) without a workgroup
// barrier, so wavefronts of one workgroup that run different lanes never
// wait on each other.
__device__ inline void wsync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Explicit address spaces. A pointer the compiler cannot trace to LDS or to
// a kernel argument (one read from memory, or passed through a struct) is
// accessed with flat instructions, which count on both the LDS and the
// memory counters: an LDS read then waits for every global load in flight.
// These casts keep LDS accesses on ds_* and HBM accesses on global_*.
template <typename T>
__device__ inline __attribute__((address_space(1))) T *gptr(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ inline __attribute__((address_space(3))) T *lptr(T *p)
{
    return (__attribute__((address_space(3))) T *)p;
}

// A lane-major array seen as one lane's vector: element i at p[i * s].
struct Strided {
    double *p;
    size_t s;
    __device__ double &operator[](size_t i) const { return p[i * s]; }
};

// Per-lane limiting state: MOS1 (vgs, vds, vbs, von) as [4][n_mos], then one
// junction voltage per diode, then (vbe, vbc) per BJT.
__device__ inline uint32_t lim_width(const DeckProg &dp) { return 4 * dp.n_mos + dp.n_dio + 2 * dp.n_bjt; }
__device__ inline uint32_t lim_bjt(const DeckProg &dp) { return 4 * dp.n_mos + dp.n_dio; }

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

template <typename S>
__device__ inline void mos_eval(const Mos &m, double xd, double xg, double xs, double xb, double *lim,
                                uint32_t stride, S st, uint32_t t)
{
    double out[kStageEntries];
    lim[3 * stride] = mos_eval_v(m, xd, xg, xs, xb, lim[0], lim[stride], lim[2 * stride], out);
#pragma unroll
    for (int e = 0; e < kStageEntries; e++) st[e * kWave + t] = out[e];
}

// ---- device evaluation of one chunk (one instance per thread) into the
// staging cells st[e * 64 + t]; src scales every independent source (the
// .op ladder's source stepping, vsource.va:327-329, isource.va:326-328)
template <typename X, typename S>
__device__ inline void eval_chunk(const DeckProg &dp, const Chunk &ch, const double *live, X x, double *lim,
                                  uint32_t lim_stride, S st, uint32_t tid, double t = 0.0, bool charges = false,
                                  double src = 1.0)
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
    case kC: {
        if (!charges) break;
        const uint32_t p = dp.c_t[2 * j], q = dp.c_t[2 * j + 1];
        const double cv = dp.c_bind[j] >= 0 ? live[dp.c_bind[j]] : dp.c_val[j];
        const double qq = cv * (x[p] - x[q]);
        st[6 * kWave + tid] = cv;
        st[7 * kWave + tid] = -cv;
        st[8 * kWave + tid] = -cv;
        st[9 * kWave + tid] = cv;
        st[10 * kWave + tid] = qq;
        st[11 * kWave + tid] = -qq;
        break;
    }
    case kV: {
        const uint32_t p = dp.v_t[3 * j], q = dp.v_t[3 * j + 1], br = dp.v_t[3 * j + 2];
        const double v1 = dp.v_bind[j] >= 0 ? live[dp.v_bind[j]] : dp.v_val[j];
        const double dc = ((dp.v_wkind[j] == 1 && t > dp.v_td[j]) ? dp.v_v2[j] : v1) * src;
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
        const double dc = (dp.i_bind[j] >= 0 ? live[dp.i_bind[j]] : dp.i_val[j]) * src;
        st[0 * kWave + tid] = dc;
        st[1 * kWave + tid] = -dc;
        break;
    }
    case kD: {
        const Dio d = dio_load(dp, j, live);
        const uint32_t a = dp.d_t[2 * j], k = dp.d_t[2 * j + 1];
        double out[12];
        dio_eval_v(d, x[a], x[k], lim[4 * lim_stride + j], charges, out);
        const int ne = charges ? 12 : 6;
        for (int e = 0; e < ne; e++) st[e * kWave + tid] = out[e];
        break;
    }
    case kMos: {
        const Mos m = mos_load(dp, j, live);
        const uint32_t *t4 = dp.m_t + 4 * j;
        mos_eval(m, x[t4[0]], x[t4[1]], x[t4[2]], x[t4[3]], lim + j, lim_stride, st, tid);
        break;
    }
    case kQ: {
        const Bjt b = bjt_load(dp, j, live);
        const uint32_t *t4 = dp.q_t + 4 * j;
        const double *ql = lim + lim_bjt(dp) + 2 * j;
        double out[16];
        bjt_eval_v(b, x[t4[0]], x[t4[1]], x[t4[2]], x[t4[3]], ql[0], ql[1], out);
        for (int e = 0; e < 16; e++) st[e * kWave + tid] = out[e];
        break;
    }
    default:
        break;
    }
}

// Limiting state of a start at x (no junction seeds): the previous terminal
// voltages are x's own.
// keep_von: MOS fetlim thresholds stay the von of the last evaluation (an
// accepted transient step: the predictor's limit pass runs against the
// accepted point with the von its last load stored, mos1.va:293-297).
__device__ inline void seed_lim_at(const DeckProg &dp, const double *live, const double *x, double *lim,
                                   uint32_t tid, bool keep_von = false)
{
    const uint32_t nm = dp.n_mos;
    for (uint32_t j = tid; j < nm; j += kWave) {
        const Mos m = mos_load(dp, j, live);
        const uint32_t *t4 = dp.m_t + 4 * j;
        const double xd = x[t4[0]], xgg = x[t4[1]], xs = x[t4[2]], xb = x[t4[3]];
        lim[j] = m.tp * (xgg - xs);
        lim[nm + j] = m.tp * (xd - xs);
        lim[2 * nm + j] = m.tp * (xb - xs);
        if (!keep_von) lim[3 * nm + j] = 0.0;
    }
    for (uint32_t j = tid; j < dp.n_dio; j += kWave) lim[4 * nm + j] = x[dp.d_t[2 * j]] - x[dp.d_t[2 * j + 1]];
    for (uint32_t j = tid; j < dp.n_bjt; j += kWave) {
        const uint32_t *t4 = dp.q_t + 4 * j;
        const double tp = dp.q_par[QF_TYPE * dp.n_bjt + j];
        lim[lim_bjt(dp) + 2 * j] = tp * (x[t4[1]] - x[t4[2]]);
        lim[lim_bjt(dp) + 2 * j + 1] = tp * (x[t4[1]] - x[t4[0]]);
    }
}

// One limit pass of a lane's devices at x against its limiting state (MOS1
// fetlim/limvds/pnjlim, diode and BJT pnjlim, eval.zig limitRange): the
// thread's part; true when any of its devices moved.
template <typename X>
__device__ inline bool limit_pass(const DeckProg &dp, const double *live, X x, double *lim, uint32_t tid)
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
    return limited;
}

// A lane's start (OpArgs modes 0 cold, 1 given x, 2 resume at the saved
// iterate) and, except on a resume, its limiting state.
// A cold start: x = 0 and the limiting state's cold seeds (the initial
// junction and threshold voltages).
__device__ inline void lane_cold(const DeckProg &dp, const double *live, double *x, double *lim, uint32_t tid)
{
    const uint32_t n = dp.n, nm = dp.n_mos;
    for (uint32_t i = tid; i < n; i += kWave) x[i] = 0.0;
    for (uint32_t j = tid; j < nm; j += kWave) {
        const Mos m = mos_load(dp, j, live);
        lim[j] = m.tp * m.vbit + m.gammax * sqrt(m.phit);
        lim[nm + j] = 0.0;
        lim[2 * nm + j] = -1.0;
        lim[3 * nm + j] = 0.0;
    }
    for (uint32_t j = tid; j < dp.n_dio; j += kWave) lim[4 * nm + j] = dio_load(dp, j, live).vcrit;
    for (uint32_t j = tid; j < dp.n_bjt; j += kWave) {
        lim[lim_bjt(dp) + 2 * j] = bjt_load(dp, j, live).vcrit;
        lim[lim_bjt(dp) + 2 * j + 1] = 0.0;
    }
    wsync();
}

__device__ inline void lane_init(const OpArgs &a, uint32_t lane, const double *live, double *x, double *lim,
                                 uint32_t tid)
{
    const DeckProg &dp = a.dp;
    const uint32_t n = dp.n;
    if (a.mode == 0) {
        lane_cold(dp, live, x, lim, tid);
        return;
    }
    const double *xg = a.x + (size_t)lane * n;
    for (uint32_t i = tid; i < n; i += kWave) x[i] = xg[i];
    wsync();
    if (a.mode == 2) return;
    seed_lim_at(dp, live, x, lim, tid);
    wsync();
}

// ---- row programs (refactor and solves) ----
// The peel-test state a row program accumulates: a failed growth test,
// non-finite quotients (bit patterns ORed) and the largest multiplier (bit
// pattern max).
struct LuFlags {
    bool bad;
    uint64_t nonfinite, worst;
};

// One row of a row program: every thread runs one op of the row's kind on
// M = [factor slots | solve vector], in the factor precision R. Rows run in
// order within the wavefront (LDS accesses of one wavefront complete in issue
// order), so a row sees the results of every earlier row. The peel tests are
// computed in FP64 on the values. kSolve: a solve program (R_MACZ and R_DIV
// rows), whose R_MACZ and idle rows run one branch-free path: an idle row, or
// an op whose multiplier is zero, writes its slot's own value back (a row
// never holds two ops on one slot; idle rows touch slot 0 or pads). The
// refactor's R_MAC rows keep the plain path, which measured faster there.
template <typename R, bool kSolve, typename MP>
__device__ inline void run_row(esl::uint2_t w, MP M, double growth, LuFlags &f)
{
    const uint32_t kind = __builtin_amdgcn_readfirstlane(esl::row_kind(w.y));
    const uint32_t d = w.x & 0xffffu, b = w.x >> 16, c = w.y & 0xffffu;
    if (kind == R_CHAIN) return;   // run after its block of rows (run_chain)
    if (kSolve && kind != R_DIV) {
        const R v = M[d], A = M[b], B = M[c];
        const R p = v - A * B;
        const bool keep = kind == R_NOP || B == R(0);
        M[d] = keep ? v : p;
    } else if (!kSolve && kind == R_MAC) {
        // (refactor programs hold no R_MACZ rows: checked on the host)
        const R v = M[d], A = M[b], B = M[c];
        M[d] = v - A * B;
    } else if (kind == R_DIV) {
        const R v = M[d], A = M[b];
        const R q = v / A;
        M[d] = q;
        if (kSolve) return;   // solve programs carry no peel tests (flags 0)
        const uint32_t fl = esl::row_flags(w.y);
        const bool u = fl & kFlagUnscaled;
        const uint64_t nf = dbits((double)(q - q));
        const bool check_growth = growth > 0;
        f.bad |= u && check_growth && !(fabs((double)A) >= growth * fabs((double)v));
        f.nonfinite |= ((u && !check_growth) || (fl & kFlagScaled)) ? nf : 0;
        const uint64_t lb = u ? dbits(fabs((double)q)) : 0;
        f.worst = lb > f.worst ? lb : f.worst;
    }
}

// One chain row (R_CHAIN word w): the thread's run of len terms on target
// d, from the segment's term words, with d in a register. A block's kChainBlk
// operand pairs are read together before its multiply-subtracts, and the
// words of the block after next load while a block runs (the segment is
// padded with kChainAhead blocks). Each term rounds as the R_MAC (refactor)
// or R_MACZ (solve: a zero multiplier leaves the target) row of the same
// term would; terms past len are skipped. Kept lean on registers: the
// kernels that run it sit at the register budget of two wavefronts per
// SIMD, and a spill there costs every phase of the Newton loop.
template <typename R, bool kSolve, typename MP>
__device__ void run_chain(esl::uint2_t w, const esl::uint2_t *segs, const uint32_t *cw, MP M, uint32_t tid)
{
    const uint32_t d = w.x & 0xffffu, len = w.x >> 16;
    const esl::uint2_t sg = segs[__builtin_amdgcn_readfirstlane(w.y & 0xffffu)];
    const uint32_t blocks = __builtin_amdgcn_readfirstlane(sg.y);
    const __attribute__((address_space(1))) esl::uint4_t *p =
        reinterpret_cast<const __attribute__((address_space(1))) esl::uint4_t *>(gptr(cw) + sg.x) + tid * 2;
    auto ld = [&](uint32_t i) {
        const __attribute__((address_space(1))) esl::uint4_t &v = p[i];
        return esl::uint4_t{v.x, v.y, v.z, v.w};
    };
    constexpr uint32_t S = kWave * 2;
    R acc = M[d];
    auto block = [&](const esl::uint4_t &u0, const esl::uint4_t &u1, uint32_t k) {
        const uint32_t ws[kChainBlk] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
        R A[kChainBlk], B[kChainBlk];
#pragma unroll
        for (int e = 0; e < kChainBlk; e++) {
            A[e] = M[ws[e] & 0xffffu];
            B[e] = M[ws[e] >> 16];
        }
#pragma unroll
        for (int e = 0; e < kChainBlk; e++) {
            const R t = A[e] * B[e];
            const R nv = acc - t;
            bool use = k * kChainBlk + e < len;
            if (kSolve) use = use && B[e] != R(0);
            acc = use ? nv : acc;
        }
    };
    esl::uint4_t a0 = ld(0), a1 = ld(1), b0 = ld(S), b1 = ld(S + 1);
    for (uint32_t k = 0; k < blocks; k += 2) {
        block(a0, a1, k);
        a0 = ld((k + 2) * S);
        a1 = ld((k + 2) * S + 1);
        if (k + 1 < blocks) {
            block(b0, b1, k + 1);
            b0 = ld((k + 3) * S);
            b1 = ld((k + 3) * S + 1);
        }
    }
    M[d] = acc;
}

// A broadcast chain row (kChainBcast): its chains one after another on the
// whole wavefront. Per round of 64 terms thread k loads term k, forms its
// product (zero for a skipped R_MACZ term and for padding) and writes it to
// pad slot k (the pads are free between rows); then every thread subtracts
// the 64 products from the target in term order, reading them at one
// address for all threads, eight reads ahead of the subtractions, so each
// target rounds exactly as its ops one by one. Headers and segments are read
// through the scalar cache; the term words of up to four rounds load
// together.
template <typename R, bool kSolve, typename MP>
__device__ void run_bchain(esl::uint2_t w, const esl::uint2_t *segs, const uint32_t *cw, MP M, uint32_t pad,
                           uint32_t tid)
{
    const __attribute__((address_space(4))) uint32_t *sc =
        (const __attribute__((address_space(4))) uint32_t *)cw;
    const __attribute__((address_space(4))) esl::uint2_t *ss =
        (const __attribute__((address_space(4))) esl::uint2_t *)segs;
    const uint32_t seg = __builtin_amdgcn_readfirstlane(w.y & 0xffffu);
    const uint32_t h0 = ss[seg].x, nc = ss[seg].y;
    const __attribute__((address_space(1))) uint32_t *g = gptr(cw) + tid;
    for (uint32_t j = 0; j < nc; j++) {
        const uint32_t d = sc[h0 + 4 * j], len = sc[h0 + 4 * j + 1], off = sc[h0 + 4 * j + 2];
        R acc = M[d];
        for (uint32_t r0 = 0; r0 < len; r0 += 4 * kWave) {
            uint32_t tw[4];
#pragma unroll
            for (int q = 0; q < 4; q++) tw[q] = r0 + q * kWave < len ? g[off + r0 + q * kWave] : 0u;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t at = r0 + q * kWave;
                if (at >= len) break;
                const R A = M[tw[q] & 0xffffu], B = M[tw[q] >> 16];
                const R p = A * B;
                const bool skip = at + tid >= len || (kSolve && B == R(0));
                M[pad + tid] = skip ? R(0) : p;
                const uint32_t groups = ((len - at < (uint32_t)kWave ? len - at : (uint32_t)kWave) + 7u) >> 3;
                for (uint32_t k8 = 0; k8 < groups; k8++) {
                    R v[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) v[e] = M[pad + k8 * 8 + e];
#pragma unroll
                    for (int e = 0; e < 8; e++) acc = acc - v[e];
                }
            }
        }
        M[d] = acc;
    }
}

// Runs a row program of `rows` rows, a multiple of kRing (its words
// followed by at least kRing more rows, which are only loaded), with kRing
// rows of words in flight. M
// (R elements) and the words may be address-space qualified pointers. The
// word loads are pinned in row order (scheduling barriers) and each refills
// the register its row just used, so the oldest load completes first and a
// row waits only for its own word, never for the whole ring. A chain row,
// always a block's last row, runs after the block (segs, cw: the program's
// chain segments and terms; null for programs without chain rows).
template <typename R, bool kSolve = false, typename WP, typename MP>
__device__ void run_rows(WP words, uint32_t rows, MP M, double growth, LuFlags &f, uint32_t tid, uint32_t pad,
                         const esl::uint2_t *segs = nullptr, const uint32_t *cw = nullptr)
{
    const WP w = words + tid;
    esl::uint2_t q[kRing];
#pragma unroll
    for (int k = 0; k < kRing; k++) {
        q[k] = esl::uint2_t{w[k * kWave].x, w[k * kWave].y};
        __builtin_amdgcn_sched_barrier(0);
    }
    for (uint32_t row = 0; row < rows; row += kRing) {
        const esl::uint2_t last = q[kRing - 1];
#pragma unroll
        for (int k = 0; k < kRing; k++) {
            // use the word, then refill its register (no register rotation)
            run_row<R, kSolve>(q[k], M, growth, f);
            const size_t at = (size_t)(row + kRing + k) * kWave;
            q[k] = esl::uint2_t{w[at].x, w[at].y};
            __builtin_amdgcn_sched_barrier(0);
        }
        if (segs && __builtin_amdgcn_readfirstlane(esl::row_kind(last.y)) == R_CHAIN) {
            if (__builtin_amdgcn_readfirstlane(last.y) & kChainBcast)
                run_bchain<R, kSolve>(last, segs, cw, M, pad, tid);
            else
                run_chain<R, kSolve>(last, segs, cw, M, tid);
        }
    }
}

}  // namespace esl
