// Device evaluation shared by the lane kernels: one instance of a chunk
// into the staging cells st[e * 64 + t]. x and st are any indexable (a
// pointer, or a lane-major view for the lane-per-thread kernels).
#pragma once
#include <stdint.h>

#include "esl_device.h"
#include "esl_internal.h"

namespace esl {

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
// staging cells st[e * 64 + t]
template <typename X, typename S>
__device__ inline void eval_chunk(const DeckProg &dp, const Chunk &ch, const double *live, X x, double *lim,
                                  uint32_t lim_stride, S st, uint32_t tid, double t = 0.0, bool charges = false)
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
        const double dc = (dp.v_wkind[j] == 1 && t > dp.v_td[j]) ? dp.v_v2[j] : v1;
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


}  // namespace esl
