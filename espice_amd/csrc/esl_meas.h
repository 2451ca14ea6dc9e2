// .meas cards measured on the device as a transient lane accepts its steps
// (src/output/measure.zig: the passes Cross :1269-1372, At :1374-1393,
// Extremum :1396-1447 and Integ :1520-1597 that `evaluate` (:575-660) reads
// for WHEN, TRIG/TARG, FIND, AVG/MIN/MAX/PP and INTEG/RMS, fed one row at a
// time as Stream (:160-330) feeds a running transient). A card resolves
// once no later row can change it (its event found, or past TO= / AT=),
// which early stop reads (CHANGELOGS.md:13-14). One thread per card; the
// state of each (lane, card) lives in HBM between rows.
#pragma once
#include <stdint.h>

#include "../../include/espice_lanes.h"

namespace esl {

constexpr double kMeasNoAt = 1e99;
constexpr int32_t kMeasUnset = -1, kMeasLast = -2;

struct MeasCross {
    double td, last, prev, prev2, prev_x, time;
    int32_t rise, fall, cross, want_rise, want_fall, want_cross, above, pending, first, found;
};
struct MeasAt {
    double px, pv, value;
    uint32_t rows, pad;
};
struct MeasExt {
    double m, m_at, xv, pv, px, span;
    int32_t first, pad;
};
struct MeasInteg {
    double px, py, p[4][2], first_x, cut_x, sum[3];
    uint32_t rows, np, count;
    int32_t inside, cut, pad;
};
struct MeasState {
    MeasCross cr[2];
    MeasAt at;
    MeasExt ext[2];
    MeasInteg integ;
    double ring[3][2], find_at;
    uint32_t ring_len, over0, over1, resolved;
};

__host__ __device__ inline bool meas_ulps_close(double a, double b)
{
    if (a == b) return true;
    int64_t ia = __builtin_bit_cast(int64_t, a), ib = __builtin_bit_cast(int64_t, b);
    if (ia < 0) ia = (int64_t)(0x8000000000000000ull - (uint64_t)ia);
    if (ib < 0) ib = (int64_t)(0x8000000000000000ull - (uint64_t)ib);
    const int64_t d = (int64_t)((uint64_t)ia - (uint64_t)ib);
    return (d < 0 ? -d : d) <= 100;
}

__device__ inline double meas_line(double x0, double y0, double x1, double y1, double at)
{
    const double slope = (y1 - y0) / (x1 - x0);
    const double intercept = y0 - slope * x0;
    return slope * at + intercept;
}

__device__ inline void meas_cross_init(MeasCross &s, const esl_meas_clause &c)
{
    s = MeasCross{};
    s.td = c.td;
    s.want_rise = c.rise;
    s.want_fall = c.fall;
    s.want_cross = c.cross;
    s.last = __builtin_nan("");
    s.time = __builtin_nan("");
}

// One row of a Cross scan for one event; true once it stops (the event, or
// past TO=).
__device__ inline bool meas_cross_feed(MeasCross &s, const esl_meas_clause &c, double xv, double v, double v2)
{
    const bool y2 = c.has_level_node != 0;
    if (xv < s.td) return false;
    if (xv < c.from) return false;
    if (c.to != 0 && xv > c.to) return true;
    const double level = y2 ? v2 : c.val;
    const double prev_level = y2 ? s.prev2 : c.val;
    if (s.first == 1) {
        s.cross = 0;
        s.above = !(v < level);
        if (!s.above && s.prev >= prev_level) {
            s.fall = 1;
            s.cross = 1;
        } else if (s.above && s.prev < prev_level) {
            s.rise = 1;
            s.cross = 1;
        }
    }
    if (s.first > 1) {
        if (!s.above && v >= level) {
            s.above = 1;
            s.cross += 1;
            s.rise += 1;
            if (c.fall != kMeasLast) s.pending = 1;
        } else if (s.above && v <= level) {
            s.above = 0;
            s.cross += 1;
            s.fall += 1;
            if (c.rise != kMeasLast) s.pending = 1;
        }
        const double hit = s.prev_x + (prev_level - s.prev) * (xv - s.prev_x) / (v - s.prev - level + prev_level);
        if (s.cross == s.want_cross || s.rise == s.want_rise || s.fall == s.want_fall) {
            s.time = hit;
            s.found = 1;
            return true;
        }
        if (s.pending) {
            if (c.cross == kMeasUnset && c.rise == kMeasUnset && c.fall == kMeasUnset) {
                s.time = hit;
                s.found = 1;
                return true;
            }
            if (c.cross == kMeasLast || c.rise == kMeasLast || c.fall == kMeasLast) s.last = hit;
            s.pending = 0;
        }
    }
    s.first += 1;
    s.prev = v;
    if (y2) s.prev2 = v2;
    s.prev_x = xv;
    return false;
}

__device__ inline double meas_cross_time(const MeasCross &s) { return s.found ? s.time : s.last; }

// At: the first segment holding `at`; past it no later one does.
__device__ inline bool meas_at_feed(MeasAt &s, double at, double xv, double v)
{
    if (s.rows > 0 && s.px <= at && xv >= at) {
        s.value = s.pv + (at - s.px) * (v - s.pv) / (xv - s.px);
        return true;
    }
    s.rows++;
    s.px = xv;
    s.pv = v;
    return xv > at;
}

// Extremum: op 0 min, 1 max, 2 the trapezoidal mean.
__device__ inline bool meas_ext_feed(MeasExt &s, const esl_meas_clause &c, int op, double xv, double v)
{
    s.xv = xv;
    if (xv < c.from) return false;
    if (c.to != 0 && xv > c.to) return true;
    if (!s.first) {
        s.first = 1;
        s.m = op == 2 ? 0.0 : v;
        s.m_at = xv;
        s.pv = v;
        s.px = xv;
        return false;
    }
    if (op == 0) {
        if (v <= s.m) {
            s.m = v;
            s.m_at = xv;
        }
    } else if (op == 1) {
        if (v >= s.m) {
            s.m = v;
            s.m_at = xv;
        }
    } else {
        s.m += 0.5 * (v + s.pv) * (xv - s.px);
        s.span += xv - s.px;
        s.pv = v;
        s.px = xv;
    }
    return false;
}

__device__ inline double meas_ext_value(const MeasExt &s, int op)
{
    return op == 2 ? s.m / (s.first ? s.span : 1.0) : s.m;
}

// Integ: composite Simpson 3/8, Simpson 1/3 and trapezoid panels as the
// widths allow, ends interpolated.
__device__ inline void meas_integ_panel(MeasInteg &s, uint32_t avail)
{
    double p[4][2];
    for (int k = 0; k < 4; k++) {
        p[k][0] = k < (int)avail ? s.p[k][0] : 0.0;
        p[k][1] = k < (int)avail ? s.p[k][1] : 0.0;
    }
    double w[3];
    for (uint32_t j = 0; j < 3; j++) w[j] = j + 1 < avail ? p[j + 1][0] - p[j][0] : 0.0;
    uint32_t used;
    if (meas_ulps_close(w[0], w[1]) && meas_ulps_close(w[0], w[2])) {
        s.sum[0] += 3 * w[0] * (p[0][1] + 3 * (p[1][1] + p[2][1]) + p[3][1]) / 8.0;
        used = 3;
    } else if (meas_ulps_close(w[0], w[1])) {
        s.sum[1] += w[0] * (p[0][1] + 4 * p[1][1] + p[2][1]) / 3.0;
        used = 2;
    } else {
        s.sum[2] += w[0] * (p[0][1] + p[1][1]) / 2;
        used = 1;
    }
    const uint32_t keep = avail > used ? avail - used : 0;
    for (uint32_t k = 0; k < keep; k++) {
        s.p[k][0] = p[avail - keep + k][0];
        s.p[k][1] = p[avail - keep + k][1];
    }
    s.np = keep;
}

__device__ inline bool meas_integ_feed(MeasInteg &s, const esl_meas_clause &c, bool rms, double xv, double v)
{
    bool ret = false;
    if (s.inside || !(xv < c.from)) {
        s.inside = 1;
        s.cut = c.to != 0 && xv > c.to;
        double pt0 = xv, pt1 = v;
        if (s.cut) {
            if (s.rows > 0 && !meas_ulps_close(xv, c.to)) {
                pt0 = c.to;
                pt1 = meas_line(s.px, s.py, xv, v, c.to);
            }
        } else if (s.count == 0 && c.from != 0 && s.rows > 0 && !meas_ulps_close(xv, c.from)) {
            pt0 = c.from;
            pt1 = meas_line(s.px, s.py, xv, v, c.from);
        }
        if (rms) pt1 = pt1 * pt1;
        if (s.count == 0) s.first_x = pt0;
        if (s.cut) s.cut_x = pt0;
        s.count++;
        s.p[s.np][0] = pt0;
        s.p[s.np][1] = pt1;
        s.np++;
        if (s.np == 4) meas_integ_panel(s, s.np);
        ret = s.cut != 0;
    }
    s.rows++;
    s.px = xv;
    s.py = v;
    return ret;
}

__device__ inline double meas_integ_value(MeasInteg &s, bool rms)
{
    if (s.rows == 0) return __builtin_nan("");
    while (s.np > 1) meas_integ_panel(s, s.np);
    const double from = s.count > (uint32_t)(s.cut ? 1 : 0) ? s.first_x : __builtin_nan("");
    const double to = s.cut ? s.cut_x : s.px;
    const double total = s.sum[0] + s.sum[1] + s.sum[2];
    return rms ? sqrt(total / (to - from)) : total;
}

__device__ inline void meas_init(const esl_meas_card &c, MeasState &s)
{
    s = MeasState{};
    meas_cross_init(s.cr[0], c.a);
    meas_cross_init(s.cr[1], c.b);
    s.at.value = __builtin_nan("");
}

// One row (t, the unknowns x); true once the card is resolved.
template <typename X>
__device__ inline bool meas_feed(const esl_meas_card &c, MeasState &s, double t, X x)
{
    if (s.resolved) return true;
    const double ya = x[c.a.node], yb = x[c.b.node];
    const double la = c.a.has_level_node ? x[c.a.level_node] : __builtin_nan("");
    const double lb = c.b.has_level_node ? x[c.b.level_node] : __builtin_nan("");
    switch (c.func) {
    case ESL_MEAS_WHEN:
        if (!s.over0) s.over0 = meas_cross_feed(s.cr[0], c.a, t, ya, la);
        s.resolved = s.over0;
        break;
    case ESL_MEAS_TRIG_TARG:
        if (c.a.at != kMeasNoAt) s.over0 = 1;
        else if (!s.over0) s.over0 = meas_cross_feed(s.cr[0], c.a, t, ya, la);
        if (c.b.at != kMeasNoAt) s.over1 = 1;
        else if (!s.over1) s.over1 = meas_cross_feed(s.cr[1], c.b, t, yb, lb);
        s.resolved = s.over0 && s.over1;
        break;
    case ESL_MEAS_FIND:
        if (c.a.at != kMeasNoAt) {
            if (!s.over0) s.over0 = meas_at_feed(s.at, c.a.at, t, ya);
            s.resolved = s.over0;
            break;
        }
        // FIND a WHEN b: the value pass starts at b's event from the last
        // three rows (Stream's ring, measure.zig:214-224)
        s.ring[0][0] = s.ring[1][0];
        s.ring[0][1] = s.ring[1][1];
        s.ring[1][0] = s.ring[2][0];
        s.ring[1][1] = s.ring[2][1];
        s.ring[2][0] = t;
        s.ring[2][1] = ya;
        if (s.ring_len < 3) s.ring_len++;
        if (!s.over1) {
            s.over1 = meas_cross_feed(s.cr[1], c.b, t, yb, lb);
            if (s.over1) {
                const double at = meas_cross_time(s.cr[1]);
                s.find_at = at;
                if (at != at) {
                    s.over0 = 1;
                } else {
                    for (uint32_t k = 3 - s.ring_len; k < 3 && !s.over0; k++)
                        s.over0 = meas_at_feed(s.at, at, s.ring[k][0], s.ring[k][1]);
                }
            }
        } else if (!s.over0) {
            s.over0 = meas_at_feed(s.at, s.find_at, t, ya);
        }
        s.resolved = s.over0 && s.over1;
        break;
    case ESL_MEAS_AVG:
    case ESL_MEAS_MIN:
    case ESL_MEAS_MAX:
        if (!s.over0)
            s.over0 = meas_ext_feed(s.ext[0], c.a, c.func == ESL_MEAS_MIN ? 0 : c.func == ESL_MEAS_MAX ? 1 : 2, t, ya);
        s.resolved = s.over0;
        break;
    case ESL_MEAS_PP:
        if (!s.over0) s.over0 = meas_ext_feed(s.ext[0], c.a, 0, t, ya);
        if (!s.over1) s.over1 = meas_ext_feed(s.ext[1], c.a, 1, t, ya);
        s.resolved = s.over0 && s.over1;
        break;
    case ESL_MEAS_INTEG:
    case ESL_MEAS_RMS:
        if (!s.over0) s.over0 = meas_integ_feed(s.integ, c.a, c.func == ESL_MEAS_RMS, t, ya);
        s.resolved = s.over0;
        break;
    default:
        s.resolved = 1;
        break;
    }
    return s.resolved != 0;
}

__device__ inline double meas_value(const esl_meas_card &c, MeasState &s)
{
    switch (c.func) {
    case ESL_MEAS_WHEN:
        return meas_cross_time(s.cr[0]);
    case ESL_MEAS_TRIG_TARG: {
        const double trig = c.a.at != kMeasNoAt ? c.a.at : meas_cross_time(s.cr[0]);
        const double targ = c.b.at != kMeasNoAt ? c.b.at : meas_cross_time(s.cr[1]);
        return targ - trig;
    }
    case ESL_MEAS_FIND:
        return s.at.value;
    case ESL_MEAS_AVG:
        return meas_ext_value(s.ext[0], 2);
    case ESL_MEAS_MIN:
        return meas_ext_value(s.ext[0], 0);
    case ESL_MEAS_MAX:
        return meas_ext_value(s.ext[0], 1);
    case ESL_MEAS_PP:
        return meas_ext_value(s.ext[1], 1) - meas_ext_value(s.ext[0], 0);
    case ESL_MEAS_INTEG:
        return meas_integ_value(s.integ, false);
    case ESL_MEAS_RMS:
        return meas_integ_value(s.integ, true);
    default:
        return __builtin_nan("");
    }
}

}  // namespace esl
