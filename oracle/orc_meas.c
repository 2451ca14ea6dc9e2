/* The oracle's .meas subset over a transient's rows, one row at a time:
 * WHEN, TRIG/TARG, FIND (AT= or WHEN), AVG, MIN, MAX, PP, INTEG and RMS,
 * each as the passes of src/output/measure.zig reads a waveform (Cross
 * :1269-1372, At :1374-1393, Extremum :1396-1447, Integ :1520-1597, the
 * line of :1251-1262), fed as its Stream feeds a running transient
 * (:160-330): a card resolves once no later row can change its value (its
 * event found, or past its TO= / AT=), which early stop reads.
 * Test infrastructure: the checker of the lanes' streamed measures. */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "orc.h"

static const double NO_AT = 1e99;

static int almost_equal_ulps(double a, double b)
{
    if (a == b) return 1;
    int64_t ia, ib;
    memcpy(&ia, &a, 8);
    memcpy(&ib, &b, 8);
    ia = ia < 0 ? (int64_t)(INT64_MIN - (uint64_t)ia) : ia;
    ib = ib < 0 ? (int64_t)(INT64_MIN - (uint64_t)ib) : ib;
    const int64_t d = (int64_t)((uint64_t)ia - (uint64_t)ib);
    return (d < 0 ? -d : d) <= 100;
}

static double line(double x0, double y0, double x1, double y1, double at)
{
    const double slope = (y1 - y0) / (x1 - x0);
    const double intercept = y0 - slope * x0;
    return slope * at + intercept;
}

/* ---- Cross (measure.zig:1269-1372), one event ---- */
static void cross_init(orc_meas_cross *s, const orc_meas_clause *c)
{
    memset(s, 0, sizeof *s);
    s->td = c->td;
    s->want_rise = c->rise;
    s->want_fall = c->fall;
    s->want_cross = c->cross;
    s->last = NAN;
    s->time = NAN;
}

/* true when the scan stops (its event found, or past TO=) */
static int cross_feed(orc_meas_cross *s, const orc_meas_clause *c, double xv, double v, double v2)
{
    const int has_y2 = c->has_level_node;
    s->rows++;
    if (xv < s->td) return 0;
    if (xv < c->from) return 0;
    if (c->to != 0 && xv > c->to) return 1;
    const double level = has_y2 ? v2 : c->val;
    const double prev_level = has_y2 ? s->prev2 : c->val;
    if (s->first == 1) {
        s->cross = 0;
        s->above = !(v < level);
        if (!s->above && s->prev >= prev_level) {
            s->fall = 1;
            s->cross = 1;
        } else if (s->above && s->prev < prev_level) {
            s->rise = 1;
            s->cross = 1;
        }
    }
    if (s->first > 1) {
        if (!s->above && v >= level) {
            s->above = 1;
            s->cross += 1;
            s->rise += 1;
            if (c->fall != ORC_MEAS_LAST) s->pending = 1;
        } else if (s->above && v <= level) {
            s->above = 0;
            s->cross += 1;
            s->fall += 1;
            if (c->rise != ORC_MEAS_LAST) s->pending = 1;
        }
        const double hit = s->prev_x + (prev_level - s->prev) * (xv - s->prev_x) / (v - s->prev - level + prev_level);
        if (s->cross == s->want_cross || s->rise == s->want_rise || s->fall == s->want_fall) {
            s->time = hit;
            s->found = 1;
            return 1;
        }
        if (s->pending) {
            if (c->cross == ORC_MEAS_UNSET && c->rise == ORC_MEAS_UNSET && c->fall == ORC_MEAS_UNSET) {
                s->time = hit;
                s->found = 1;
                return 1;
            }
            if (c->cross == ORC_MEAS_LAST || c->rise == ORC_MEAS_LAST || c->fall == ORC_MEAS_LAST) s->last = hit;
            s->pending = 0;
        }
    }
    s->first += 1;
    s->prev = v;
    if (has_y2) s->prev2 = v2;
    s->prev_x = xv;
    return 0;
}

static double cross_time(const orc_meas_cross *s) { return s->found ? s->time : s->last; }

/* ---- At (measure.zig:1374-1393) ---- */
static int at_feed(orc_meas_at *s, double at, double xv, double v)
{
    if (s->rows > 0 && s->px <= at && xv >= at) {
        s->value = s->pv + (at - s->px) * (v - s->pv) / (xv - s->px);
        return 1;
    }
    s->rows++;
    s->px = xv;
    s->pv = v;
    return xv > at;   /* a transient's times rise: past AT=, no later segment holds it */
}

/* ---- Extremum (measure.zig:1396-1447): op 0 min, 1 max, 2 avg ---- */
static int ext_feed(orc_meas_ext *s, const orc_meas_clause *c, int op, double xv, double v)
{
    s->xv = xv;
    if (xv < c->from) return 0;
    if (c->to != 0 && xv > c->to) return 1;
    if (!s->first) {
        s->first = 1;
        s->m = op == 2 ? 0.0 : v;
        s->m_at = xv;
        s->pv = v;
        s->px = xv;
        return 0;
    }
    if (op == 0) {
        if (v <= s->m) { s->m = v; s->m_at = xv; }
    } else if (op == 1) {
        if (v >= s->m) { s->m = v; s->m_at = xv; }
    } else {
        s->m += 0.5 * (v + s->pv) * (xv - s->px);
        s->span += xv - s->px;
        s->pv = v;
        s->px = xv;
    }
    return 0;
}

static double ext_value(const orc_meas_ext *s, int op)
{
    return op == 2 ? s->m / (s->first ? s->span : 1.0) : s->m;
}

/* ---- Integ (measure.zig:1520-1597) ---- */
static void integ_panel(orc_meas_integ *s, uint32_t avail)
{
    double p[4][2];
    memcpy(p, s->p, sizeof p);
    for (uint32_t k = avail; k < 4; k++) p[k][0] = p[k][1] = 0.0;
    double width[3];
    for (uint32_t j = 0; j < 3; j++) width[j] = j + 1 < avail ? p[j + 1][0] - p[j][0] : 0.0;
    uint32_t used;
    if (almost_equal_ulps(width[0], width[1]) && almost_equal_ulps(width[0], width[2])) {
        s->sum[0] += 3 * width[0] * (p[0][1] + 3 * (p[1][1] + p[2][1]) + p[3][1]) / 8.0;
        used = 3;
    } else if (almost_equal_ulps(width[0], width[1])) {
        s->sum[1] += width[0] * (p[0][1] + 4 * p[1][1] + p[2][1]) / 3.0;
        used = 2;
    } else {
        s->sum[2] += width[0] * (p[0][1] + p[1][1]) / 2;
        used = 1;
    }
    s->np = avail > used ? avail - used : 0;
    memmove(s->p, s->p + (avail - s->np), (size_t)s->np * sizeof s->p[0]);
}

static int integ_feed(orc_meas_integ *s, const orc_meas_clause *c, int rms, double xv, double v)
{
    int ret = 0;
    if (!s->inside && xv < c->from) goto done;
    s->inside = 1;
    s->cut = c->to != 0 && xv > c->to;
    double pt[2] = {xv, v};
    if (s->cut) {
        if (s->rows > 0 && !almost_equal_ulps(xv, c->to)) {
            pt[0] = c->to;
            pt[1] = line(s->px, s->py, xv, v, c->to);
        }
    } else if (s->count == 0 && c->from != 0 && s->rows > 0 && !almost_equal_ulps(xv, c->from)) {
        pt[0] = c->from;
        pt[1] = line(s->px, s->py, xv, v, c->from);
    }
    if (rms) pt[1] = pt[1] * pt[1];
    if (s->count == 0) s->first_x = pt[0];
    if (s->cut) s->cut_x = pt[0];
    s->count++;
    s->p[s->np][0] = pt[0];
    s->p[s->np][1] = pt[1];
    s->np++;
    if (s->np == 4) integ_panel(s, s->np);
    ret = s->cut;
done:
    s->rows++;
    s->px = xv;
    s->py = v;
    return ret;
}

static double integ_value(orc_meas_integ *s, int rms)
{
    if (s->rows == 0) return NAN;
    while (s->np > 1) integ_panel(s, s->np);
    const double from = s->count > (uint32_t)(s->cut ? 1 : 0) ? s->first_x : NAN;
    const double to = s->cut ? s->cut_x : s->px;
    const double total = s->sum[0] + s->sum[1] + s->sum[2];
    return rms ? sqrt(total / (to - from)) : total;
}

/* ---- cards ---- */
void orc_meas_init(const orc_meas_card *c, orc_meas_state *s)
{
    memset(s, 0, sizeof *s);
    cross_init(&s->cr[0], &c->a);
    cross_init(&s->cr[1], &c->b);
    s->at.value = NAN;
}

/* one row: t and the unknowns x; 1 once the card is resolved */
int orc_meas_feed(const orc_meas_card *c, orc_meas_state *s, double t, const double *x)
{
    if (s->resolved) return 1;
    const double ya = x[c->a.node], yb = x[c->b.node];
    const double la = c->a.has_level_node ? x[c->a.level_node] : NAN;
    const double lb = c->b.has_level_node ? x[c->b.level_node] : NAN;
    switch (c->func) {
    case ORC_MEAS_WHEN:
        if (!s->over[0]) s->over[0] = cross_feed(&s->cr[0], &c->a, t, ya, la);
        s->resolved = s->over[0];
        break;
    case ORC_MEAS_TRIG_TARG:
        if (c->a.at != NO_AT) s->over[0] = 1;
        else if (!s->over[0]) s->over[0] = cross_feed(&s->cr[0], &c->a, t, ya, la);
        if (c->b.at != NO_AT) s->over[1] = 1;
        else if (!s->over[1]) s->over[1] = cross_feed(&s->cr[1], &c->b, t, yb, lb);
        s->resolved = s->over[0] && s->over[1];
        break;
    case ORC_MEAS_FIND:
        if (c->a.at != NO_AT) {
            if (!s->over[0]) s->over[0] = at_feed(&s->at, c->a.at, t, ya);
            s->resolved = s->over[0];
            break;
        }
        /* FIND a WHEN b: the value of a at b's event, its pass starting at
         * the event from the last three rows (measure.zig:214-224) */
        for (int k = 0; k < 2; k++) {
            s->ring[k][0] = s->ring[k + 1][0];
            s->ring[k][1] = s->ring[k + 1][1];
        }
        s->ring[2][0] = t;
        s->ring[2][1] = ya;
        if (s->ring_len < 3) s->ring_len++;
        if (!s->over[1]) {
            s->over[1] = cross_feed(&s->cr[1], &c->b, t, yb, lb);
            if (s->over[1]) {
                const double at = cross_time(&s->cr[1]);
                s->find_at = at;
                if (isnan(at)) {
                    s->over[0] = 1;
                } else {
                    for (uint32_t k = 3 - s->ring_len; k < 3 && !s->over[0]; k++)
                        s->over[0] = at_feed(&s->at, at, s->ring[k][0], s->ring[k][1]);
                }
            }
        } else if (!s->over[0]) {
            s->over[0] = at_feed(&s->at, s->find_at, t, ya);
        }
        s->resolved = s->over[0] && s->over[1];
        break;
    case ORC_MEAS_AVG:
    case ORC_MEAS_MIN:
    case ORC_MEAS_MAX:
        if (!s->over[0])
            s->over[0] = ext_feed(&s->ext[0], &c->a, c->func == ORC_MEAS_MIN ? 0 : c->func == ORC_MEAS_MAX ? 1 : 2, t, ya);
        s->resolved = s->over[0];
        break;
    case ORC_MEAS_PP:
        if (!s->over[0]) s->over[0] = ext_feed(&s->ext[0], &c->a, 0, t, ya);
        if (!s->over[1]) s->over[1] = ext_feed(&s->ext[1], &c->a, 1, t, ya);
        s->resolved = s->over[0] && s->over[1];
        break;
    case ORC_MEAS_INTEG:
    case ORC_MEAS_RMS:
        if (!s->over[0]) s->over[0] = integ_feed(&s->integ, &c->a, c->func == ORC_MEAS_RMS, t, ya);
        s->resolved = s->over[0];
        break;
    default:
        s->resolved = 1;
        break;
    }
    return s->resolved;
}

/* the card's value over the rows fed (NaN when its event never came) */
double orc_meas_value(const orc_meas_card *c, orc_meas_state *s)
{
    switch (c->func) {
    case ORC_MEAS_WHEN:
        return cross_time(&s->cr[0]);
    case ORC_MEAS_TRIG_TARG: {
        const double trig = c->a.at != NO_AT ? c->a.at : cross_time(&s->cr[0]);
        const double targ = c->b.at != NO_AT ? c->b.at : cross_time(&s->cr[1]);
        return targ - trig;
    }
    case ORC_MEAS_FIND:
        return s->at.value;
    case ORC_MEAS_AVG:
        return ext_value(&s->ext[0], 2);
    case ORC_MEAS_MIN:
        return ext_value(&s->ext[0], 0);
    case ORC_MEAS_MAX:
        return ext_value(&s->ext[0], 1);
    case ORC_MEAS_PP:
        return ext_value(&s->ext[1], 1) - ext_value(&s->ext[0], 0);
    case ORC_MEAS_INTEG:
        return integ_value(&s->integ, 0);
    case ORC_MEAS_RMS:
        return integ_value(&s->integ, 1);
    default:
        return NAN;
    }
}

/* every card over recorded rows t[r], x[r * stride + node]; out[k] each
 * card's value; returns the rows read before every card resolved (early)
 * or all of them */
size_t orc_meas_rows(const orc_meas_card *cards, uint32_t n_cards, const double *t, const double *x, size_t rows,
                     size_t stride, int early, double *out)
{
    orc_meas_state st[ORC_MEAS_MAX_CARDS];
    if (n_cards > ORC_MEAS_MAX_CARDS) return 0;
    for (uint32_t k = 0; k < n_cards; k++) orc_meas_init(&cards[k], &st[k]);
    size_t r = 0;
    for (; r < rows; r++) {
        int all = 1;
        for (uint32_t k = 0; k < n_cards; k++) all &= orc_meas_feed(&cards[k], &st[k], t[r], x + r * stride);
        if (early && all) {
            r++;
            break;
        }
    }
    for (uint32_t k = 0; k < n_cards; k++) out[k] = orc_meas_value(&cards[k], &st[k]);
    return r;
}
