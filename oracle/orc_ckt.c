/* Oracle circuits, device models and Newton for operating points.
 *
 * MNA (src/analysis/Circuit.zig:89-214, 570-573): unknown 0 is ground and is
 * pinned by g[diag0] += 1, rhs[0] += x[0]; devices never stamp a ground row
 * or column. Node unknowns come first, then one branch current per voltage
 * source in stamp order. The pattern holds every entry a device can stamp
 * plus every diagonal. rhs is the KCL residual F (currents leaving a node)
 * and the branch rows' KVL residual; Newton solves dx = -J^-1 F.
 *
 * Stamp order: by kind (R, C, V, I, D, MOS1), deck order within a kind; each
 * device adds its rows and Jacobian entries in a fixed entry order
 * (src/device/eval.zig:860-880 stamps per instance, rows in order).
 *
 * Device models (restated from the models/ Verilog-A files):
 *   R   resistor.va:44-52   I(p,n) = V(p,n)/r           (tnom = temp)
 *   V   vsource.va:329      V(p,n) = dc; branch row F = (vp - vn) - dc
 *   I   isource.va:328      I(p,n) = dc
 *   MOS1 mos1.va:176-463    DC part: bias-independent prep (mos1temp), the
 *       fetlimds/limvds/pnjlimds limiting, bulk junctions jcur, mode swap,
 *       Shichman-Hodges current with sbody and lambda. Derivatives are the
 *       analytic partials of those expressions (the reference takes them by
 *       forward AD; both differentiate the same branches).
 * Limiting (src/device/eval.zig:837-841, 1005-1039; ngspice mos1load.c:
 * 361-407 as described in mos1.va:282-304): a device keeps limited branch
 * voltages between iterations; after each update it limits the new branch
 * voltages against them, snaps values within 16 eps of the probe back to it
 * (eval.zig:984-993), and reports "limited" when any value moved. The next
 * evaluation runs at the limited voltages and corrects its residual by the
 * Jacobian times (probe - limited). A cold start seeds the limited voltages
 * (eval.zig:1841-1853): vgs = type vbit + gamma sqrt(phit), vds = 0,
 * vbs = -1 (mos1.va:297-300).
 *
 * Newton (converger.zig:199-315) with finalizeStep's gate order (454-522):
 * update + delta score, limiting, first-iterate veto, delta < 1, runaway
 * bound, row-scaled residual gate; INITFIX on a cold start; a converged
 * point publishes x_old.
 *
 * Parity: the models are pinned against the ngspice-44.2 operating points in
 * tests/fixtures of the reference (transcribed into tests/golden) and by
 * finite differences of the residual; Newton by closed forms (a divider,
 * inverter rails); the lane machinery by agreement with the GPU.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define K_R 0
#define K_C 1
#define K_V 2
#define K_I 3
#define K_D 4
#define K_MOS 5
#define K_Q 6
#define NKIND 7
#define NPARAM 32
#define TRASH 0xffffffffu

#define KB 1.38064852e-23
#define QE 1.6021766208e-19

typedef struct {
    uint32_t kind, dev;           /* deck index */
    uint32_t t[4];                /* terminal unknowns */
    uint32_t br;                  /* branch unknown (V) */
    uint32_t slot[12];            /* pattern slot per entry, TRASH for ground */
    uint32_t nslot;
} orc_inst;

struct orc_ckt {
    orc_deck d;
    uint32_t *kind, *nodes, *bind_dev, *bind_param, *bind_live;
    double *params;
    uint32_t n, nnz, n_nodes, n_inst;
    uint32_t *col_ptr, *row_idx, *diag;
    uint32_t *csr_ptr, *csr_col, *csr_slot;   /* the pattern by rows, columns ascending */
    uint8_t *current_row;
    orc_inst *inst;
};

/* ---- MOS1 bias-independent prep (mos1.va:176-263) ---- */
typedef struct {
    double type, vbit, gammax, phit, beta, lambda, vt, isats, isatd;
    double vcrits, vcritd, gmin, seed_vgs;
} mos_prep;

static void mos1_prep(const double *p, mos_prep *o)
{
    const double tp = p[0];
    const unsigned given = (unsigned)p[24];
    const double tnom = p[23], temp = p[22];
    const double tnomk = tnom + 273.15;
    const double vtnom = tnomk * (KB / QE);
    const double egfet1 = 1.16 - (7.02e-4 * tnomk * tnomk) / (tnomk + 1108.0);
    const double tox = p[10];
    double coxp = ((given & 1u) && tox > 0.0) ? 3.9 * 8.854214871e-12 / tox : 0.0;
    const double phix = p[4];
    const double gammax = p[3];
    const double vtox = p[1];
    double kpx;
    if ((given & 2u) || coxp <= 0.0) kpx = p[2];
    else kpx = p[11] * coxp * 1.0e-4;
    const double tk = temp + 273.15;
    const double vt = tk * (KB / QE);
    const double fact1 = tnomk / 300.15;
    const double fact2 = tk / 300.15;
    const double egfet = 1.16 - (7.02e-4 * tk * tk) / (tk + 1108.0);
    const double arg1 = -egfet1 / (2.0 * KB * tnomk) + 1.1150877 / (2.0 * KB * 300.15);
    const double argt = -egfet / (2.0 * KB * tk) + 1.1150877 / (2.0 * KB * 300.15);
    const double pbfact1 = -2.0 * vtnom * (1.5 * orc_log(fact1) + QE * arg1);
    const double pbfact = -2.0 * vt * (1.5 * orc_log(fact2) + QE * argt);
    const double ratio4 = (tk / tnomk) * sqrt(tk / tnomk);
    const double kpt = kpx / ratio4;
    const double phio = (phix - pbfact1) / fact1;
    const double phit = fact2 * phio + pbfact;
    const double vbit = vtox - tp * (gammax * sqrt(phix)) + 0.5 * (egfet1 - egfet)
                        + tp * 0.5 * (phit - phix);
    const double ist = p[8] * orc_exp(-egfet / vt + egfet1 / vtnom);
    double leff = p[7] - 2.0 * p[9];
    if (leff < 1.0e-10) leff = 1.0e-10;
    const double m = p[12];
    o->type = tp;
    o->vbit = vbit;
    o->gammax = gammax;
    o->phit = phit;
    o->beta = kpt * m * p[6] / leff;
    o->lambda = p[5];
    o->vt = vt;
    o->isats = ist * m;
    o->isatd = ist * m;
    o->vcrits = vt * orc_log(vt / (1.4142135623730951 * o->isats));
    o->vcritd = vt * orc_log(vt / (1.4142135623730951 * o->isatd));
    o->gmin = p[21];
    o->seed_vgs = tp * vbit + gammax * sqrt(phit);
}

/* ---- ngspice limiters (devsup.c as given in
 * dev/solvers/newton-raphson-convergence.md:42-69,151-162) ---- */
static double pnjlim(double vnew, double vold, double vt, double vcrit)
{
    if (vnew > vcrit && fabs(vnew - vold) > vt + vt) {
        if (vold > 0) {
            double arg = (vnew - vold) / vt;
            if (arg > 0) vnew = vold + vt * (2.0 + orc_log(arg - 2.0));
            else vnew = vold - vt * (2.0 + orc_log(2.0 - arg));
        } else {
            vnew = vt * orc_log(vnew / vt);
        }
    } else if (vnew < 0) {
        double arg = vold > 0 ? -1.0 * vold - 1.0 : 2.0 * vold - 1.0;
        if (vnew < arg) vnew = arg;
    }
    return vnew;
}

static double fetlim(double vnew, double vold, double vto)
{
    double vtsthi = fabs(2.0 * (vold - vto)) + 2.0;
    double vtstlo = fabs(vold - vto) + 1.0;
    double vtox = vto + 3.5;
    double delv = vnew - vold;
    if (vold >= vto) {
        if (vold >= vtox) {
            if (delv <= 0) {
                if (vnew >= vtox) {
                    if (-delv > vtstlo) vnew = vold - vtstlo;
                } else {
                    vnew = fmax(vnew, vto + 2.0);
                }
            } else {
                if (delv >= vtsthi) vnew = vold + vtsthi;
            }
        } else {
            if (delv <= 0) vnew = fmax(vnew, vto - 0.5);
            else vnew = fmin(vnew, vto + 4.0);
        }
    } else {
        if (delv <= 0) {
            if (-delv > vtsthi) vnew = vold - vtsthi;
        } else {
            double vtemp = vto + 0.5;
            if (vnew <= vtemp) {
                if (delv > vtstlo) vnew = vold + vtstlo;
            } else {
                vnew = vtemp;
            }
        }
    }
    return vnew;
}

static double limvds(double vnew, double vold)
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

/* jcur (mos1.va:99-110) and its derivative. */
/* ---- junction diode (diode.va, the RS = 0 subset: no breakdown,
 * recombination, knee currents or sidewall; TLEV 0 / TLEVC 0) ---- */
typedef struct {
    double is, vte, vcrit, gmin, cjo, vj, mt, qexp, vfc, f1, inv_f2, f3, tt;
    double bv, xbv, vtebrk;       /* breakdown: armed when bv > 0 */
    double ikf;                   /* high-injection knee IKF * area, 0 off */
    double isr, vterec;           /* recombination current ISR(T), NR vt; isr 0 off */
} dio_prep;

static double pw(double a, double b) { return orc_exp(b * orc_log(a)); }

static void dio1_prep(const double *p, dio_prep *o)
{
    const double n = p[1], area = p[7], mt = p[4], fc = p[6];
    const double tk = p[9] + 273.15, tnomk = p[10] + 273.15;
    const double ratio = tk / tnomk;
    const double vt = KB / QE * tk, vtnom = KB / QE * tnomk;
    const double vte = n * vt;
    const double egfet = 1.16 - (7.02e-4 * tk * tk) / (tk + 1108.0);
    const double egfet1 = 1.16 - (7.02e-4 * tnomk * tnomk) / (tnomk + 1108.0);
    const double fact1 = tnomk / 300.15, fact2 = tk / 300.15;
    const double arg1 = -egfet1 / (2.0 * KB * tnomk) + 1.1150877 / (2.0 * KB * 300.15);
    const double argt = -egfet / (2.0 * KB * tk) + 1.1150877 / (2.0 * KB * 300.15);
    const double pbfact1 = -2.0 * vtnom * (1.5 * orc_log(fact1) + QE * arg1);
    const double pbfact = -2.0 * vt * (1.5 * orc_log(fact2) + QE * argt);
    const double vj = p[3];
    const double pbo = (vj - pbfact1) / fact1;
    const double gmaold = (vj - pbo) / pbo;
    const double vj_t = pbfact + fact2 * pbo;
    const double gmanew = (vj_t - pbo) / pbo;
    const double cjo_t = p[2] * (1.0 + mt * (400e-6 * (tk - 300.15) - gmanew)) /
                         (1.0 + mt * (400e-6 * (tnomk - 300.15) - gmaold)) * area;
    const double xti = p[12], eg = p[11];
    const double scale = area * orc_exp((ratio - 1.0) * eg / vte + (xti / n) * orc_log(ratio));
    o->is = p[0] * scale;
    o->vte = vte;
    o->vcrit = vte * orc_log(vte / (1.4142135623730951 * o->is));
    o->gmin = p[8];
    o->cjo = cjo_t;
    o->vj = vj_t;
    o->mt = mt;
    o->qexp = 1.0 - mt;
    o->vfc = fc * vj_t;
    o->f1 = vj_t * (1.0 - pw(1.0 - fc, 1.0 - mt)) / (1.0 - mt);
    o->inv_f2 = 1.0 / pw(1.0 - fc, 1.0 + mt);
    o->f3 = 1.0 - fc * (1.0 + mt);
    o->tt = p[5];
    /* breakdown (diotemp.c): the knee xbv where -is(T) exp(-(xbv + vd) /
     * vtebrk) carries IBV at vd = -BV; NBV defaults to N (diosetup.c);
     * TCV = 0, so BV(T) = BV */
    const double bv = p[13], ibv = p[14], nbv = p[15] > 0 ? p[15] : n;
    o->bv = bv;
    o->vtebrk = nbv * vt;
    o->xbv = bv;
    if (bv > 0.0 && ibv >= o->is * bv / vt) {
        double xbv = bv - o->vtebrk * orc_log(1.0 + ibv / o->is);
        xbv = bv - o->vtebrk * orc_log(ibv / o->is + 1.0 - xbv / vt);
        o->xbv = xbv;
    }
    o->ikf = p[16] * area;
    /* ISR(T) rides NR and reuses XTI (TLEV 0) */
    const double nr = p[18] > 0 ? p[18] : 2.0;
    o->vterec = nr * vt;
    o->isr = p[17] > 0 ? p[17] * area * orc_exp((ratio - 1.0) * eg / o->vterec + (xti / nr) * orc_log(ratio)) : 0.0;
}

/* I = ihi + gmin vd (diode.va:264-314): the junction current with the 80 vte
 * tangent and the cubic reverse tail, the ISR/NR recombination current in
 * forward bias, the breakdown branch past -xbv, the IKF knee roll-off;
 * g = dI/dvd */
static double dio1_current(const dio_prep *d, double vd, double *g)
{
    double id, gd;
    if (vd >= -3.0 * d->vte) {
        const double u = vd / d->vte;
        const double ul = fmin(u, 80.0);
        const double e = orc_exp(ul);
        id = d->is * (e * (1.0 + (u - ul)) - 1.0);
        gd = d->is * e / d->vte;
    } else {
        const double w = 3.0 * d->vte / (vd * 2.718281828459045);
        const double w3 = w * w * w;
        id = -d->is * (1.0 + w3);
        gd = 3.0 * d->is * w3 / vd;
    }
    double irec = 0.0, grec = 0.0;
    if (d->isr > 0.0 && vd >= -3.0 * d->vte) {
        const double ur = vd / d->vterec;
        const double url = fmin(ur, 80.0);
        const double er = orc_exp(url);
        const double a = 1.0 - vd / d->vj;
        const double gg = a * a + 0.005;
        const double genfac = orc_exp((d->mt / 2.0) * orc_log(gg));
        const double base = er * (1.0 + (ur - url)) - 1.0;
        irec = d->isr * base * genfac;
        /* d genfac / dvd = genfac (mt / 2) (-2 a / vj) / gg */
        const double dgen = genfac * (d->mt / 2.0) * (-2.0 * a / d->vj) / gg;
        grec = d->isr * (er / d->vterec * genfac + base * dgen);
    }
    if (d->bv > 0.0 && vd < -3.0 * d->vte && vd < -d->xbv) {
        const double wb = -(d->xbv + vd) / d->vtebrk;
        const double wl = fmin(wb, 80.0);
        const double e = orc_exp(wl);
        id = -d->is * e * (1.0 + (wb - wl));
        gd = d->is * e / d->vtebrk;
    }
    double ihi = id + irec, ghi = gd + grec;
    if (d->ikf > 0.0 && ihi > 0.0) {
        /* ihi / (1 + s), s = sqrt(ihi / ikf): d/dihi = (1 + s/2) / (1 + s)^2 */
        const double sq = sqrt(ihi / d->ikf);
        const double den = 1.0 + sq;
        ghi = ghi * ((1.0 + 0.5 * sq) / (den * den));
        ihi = ihi / den;
    }
    *g = ghi + d->gmin;
    return ihi + d->gmin * vd;
}

/* depletion charge (two-branch form) + tt I; *c = dq/dvd */
static double dio1_charge(const dio_prep *d, double vd, double i, double g, double *c)
{
    double q = 0.0, cap = 0.0;
    if (d->cjo > 0) {
        if (vd < d->vfc) {
            const double a = 1.0 - vd / d->vj;
            const double pp = orc_exp(d->qexp * orc_log(a));
            q = d->cjo * d->vj * (1.0 - pp) / d->qexp;
            cap = d->cjo * pp / a;
        } else {
            q = d->cjo * (d->f1 + d->inv_f2 * (d->f3 * (vd - d->vfc) + (d->mt / (2.0 * d->vj)) * (vd * vd - d->vfc * d->vfc)));
            cap = d->cjo * d->inv_f2 * (d->f3 + (d->mt / d->vj) * vd);
        }
    }
    if (d->tt > 0) {
        q = q + d->tt * i;
        cap = cap + d->tt * g;
    }
    *c = cap;
    return q;
}

static int dio1_limit(const dio_prep *d, const double *x, const uint32_t *t, double *lvd)
{
    const double xa = x[t[0]], xk = x[t[1]];
    const double vd = xa - xk;
    double v = pnjlim(vd, *lvd, d->vte, d->vcrit);
    const double tol = 16.0 * 2.220446049250313e-16 * fmax(fabs(xa), fabs(xk));
    if (fabs(v - vd) <= tol) v = vd;
    *lvd = v;
    return v != vd;
}

/* out[0..3] Jacobian (a,a) (a,k) (k,a) (k,k), out[4..5] rows a, k; with
 * charges out[6..9] the charge Jacobian, out[10..11] the row charges */
static void dio1_eval(const dio_prep *d, double xa, double xk, double vl, int charges, double *out)
{
    double g;
    double i = dio1_current(d, vl, &g);
    const double dv = (xa - xk) - vl;
    double q = 0.0, c = 0.0;
    if (charges) q = dio1_charge(d, vl, i, g, &c);
    if (dv != 0) {
        i = i + g * dv;
        if (charges) q = q + c * dv;
    }
    out[0] = g; out[1] = -g; out[2] = -g; out[3] = g;
    out[4] = i; out[5] = -i;
    if (charges) {
        out[6] = c; out[7] = -c; out[8] = -c; out[9] = c;
        out[10] = q; out[11] = -q;
    }
}

/* ---- Gummel-Poon BJT (gummel_poon.va, DC part: RB = RC = RE = 0, no excess
 * phase, vertical substrate) ---- */
typedef struct {
    double tp, is, bf, br, vt, nfvt, nrvt, ise, nevt, isc, ncvt, iss, nsvt, inv_vaf, inv_var, inv_ikf, inv_ikr, nkf,
        mgmin, vert, vcrit;
} bjt_prep;

static void bjt1_prep(const double *p, bjt_prep *o)
{
    /* slots: 0 type 1 is 2 bf 3 nf 4 vaf 5 ikf 6 ise 7 ne 8 br 9 nr 10 var 11 ikr 12 isc 13 nc 14 nkf 15 iss
     * 16 ns 17 area 18 m 19 gmin 20 temp 21 tnom 22 eg 23 xti 24 xtb 25 areab 26 areac 27 subs */
    const double tdev = p[20] + 273.15, tnom = p[21] + 273.15;
    const double vt = KB / QE * tdev;
    const double ratlog = orc_log(tdev / tnom);
    const double factlog = (tdev / tnom - 1.0) * p[22] / vt + p[23] * ratlog;
    const double factor = orc_exp(factlog);
    const double bfac = orc_exp(p[24] * ratlog);
    const double m = p[18], ai = p[17] * m;
    const double vert = p[27] > 0 ? 1.0 : 0.0;
    const double abc = m * (vert * p[25] + (1.0 - vert) * p[26]);
    const double ne = p[7], nc = p[13];
    o->tp = p[0];
    o->is = p[1] * (factor * ai);
    o->bf = p[2] * bfac;
    o->br = p[8] * bfac;
    o->vt = vt;
    o->nfvt = p[3] * vt;
    o->nrvt = p[9] * vt;
    o->ise = p[6] * orc_exp(factlog / ne) / bfac * ai;
    o->nevt = ne * vt;
    o->isc = p[12] * orc_exp(factlog / nc) / bfac * abc;
    o->ncvt = nc * vt;
    o->iss = p[15] * factor * ai;
    o->nsvt = p[16] * vt;
    o->inv_vaf = p[4] > 0 ? 1.0 / p[4] : 0.0;
    o->inv_var = p[10] > 0 ? 1.0 / p[10] : 0.0;
    o->inv_ikf = p[5] > 0 ? 1.0 / (p[5] * ai) : 0.0;
    o->inv_ikr = p[11] > 0 ? 1.0 / (p[11] * ai) : 0.0;
    o->nkf = p[14];
    o->mgmin = m * p[19];
    o->vert = vert;
    o->vcrit = vt * orc_log(vt / (1.4142135623730951 * o->is));
}

static double bjt1_pnj(double v, double isat, double vtn, double *g)
{
    if (v >= -3.0 * vtn) {
        const double u = fmin(v / vtn, 80.0);
        const double e = orc_exp(u);
        *g = isat * e / vtn;
        return isat * (e * (1.0 + (v / vtn - u)) - 1.0);
    }
    const double u = 3.0 * vtn / (2.718281828459045 * v);
    const double u3 = u * u * u;
    *g = 3.0 * isat * u3 / v;
    return -isat * (1.0 + u3);
}

static int bjt1_limit(const bjt_prep *b, const double *x, const uint32_t *t, double *lim)
{
    const double xc = x[t[0]], xb = x[t[1]], xe = x[t[2]];
    const double vbe = b->tp * (xb - xe), vbc = b->tp * (xb - xc);
    double ve = pnjlim(vbe, lim[0], b->vt, b->vcrit);
    double vc = pnjlim(vbc, lim[1], b->vt, b->vcrit);
    const double tol = 16.0 * 2.220446049250313e-16 * fmax(fmax(fabs(xc), fabs(xb)), fabs(xe));
    if (fabs(ve - vbe) <= tol) ve = vbe;
    if (fabs(vc - vbc) <= tol) vc = vbc;
    lim[0] = ve;
    lim[1] = vc;
    return ve != vbe || vc != vbc;
}

/* out[0..11] Jacobian (c,c) (c,b) (c,e) (c,s) (b,c) (b,b) (b,e) (e,c) (e,b)
 * (e,e) (s,c) (s,s); out[12..15] rows c, b, e, s */
static void bjt1_eval(const bjt_prep *b, const double *x, const uint32_t *t, const double *lim, double *out)
{
    const double xc = x[t[0]], xb = x[t[1]], xe = x[t[2]], xs = x[t[3]];
    const double tp = b->tp, vbe = lim[0], vbc = lim[1];
    const double dbe = tp * (xb - xe) - vbe, dbc = tp * (xb - xc) - vbc;
    const double vsub = tp * (xs - xc);
    double gbe, gben, gbc, gbcn, gsub;
    const double cbe = bjt1_pnj(vbe, b->is, b->nfvt, &gbe);
    const double cben = bjt1_pnj(vbe, b->ise, b->nevt, &gben) + b->mgmin * vbe;
    gben = gben + b->mgmin;
    const double cbc = bjt1_pnj(vbc, b->is, b->nrvt, &gbc);
    const double cbcn = bjt1_pnj(vbc, b->isc, b->ncvt, &gbcn) + b->mgmin * vbc;
    gbcn = gbcn + b->mgmin;
    const double cdsub = bjt1_pnj(vsub, b->iss, b->nsvt, &gsub) + b->mgmin * vsub;
    gsub = gsub + b->mgmin;
    const double q1 = 1.0 / (1.0 - vbc * b->inv_vaf - vbe * b->inv_var);
    const double q2 = cbe * b->inv_ikf + cbc * b->inv_ikr;
    const double arg = fmax(1e-30, 1.0 + 4.0 * q2);
    const double pp = b->nkf == 0.5 ? sqrt(arg) : orc_exp(b->nkf * orc_log(arg));
    const double qb = 0.5 * q1 * (1.0 + pp);
    const double dp_dq2 = (1.0 + 4.0 * q2 > 1e-30) ? 4.0 * b->nkf * pp / arg : 0.0;
    const double dqb_be = 0.5 * (q1 * q1 * b->inv_var * (1.0 + pp) + q1 * dp_dq2 * (gbe * b->inv_ikf));
    const double dqb_bc = 0.5 * (q1 * q1 * b->inv_vaf * (1.0 + pp) + q1 * dp_dq2 * (gbc * b->inv_ikr));
    double icc = cbe / qb - cbc / qb;
    const double dcc = (cbe - cbc) / (qb * qb);
    const double dicc_be = gbe / qb - dcc * dqb_be;
    const double dicc_bc = -gbc / qb - dcc * dqb_bc;
    double ibe = cbe / b->bf + cben;
    const double gbe_t = gbe / b->bf + gben;
    double ibc = cbc / b->br + cbcn;
    const double gbc_t = gbc / b->br + gbcn;
    const double isc_ = b->vert * cdsub;
    const double gs = b->vert * gsub;
    if (dbe != 0 || dbc != 0) {
        icc = icc + (dicc_be * dbe + dicc_bc * dbc);
        ibe = ibe + gbe_t * dbe;
        ibc = ibc + gbc_t * dbc;
    }
    out[0] = (-dicc_bc + gbc_t) + gs;
    out[1] = (dicc_be + dicc_bc) - gbc_t;
    out[2] = -dicc_be;
    out[3] = -gs;
    out[4] = -gbc_t;
    out[5] = gbe_t + gbc_t;
    out[6] = -gbe_t;
    out[7] = dicc_bc;
    out[8] = -(dicc_be + dicc_bc) - gbe_t;
    out[9] = dicc_be + gbe_t;
    out[10] = -gs;
    out[11] = gs;
    out[12] = (tp * icc - tp * ibc) - tp * isc_;
    out[13] = tp * ibe + tp * ibc;
    out[14] = -tp * icc - tp * ibe;
    out[15] = tp * isc_;
}

static const uint8_t BJT_ENT[12][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 0}, {1, 1},
                                       {1, 2}, {2, 0}, {2, 1}, {2, 2}, {3, 0}, {3, 3}};

static void jcur(double v, double isat, double vtx, double gm, double *i, double *g)
{
    if (v > -3.0 * vtx) {
        double u = fmin(v / vtx, 80.0);
        double e = orc_exp(u);
        *i = isat * (e * (1.0 + (v / vtx - u)) - 1.0) + gm * v;
        *g = isat * e / vtx + gm;
    } else {
        *i = gm * v - isat;
        *g = gm;
    }
}

/* MOS1 limit step: limited (vgs, vds, vbs) from the probe and the last
 * limited values; returns 1 when any value moved. lim: [vgs, vds, vbs, von]. */
static int mos1_limit(const mos_prep *m, const double *x, const uint32_t *t, double *lim)
{
    const double xd = x[t[0]], xg = x[t[1]], xs = x[t[2]], xb = x[t[3]];
    const double tp = m->type;
    const double cgs = tp * (xg - xs), cds = tp * (xd - xs), cbs = tp * (xb - xs);
    const double ogs = lim[0], ods = lim[1], obs = lim[2], von = lim[3];
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
        vbs = pnjlim(vbs, obs, m->vt, m->vcrits);
    } else {
        double vbd = pnjlim(vbs - vds, obs - ods, m->vt, m->vcritd);
        vbs = vbd + vds;
    }
    /* snapLimited: 16 eps of the largest node voltage the device reads. */
    double scale = fmax(fmax(fabs(xd), fabs(xg)), fmax(fabs(xs), fabs(xb)));
    double tol = 16.0 * 2.220446049250313e-16 * scale;
    if (fabs(vgs - cgs) <= tol) vgs = cgs;
    if (fabs(vds - cds) <= tol) vds = cds;
    if (fabs(vbs - cbs) <= tol) vbs = cbs;
    lim[0] = vgs;
    lim[1] = vds;
    lim[2] = vbs;
    return vgs != cgs || vds != cds || vbs != cbs;
}

/* MOS1 DC evaluation at the limited voltages with the residual correction.
 * Stamps rows d, s, b and 11 Jacobian entries (see entry order below). */
static void mos1_eval(const mos_prep *m, const double *x, const uint32_t *t, double *lim,
                      double *f3, double *j11)
{
    const double tp = m->type;
    const double xd = x[t[0]], xg = x[t[1]], xs = x[t[2]], xb = x[t[3]];
    const double vgs = lim[0], vds = lim[1], vbs = lim[2];
    const double dgs = tp * (xg - xs) - vgs;
    const double dds = tp * (xd - xs) - vds;
    const double dbs = tp * (xb - xs) - vbs;
    const double vbd = vbs - vds, vgd = vgs - vds;
    double ibs, gbs, ibd, gbd;
    jcur(vbs, m->isats, m->vt, m->gmin, &ibs, &gbs);
    jcur(vbd, m->isatd, m->vt, m->gmin, &ibd, &gbd);
    double mode, vbx, vgsx;
    if (vds >= 0) { mode = 1.0; vbx = vbs; vgsx = vgs; }
    else { mode = -1.0; vbx = vbd; vgsx = vgd; }
    const double vdsx = mode * vds;
    double sarg;
    if (vbx <= 0) sarg = sqrt(m->phit - vbx);
    else sarg = fmax(sqrt(m->phit) - vbx / (2.0 * sqrt(m->phit)), 0.0);
    const double dsarg = sarg > 0 ? -0.5 / sarg : 0.0;
    const double von = m->vbit * tp + m->gammax * sarg;
    const double vgst = vgsx - von;
    const double bp = m->beta * (1.0 + m->lambda * vdsx);
    double id, gm, gds;
    if (vgst <= 0) {
        id = 0; gm = 0; gds = 0;
    } else if (vgst <= vdsx) {
        id = 0.5 * bp * vgst * vgst;
        gm = bp * vgst;
        gds = 0.5 * m->beta * m->lambda * vgst * vgst;
    } else {
        id = bp * vdsx * (vgst - 0.5 * vdsx);
        gm = bp * vdsx;
        gds = m->beta * m->lambda * vdsx * (vgst - 0.5 * vdsx) + bp * (vgst - vdsx);
    }
    const double gmbs = -gm * (m->gammax * dsarg);
    lim[3] = von;
    /* frame partials of id */
    double ags, ads, abs_;
    if (mode > 0) { ags = gm; ads = gds; abs_ = gmbs; }
    else { ags = gm; ads = -(gm + gds + gmbs); abs_ = gmbs; }
    double ich = tp * mode * id;
    double ib_d = tp * ibd, ib_s = tp * ibs;
    if (dgs != 0 || dds != 0 || dbs != 0) {
        ich = ich + tp * mode * ((ags * dgs + ads * dds) + abs_ * dbs);
        ib_d = ib_d + tp * (gbd * (dbs - dds));
        ib_s = ib_s + tp * (gbs * dbs);
    }
    const double cg = mode * ags, cd = mode * ads, cb = mode * abs_;
    const double cs = -((cg + cd) + cb);
    f3[0] = ich - ib_d;
    f3[1] = -ich - ib_s;
    f3[2] = ib_d + ib_s;
    j11[0] = cd + gbd;   /* (d,d) */
    j11[1] = cg;         /* (d,g) */
    j11[2] = cs;         /* (d,s) */
    j11[3] = cb - gbd;   /* (d,b) */
    j11[4] = -cd;        /* (s,d) */
    j11[5] = -cg;        /* (s,g) */
    j11[6] = -cs + gbs;  /* (s,s) */
    j11[7] = -cb - gbs;  /* (s,b) */
    j11[8] = -gbd;       /* (b,d) */
    j11[9] = -gbs;       /* (b,s) */
    j11[10] = gbd + gbs; /* (b,b) */
}

/* Entry (row terminal, col terminal) lists per kind. */
static const uint8_t R_ENT[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
static const uint8_t MOS_ENT[11][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {2, 0}, {2, 1},
                                       {2, 2}, {2, 3}, {3, 0}, {3, 2}, {3, 3}};

/* ---- construction ---- */
typedef struct { uint32_t r, c; } rc_t;
static int rc_cmp(const void *a, const void *b)
{
    const rc_t *x = a, *y = b;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    return 0;
}

static uint32_t find_slot(const orc_ckt *c, uint32_t r, uint32_t col)
{
    if (r == 0 || col == 0) return TRASH;
    uint32_t lo = c->col_ptr[col], hi = c->col_ptr[col + 1];
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (c->row_idx[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

static void set_err(char *err, size_t cap, const char *msg)
{
    if (err && cap) { strncpy(err, msg, cap - 1); err[cap - 1] = 0; }
}

orc_ckt *orc_ckt_new(const orc_deck *deck, char *err, size_t err_cap)
{
    if (deck->n_nodes < 1) { set_err(err, err_cap, "no nodes"); return NULL; }
    for (uint32_t i = 0; i < deck->n_dev; i++) {
        if (deck->kind[i] >= NKIND) { set_err(err, err_cap, "unknown device kind"); return NULL; }
        for (int t = 0; t < 4; t++)
            if (deck->nodes[i * 4 + t] >= deck->n_nodes) { set_err(err, err_cap, "node out of range"); return NULL; }
    }
    for (uint32_t b = 0; b < deck->n_bind; b++)
        if (deck->bind_dev[b] >= deck->n_dev || deck->bind_param[b] >= NPARAM || deck->bind_live[b] >= deck->n_live) {
            set_err(err, err_cap, "binding out of range");
            return NULL;
        }
    orc_ckt *c = calloc(1, sizeof(orc_ckt));
    c->d = *deck;
    size_t nd = deck->n_dev;
    c->kind = malloc(nd * 4 + 4); memcpy(c->kind, deck->kind, nd * 4);
    c->nodes = malloc(nd * 16 + 4); memcpy(c->nodes, deck->nodes, nd * 16);
    c->params = malloc(nd * NPARAM * 8 + 8); memcpy(c->params, deck->params, nd * NPARAM * 8);
    size_t nb = deck->n_bind;
    c->bind_dev = malloc(nb * 4 + 4); memcpy(c->bind_dev, deck->bind_dev, nb * 4);
    c->bind_param = malloc(nb * 4 + 4); memcpy(c->bind_param, deck->bind_param, nb * 4);
    c->bind_live = malloc(nb * 4 + 4); memcpy(c->bind_live, deck->bind_live, nb * 4);
    c->d.kind = c->kind; c->d.nodes = c->nodes; c->d.params = c->params;
    c->d.bind_dev = c->bind_dev; c->d.bind_param = c->bind_param; c->d.bind_live = c->bind_live;
    c->n_nodes = deck->n_nodes;

    /* instances in stamp order; V branches numbered in that order */
    c->inst = calloc(nd + 1, sizeof(orc_inst));
    uint32_t ni = 0, nbr = 0;
    for (uint32_t k = 0; k < NKIND; k++)
        for (uint32_t i = 0; i < nd; i++) {
            if (c->kind[i] != k) continue;
            orc_inst *in = &c->inst[ni++];
            in->kind = k;
            in->dev = i;
            for (int t = 0; t < 4; t++) in->t[t] = c->nodes[i * 4 + t];
            in->br = k == K_V ? deck->n_nodes + nbr++ : 0;
        }
    c->n_inst = ni;
    c->n = deck->n_nodes + nbr;
    const uint32_t n = c->n;
    c->current_row = calloc(n, 1);
    for (uint32_t i = deck->n_nodes; i < n; i++) c->current_row[i] = 1;

    /* pattern */
    size_t cap = n + ni * 16 + 16, cnt = 0;
    rc_t *ent = malloc(cap * sizeof(rc_t));
    for (uint32_t i = 0; i < n; i++) ent[cnt++] = (rc_t){i, i};
    for (uint32_t i = 0; i < ni; i++) {
        orc_inst *in = &c->inst[i];
        if (in->kind == K_R || in->kind == K_C || in->kind == K_D) {
            for (int e = 0; e < 4; e++) ent[cnt++] = (rc_t){in->t[R_ENT[e][0]], in->t[R_ENT[e][1]]};
        } else if (in->kind == K_V) {
            ent[cnt++] = (rc_t){in->t[0], in->br};
            ent[cnt++] = (rc_t){in->t[1], in->br};
            ent[cnt++] = (rc_t){in->br, in->t[0]};
            ent[cnt++] = (rc_t){in->br, in->t[1]};
        } else if (in->kind == K_MOS) {
            for (int e = 0; e < 11; e++) ent[cnt++] = (rc_t){in->t[MOS_ENT[e][0]], in->t[MOS_ENT[e][1]]};
        } else if (in->kind == K_Q) {
            for (int e = 0; e < 12; e++) ent[cnt++] = (rc_t){in->t[BJT_ENT[e][0]], in->t[BJT_ENT[e][1]]};
        }
    }
    size_t keep = 0;
    for (size_t e = 0; e < cnt; e++)
        if ((ent[e].r != 0 && ent[e].c != 0) || (ent[e].r == 0 && ent[e].c == 0)) ent[keep++] = ent[e];
    qsort(ent, keep, sizeof(rc_t), rc_cmp);
    size_t u = 0;
    for (size_t e = 0; e < keep; e++)
        if (u == 0 || ent[e].r != ent[u - 1].r || ent[e].c != ent[u - 1].c) ent[u++] = ent[e];
    c->nnz = (uint32_t)u;
    c->col_ptr = calloc(n + 1, 4);
    c->row_idx = malloc(u * 4 + 4);
    for (size_t e = 0; e < u; e++) { c->col_ptr[ent[e].c + 1]++; c->row_idx[e] = ent[e].r; }
    for (uint32_t j = 0; j < n; j++) c->col_ptr[j + 1] += c->col_ptr[j];
    free(ent);
    c->diag = malloc(n * 4);
    for (uint32_t i = 0; i < n; i++) c->diag[i] = find_slot(c, i, i);
    c->diag[0] = c->col_ptr[0];   /* (0,0) is column 0's only entry */
    c->csr_ptr = calloc(n + 1, 4);
    c->csr_col = malloc((c->nnz + 1) * 4);
    c->csr_slot = malloc((c->nnz + 1) * 4);
    for (uint32_t p = 0; p < c->nnz; p++) c->csr_ptr[c->row_idx[p] + 1]++;
    for (uint32_t i = 0; i < n; i++) c->csr_ptr[i + 1] += c->csr_ptr[i];
    {
        uint32_t *at = malloc((n + 1) * 4);
        memcpy(at, c->csr_ptr, n * 4);
        for (uint32_t j = 0; j < n; j++)
            for (uint32_t p = c->col_ptr[j]; p < c->col_ptr[j + 1]; p++) {
                const uint32_t e = at[c->row_idx[p]]++;
                c->csr_col[e] = j;
                c->csr_slot[e] = p;
            }
        free(at);
    }
    for (uint32_t i = 0; i < ni; i++) {
        orc_inst *in = &c->inst[i];
        if (in->kind == K_R || in->kind == K_C || in->kind == K_D) {
            in->nslot = 4;
            for (int e = 0; e < 4; e++) in->slot[e] = find_slot(c, in->t[R_ENT[e][0]], in->t[R_ENT[e][1]]);
        } else if (in->kind == K_V) {
            in->nslot = 4;
            in->slot[0] = find_slot(c, in->t[0], in->br);
            in->slot[1] = find_slot(c, in->t[1], in->br);
            in->slot[2] = find_slot(c, in->br, in->t[0]);
            in->slot[3] = find_slot(c, in->br, in->t[1]);
        } else if (in->kind == K_MOS) {
            in->nslot = 11;
            for (int e = 0; e < 11; e++) in->slot[e] = find_slot(c, in->t[MOS_ENT[e][0]], in->t[MOS_ENT[e][1]]);
        } else if (in->kind == K_Q) {
            in->nslot = 12;
            for (int e = 0; e < 12; e++) in->slot[e] = find_slot(c, in->t[BJT_ENT[e][0]], in->t[BJT_ENT[e][1]]);
        }
    }
    return c;
}

void orc_ckt_free(orc_ckt *c)
{
    if (!c) return;
    free(c->kind); free(c->nodes); free(c->params); free(c->bind_dev); free(c->bind_param);
    free(c->bind_live); free(c->inst); free(c->current_row); free(c->col_ptr); free(c->row_idx);
    free(c->diag); free(c->csr_ptr); free(c->csr_col); free(c->csr_slot); free(c);
}

uint32_t orc_ckt_n(const orc_ckt *c) { return c->n; }
uint32_t orc_ckt_nnz(const orc_ckt *c) { return c->nnz; }

void orc_ckt_pattern(const orc_ckt *c, uint32_t *col_ptr, uint32_t *row_idx,
                     uint32_t *diag_slots, uint8_t *current_row)
{
    if (col_ptr) memcpy(col_ptr, c->col_ptr, (size_t)(c->n + 1) * 4);
    if (row_idx) memcpy(row_idx, c->row_idx, (size_t)c->nnz * 4);
    if (diag_slots) memcpy(diag_slots, c->diag, (size_t)c->n * 4);
    if (current_row) memcpy(current_row, c->current_row, c->n);
}

/* ---- one point's state ---- */
typedef struct {
    double *p;          /* [n_dev][NPARAM] with bindings applied */
    mos_prep *mp;       /* per instance (MOS only used) */
    dio_prep *dp;       /* per instance (diodes only used) */
    bjt_prep *bp;       /* per instance (BJTs only used) */
    double *lim;        /* [n_inst][4]: MOS (vgs, vds, vbs, von), diode vd */
    double src;         /* independent sources scaled by src (the .op ladder's
                         * source rung; 1 everywhere else): vsource.va:327-329,
                         * isource.va:326-328 multiply the output by
                         * $simparam("sourceScaleFactor") */
} lane_dev;

static void lane_dev_init(const orc_ckt *c, const double *live, lane_dev *ld)
{
    size_t nd = c->d.n_dev;
    ld->src = 1.0;
    ld->p = malloc(nd * NPARAM * 8 + 8);
    memcpy(ld->p, c->params, nd * NPARAM * 8);
    for (uint32_t b = 0; b < c->d.n_bind; b++)
        ld->p[(size_t)c->bind_dev[b] * NPARAM + c->bind_param[b]] = live[c->bind_live[b]];
    ld->mp = calloc(c->n_inst + 1, sizeof(mos_prep));
    ld->lim = calloc((size_t)c->n_inst * 4 + 4, 8);
    ld->dp = calloc(c->n_inst + 1, sizeof(dio_prep));
    ld->bp = calloc(c->n_inst + 1, sizeof(bjt_prep));
    for (uint32_t i = 0; i < c->n_inst; i++) {
        if (c->inst[i].kind == K_MOS) mos1_prep(ld->p + (size_t)c->inst[i].dev * NPARAM, &ld->mp[i]);
        if (c->inst[i].kind == K_D) dio1_prep(ld->p + (size_t)c->inst[i].dev * NPARAM, &ld->dp[i]);
        if (c->inst[i].kind == K_Q) bjt1_prep(ld->p + (size_t)c->inst[i].dev * NPARAM, &ld->bp[i]);
    }
}

static void lane_dev_free(lane_dev *ld) { free(ld->p); free(ld->mp); free(ld->dp); free(ld->bp); free(ld->lim); }

/* cold start: MODEINITJCT seeds into the limiting state */
static void seed_cold(const orc_ckt *c, lane_dev *ld)
{
    for (uint32_t i = 0; i < c->n_inst; i++) {
        if (c->inst[i].kind == K_D) ld->lim[(size_t)i * 4] = ld->dp[i].vcrit;
        if (c->inst[i].kind == K_Q) {
            ld->lim[(size_t)i * 4] = ld->bp[i].vcrit;
            ld->lim[(size_t)i * 4 + 1] = 0.0;
        }
        if (c->inst[i].kind != K_MOS) continue;
        double *lim = ld->lim + (size_t)i * 4;
        lim[0] = ld->mp[i].seed_vgs;
        lim[1] = 0.0;
        lim[2] = -1.0;
        lim[3] = 0.0;
    }
}

/* warm start: the limiting state is the start point's own branch voltages.
 * keep_von: MOS fetlim thresholds stay the von the last evaluation stored
 * (an accepted transient step: the predictor's limit pass runs against the
 * accepted point with the von of its last load, mos1.va:293-297) */
static void seed_at_von(const orc_ckt *c, lane_dev *ld, const double *x, int keep_von)
{
    for (uint32_t i = 0; i < c->n_inst; i++) {
        const orc_inst *in = &c->inst[i];
        if (in->kind == K_D) ld->lim[(size_t)i * 4] = x[in->t[0]] - x[in->t[1]];
        if (in->kind == K_Q) {
            ld->lim[(size_t)i * 4] = ld->bp[i].tp * (x[in->t[1]] - x[in->t[2]]);
            ld->lim[(size_t)i * 4 + 1] = ld->bp[i].tp * (x[in->t[1]] - x[in->t[0]]);
        }
        if (in->kind != K_MOS) continue;
        double *lim = ld->lim + (size_t)i * 4;
        const double tp = ld->mp[i].type;
        const double xd = x[in->t[0]], xg = x[in->t[1]], xs = x[in->t[2]], xb = x[in->t[3]];
        lim[0] = tp * (xg - xs);
        lim[1] = tp * (xd - xs);
        lim[2] = tp * (xb - xs);
        if (!keep_von) lim[3] = 0.0;
    }
}

static void seed_at(const orc_ckt *c, lane_dev *ld, const double *x) { seed_at_von(c, ld, x, 0); }

/* Static stamps into g (pattern slots) and rhs; with charges, the charge
 * Jacobian into cp (pattern slots) and the row charges into q. t is the
 * source time (0 for the operating point). */
static void assemble_t(const orc_ckt *c, lane_dev *ld, const double *x, double t, int charges, double *g,
                       double *rhs, double *cp, double *q)
{
    const uint32_t n = c->n;
    memset(g, 0, (size_t)(c->nnz + 1) * 8);
    memset(rhs, 0, (size_t)(n + 1) * 8);
    if (charges) {
        memset(cp, 0, (size_t)(c->nnz + 1) * 8);
        memset(q, 0, (size_t)(n + 1) * 8);
    }
    /* ground rows land in rhs[n], ground entries in g[nnz]: the trash */
#define ROW(u) ((u) == 0 ? n : (u))
#define SLOT(s) ((s) == TRASH ? c->nnz : (s))
    for (uint32_t i = 0; i < c->n_inst; i++) {
        const orc_inst *in = &c->inst[i];
        const double *p = ld->p + (size_t)in->dev * NPARAM;
        switch (in->kind) {
        case K_R: {
            const double r = p[0];
            const double cur = (x[in->t[0]] - x[in->t[1]]) / r;
            const double gg = 1.0 / r;
            rhs[ROW(in->t[0])] += cur;
            rhs[ROW(in->t[1])] += -cur;
            g[SLOT(in->slot[0])] += gg;
            g[SLOT(in->slot[1])] += -gg;
            g[SLOT(in->slot[2])] += -gg;
            g[SLOT(in->slot[3])] += gg;
            break;
        }
        case K_C: {
            if (!charges) break;
            const double cv = p[0];
            const double qq = cv * (x[in->t[0]] - x[in->t[1]]);
            cp[SLOT(in->slot[0])] += cv;
            cp[SLOT(in->slot[1])] += -cv;
            cp[SLOT(in->slot[2])] += -cv;
            cp[SLOT(in->slot[3])] += cv;
            q[ROW(in->t[0])] += qq;
            q[ROW(in->t[1])] += -qq;
            break;
        }
        case K_V: {
            const double ib = x[in->br];
            /* waveform (ESL_V_WAVE 1: step to p[2] after p[3]) */
            const double dc = ((p[1] == 1.0 && t > p[3]) ? p[2] : p[0]) * ld->src;
            rhs[ROW(in->t[0])] += ib;
            rhs[ROW(in->t[1])] += -ib;
            rhs[in->br] += (x[in->t[0]] - x[in->t[1]]) - dc;
            g[SLOT(in->slot[0])] += 1.0;
            g[SLOT(in->slot[1])] += -1.0;
            g[SLOT(in->slot[2])] += 1.0;
            g[SLOT(in->slot[3])] += -1.0;
            break;
        }
        case K_I: {
            const double iv = p[0] * ld->src;
            rhs[ROW(in->t[0])] += iv;
            rhs[ROW(in->t[1])] += -iv;
            break;
        }
        case K_D: {
            double o[12] = {0};
            dio1_eval(&ld->dp[i], x[in->t[0]], x[in->t[1]], ld->lim[(size_t)i * 4], charges, o);
            for (int e = 0; e < 4; e++) g[SLOT(in->slot[e])] += o[e];
            rhs[ROW(in->t[0])] += o[4];
            rhs[ROW(in->t[1])] += o[5];
            if (charges) {
                for (int e = 0; e < 4; e++) cp[SLOT(in->slot[e])] += o[6 + e];
                q[ROW(in->t[0])] += o[10];
                q[ROW(in->t[1])] += o[11];
            }
            break;
        }
        case K_Q: {
            double o[16];
            bjt1_eval(&ld->bp[i], x, in->t, ld->lim + (size_t)i * 4, o);
            for (int e = 0; e < 12; e++) g[SLOT(in->slot[e])] += o[e];
            for (int r = 0; r < 4; r++) rhs[ROW(in->t[r])] += o[12 + r];
            break;
        }
        case K_MOS: {
            double f3[3], j11[11];
            mos1_eval(&ld->mp[i], x, in->t, ld->lim + (size_t)i * 4, f3, j11);
            rhs[ROW(in->t[0])] += f3[0];
            rhs[ROW(in->t[2])] += f3[1];
            rhs[ROW(in->t[3])] += f3[2];
            for (int e = 0; e < 11; e++) g[SLOT(in->slot[e])] += j11[e];
            break;
        }
        default:
            break;
        }
    }
#undef ROW
#undef SLOT
    g[c->diag[0]] += 1.0;
    rhs[0] += x[0];
}

static void assemble(const orc_ckt *c, lane_dev *ld, const double *x, double *g, double *rhs)
{
    assemble_t(c, ld, x, 0.0, 0, g, rhs, NULL, NULL);
}

static int apply_limits(const orc_ckt *c, lane_dev *ld, const double *x)
{
    int any = 0;
    for (uint32_t i = 0; i < c->n_inst; i++) {
        if (c->inst[i].kind == K_MOS) any |= mos1_limit(&ld->mp[i], x, c->inst[i].t, ld->lim + (size_t)i * 4);
        if (c->inst[i].kind == K_D) any |= dio1_limit(&ld->dp[i], x, c->inst[i].t, ld->lim + (size_t)i * 4);
        if (c->inst[i].kind == K_Q) any |= bjt1_limit(&ld->bp[i], x, c->inst[i].t, ld->lim + (size_t)i * 4);
    }
    return any;
}

void orc_ckt_assemble_cold(orc_ckt *c, const double *live, double *g, double *rhs)
{
    lane_dev ld;
    lane_dev_init(c, live, &ld);
    seed_cold(c, &ld);
    double *x = calloc(c->n + 1, 8);
    double *gg = malloc((size_t)(c->nnz + 1) * 8), *rr = malloc((size_t)(c->n + 1) * 8);
    assemble(c, &ld, x, gg, rr);
    memcpy(g, gg, (size_t)c->nnz * 8);
    memcpy(rhs, rr, (size_t)c->n * 8);
    free(x); free(gg); free(rr);
    lane_dev_free(&ld);
}

void orc_ckt_assemble_at(orc_ckt *c, const double *live, const double *x, double *g, double *rhs)
{
    orc_ckt_assemble_scaled(c, live, x, 1.0, g, rhs);
}

void orc_ckt_assemble_scaled(orc_ckt *c, const double *live, const double *x, double src, double *g, double *rhs)
{
    lane_dev ld;
    lane_dev_init(c, live, &ld);
    ld.src = src;
    seed_at(c, &ld, x);
    double *gg = malloc((size_t)(c->nnz + 1) * 8), *rr = malloc((size_t)(c->n + 1) * 8);
    assemble(c, &ld, x, gg, rr);
    memcpy(g, gg, (size_t)c->nnz * 8);
    memcpy(rhs, rr, (size_t)c->n * 8);
    free(gg); free(rr);
    lane_dev_free(&ld);
}

void orc_ckt_assemble_charges(orc_ckt *c, const double *live, const double *x, double t, double *g, double *rhs,
                              double *cp, double *q)
{
    lane_dev ld;
    lane_dev_init(c, live, &ld);
    seed_at(c, &ld, x);
    double *gg = malloc((size_t)(c->nnz + 1) * 8), *rr = malloc((size_t)(c->n + 1) * 8);
    double *cc = malloc((size_t)(c->nnz + 1) * 8), *qq = malloc((size_t)(c->n + 1) * 8);
    assemble_t(c, &ld, x, t, 1, gg, rr, cc, qq);
    memcpy(g, gg, (size_t)c->nnz * 8);
    memcpy(rhs, rr, (size_t)c->n * 8);
    memcpy(cp, cc, (size_t)c->nnz * 8);
    memcpy(q, qq, (size_t)c->n * 8);
    free(gg); free(rr); free(cc); free(qq);
    lane_dev_free(&ld);
}

/* ---- Newton ---- */
static uint64_t db(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }
static double bd(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }

/* plausibleLimit (converger.zig:377-384) */
static double plausible_limit(const orc_ckt *c, const double *x, const double *f, const double *g)
{
    double gm = 1.0;
    for (uint32_t i = 0; i < c->n; i++) gm = fmax(gm, fabs(g[c->diag[i]]));
    uint64_t xm = 0, fm = 0;
    for (uint32_t i = 0; i < c->n; i++) {
        uint64_t a = db(fabs(x[i])), b = db(fabs(f[i]));
        if (a > xm) xm = a;
        if (b > fm) fm = b;
    }
    double scale = fmin(fmax(bd(xm), bd(fm) * gm), 1e18);
    return !(1e3 * scale > 1e9) ? 1e9 : 1e3 * scale;
}

typedef struct {
    const orc_ckt *c;
    const double *values;
    size_t first, count;
    const orc_tol *tol;
    const orc_lu *lead;        /* the lead tape (read-only) */
    const uint32_t *q;
    const double *start_x;
    double *x;
    uint32_t *iterations, *reasons, *repivots;
    orc_prec prec;
    uint32_t *fp64_lanes, *fell_back;
} job_t;

/* One Newton linear solve with FP32 factors and FP64 refinement:
 * dx = -A^-1 F. r = -F - A dx in FP64, row by row over the pattern in
 * ascending column order. Returns 0 when the refinement converged, 1 when it
 * stalled or ran out of steps. */
static int refine_solve(const orc_ckt *c, const orc_lu *lu, orc_lu32 *f, const double *g, const double *rhs,
                        const orc_prec *pr, double *dx, double *r, float *b32, float *x32)
{
    const uint32_t n = c->n;
    uint64_t fm = 0;
    for (uint32_t i = 0; i < n; i++) {
        r[i] = -rhs[i];
        dx[i] = 0;
        const uint64_t b = db(fabs(rhs[i]));
        if (b > fm) fm = b;
    }
    double prev = INFINITY;
    for (uint32_t s = 0; s < pr->refine_max; s++) {
        for (uint32_t i = 0; i < n; i++) b32[i] = (float)r[i];
        orc_lu_solve32(lu, f, b32, x32);
        for (uint32_t i = 0; i < n; i++) dx[i] = dx[i] + (double)x32[i];
        uint64_t rm = 0, am = 0;
        for (uint32_t i = 0; i < n; i++) {
            double acc = -rhs[i], aa = 0;
            for (uint32_t e = c->csr_ptr[i]; e < c->csr_ptr[i + 1]; e++) {
                const double t = g[c->csr_slot[e]] * dx[c->csr_col[e]];
                acc = acc - t;
                aa = aa + fabs(t);
            }
            r[i] = acc;
            const uint64_t rb = db(fabs(acc)), ab = db(aa);
            if (rb > rm) rm = rb;
            if (ab > am) am = ab;
        }
        const double target = pr->refine_tol * fmax(bd(fm), bd(am));
        if (bd(rm) <= target) return 0;
        if (!(bd(rm) <= 0.5 * prev)) return 1;
        prev = bd(rm);
    }
    return 1;
}

/* ---- the .op ladder's gmin rung (dc/op.zig:160-222, ngspice cktop.c
 * dynamic_gmin): after a failed plain rung, Newton runs with a conductance
 * gmin from every node row to ground (A_ii += gmin, F_i += gmin x_i, the
 * load of converger.zig:699-713 on the node diagonals), gmin descending from
 * 1e-3 by an adaptive factor to 1e-12, each run at most itl2 = 50
 * iterations; then a clean run without the shunts from the last point. A
 * run that fails backs up toward the last good gmin by the 4th root of the
 * factor and restarts from the last converged point (cold while none).
 * The GPU host path runs this control, its rungs on the device. ---- */
#define ORC_GMIN_TARGET 1e-12
#define ORC_GMIN_START 1e-2
#define ORC_ITL2 50u

/* Scratch of one lane's FP64 Newton runs. */
typedef struct {
    double *x_old, *dx, *g, *rhs, *sdiag, *x_good;
} newton_ws;

/* One plain FP64 Newton run of a lane: cold (x = 0, cold limiting state,
 * INITFIX) or from x as given (limiting state seeded at x), at most max_iter
 * iterations, gmin shunts on the node diagonals when gmin > 0; the gates of
 * solve_lane. Returns the reason (0 converged, 1 not converged, 3 singular)
 * and the iterations in *its; re-pivots the lane's tape lu in place. */
static uint32_t newton_run(const orc_ckt *c, lane_dev *ld, orc_lu *lu, const uint32_t *q, const orc_tol *tol,
                           uint32_t max_iter, double gmin, int cold, double *x, newton_ws *w, uint32_t *its,
                           uint32_t *repiv)
{
    const uint32_t n = c->n;
    const double wide = 1.0 / 1e-3;
    int init_fix;
    if (cold) {
        memset(x, 0, (size_t)n * 8);
        seed_cold(c, ld);
        init_fix = 1;
    } else {
        seed_at(c, ld, x);
        init_fix = 0;
    }
    double limit = 1e9;
    uint32_t iter = 0;
    for (; iter < max_iter; iter++) {
        assemble(c, ld, x, w->g, w->rhs);
        if (gmin > 0)
            for (uint32_t i = 1; i < n; i++)
                if (!c->current_row[i]) {
                    w->g[c->diag[i]] += gmin;
                    w->rhs[i] += gmin * x[i];
                }
        if (iter == 0) limit = plausible_limit(c, x, w->rhs, w->g);
        for (uint32_t i = 0; i < n; i++) w->sdiag[i] = w->g[c->diag[i]];
        /* a factor that met a singular matrix leaves no usable LU: the next
         * run factors afresh instead of replaying (direct.zig:268-298) */
        if (!lu->factored || orc_lu_lane_refactor(lu, c->col_ptr, w->g, 1e-12, wide)) {
            (*repiv)++;
            if (orc_lu_factor(lu, c->col_ptr, c->row_idx, w->g, q, 1e-3) != 0) {
                *its = iter + 1;
                return 3;
            }
        }
        orc_lu_solve_neg(lu, w->rhs, w->dx);
        uint64_t worst = 0;
        for (uint32_t i = 0; i < n; i++) {
            const double xo = x[i], d = w->dx[i], xn = xo + d;
            w->x_old[i] = xo;
            x[i] = xn;
            const double t = tol->reltol * fmax(fabs(xn), fabs(xo)) + (c->current_row[i] ? tol->abstol : tol->vntol);
            const int finite = fabs(xn) < INFINITY && fabs(d) < INFINITY;
            const double qq = fabs(d) / t;
            const double score = !finite ? INFINITY : (qq == qq ? qq : 0.0);
            const uint64_t sb = db(score);
            if (sb > worst) worst = sb;
        }
        const int limited = apply_limits(c, ld, x);
        int ok = !limited && iter != 0 && bd(worst) < 1.0;
        if (ok)
            for (uint32_t i = 0; i < n; i++)
                if (!(fabs(x[i]) <= limit)) { ok = 0; break; }
        if (ok)
            for (uint32_t i = 0; i < n; i++) {
                const double s0 = fabs(w->sdiag[i]);
                const double s = s0 == 0 ? 1.0 : s0;
                const double rt = fmax(tol->residual_tol, 10.0 * s * (tol->reltol * fabs(x[i]) + tol->vntol));
                if (!(fabs(w->rhs[i]) <= rt)) { ok = 0; break; }
            }
        if (ok && init_fix) {
            init_fix = 0;
        } else if (ok) {
            memcpy(x, w->x_old, (size_t)n * 8);
            *its = iter + 1;
            return 0;
        }
    }
    *its = iter;
    return 1;
}

/* The ladder after a failed plain rung of `plain` iterations (the oracle's
 * tests can also run it after a converged one, without the gmin rung, or
 * without the source rung and final run: ORC_TOL_LADDER_ALWAYS,
 * ORC_TOL_LADDER_NO_GMIN, ORC_TOL_LADDER_NO_SRC); x is the
 * lane's state (overwritten). Returns the reason (0 or 1) and the lane's
 * total iterations in *total.
 *
 * Rung 2, dynamic gmin stepping (dc/op.zig:160-222): see above; a clean run
 * that converges ends the ladder.
 * Rung 3, source stepping (dc/op.zig:228-278): from a cold start, Newton
 * runs of at most itl2 iterations with every independent source scaled by
 * lambda, lambda from 0 by a step of 0.25 that grows 1.5x after each
 * converged run and halves after a failed one (the run retried from the
 * last converged lambda and point); it stops at lambda = 1 converged, when
 * the step falls below 1e-4, after 100 runs, or when the cold lambda = 0
 * run fails. Then a final run at full sources from where it stopped
 * (dc/op.zig:279-288); that run's outcome is the point's. */
static uint32_t op_ladder(const orc_ckt *c, lane_dev *ld, orc_lu *lu, const uint32_t *q, const orc_tol *tol,
                          double *x, newton_ws *w, uint32_t plain, uint32_t *total, uint32_t *repiv)
{
    const uint32_t n = c->n;
    double factor = 10.0, good = ORC_GMIN_START, gv = good / factor;
    int have_good = 0, cold = 1;
    uint32_t tot = plain, its = 0;
    const uint32_t gmin_runs = (tol->reserved & ORC_TOL_LADDER_NO_GMIN) ? 0u : 100u;
    for (uint32_t solves = 0; solves < gmin_runs; solves++) {
        const uint32_t r = newton_run(c, ld, lu, q, tol, ORC_ITL2, gv, cold, x, w, &its, repiv);
        tot += its;
        if (r == 0) {
            if (gv <= ORC_GMIN_TARGET) {
                const uint32_t rc = newton_run(c, ld, lu, q, tol, tol->max_iter, 0.0, 0, x, w, &its, repiv);
                tot += its;
                if (rc == 0) {
                    *total = tot;
                    return 0;
                }
                break;
            }
            memcpy(w->x_good, x, (size_t)n * 8);
            have_good = 1;
            good = gv;
            if (its <= ORC_ITL2 / 4) factor = fmin(factor * sqrt(factor), 10.0);
            else if (its > 3 * ORC_ITL2 / 4) factor = fmax(sqrt(factor), 1.00005);
            if (gv < factor * ORC_GMIN_TARGET) {
                factor = gv / ORC_GMIN_TARGET;
                gv = ORC_GMIN_TARGET;
            } else {
                gv /= factor;
            }
            cold = 0;
        } else {
            if (factor < 1.00005) break;
            factor = sqrt(sqrt(factor));
            gv = good / factor;
            if (have_good) memcpy(x, w->x_good, (size_t)n * 8);
            cold = !have_good;
        }
    }
    if (tol->reserved & ORC_TOL_LADDER_NO_SRC) {
        *total = tot;
        return 1;
    }
    /* rung 3: source stepping */
    double lambda = 0.0, lambda_good = -1.0, delta = 0.25;
    for (uint32_t solves = 0; solves < 100; solves++) {
        ld->src = lambda;
        const uint32_t r = newton_run(c, ld, lu, q, tol, ORC_ITL2, 0.0, lambda_good < 0.0, x, w, &its, repiv);
        tot += its;
        if (r == 0) {
            if (lambda >= 1.0) break;
            lambda_good = lambda;
            memcpy(w->x_good, x, (size_t)n * 8);
            delta *= 1.5;
            lambda = fmin(lambda + delta, 1.0);
        } else {
            delta *= 0.5;
            if (delta < 1e-4) break;
            if (lambda_good < 0.0) break;
            memcpy(x, w->x_good, (size_t)n * 8);
            lambda = fmin(lambda_good + delta, 1.0);
        }
    }
    ld->src = 1.0;
    const uint32_t rf = newton_run(c, ld, lu, q, tol, tol->max_iter, 0.0, 0, x, w, &its, repiv);
    tot += its;
    *total = tot;
    return rf == 0 ? 0 : 1;
}

/* one lane's Newton (converger.newtonLanes at W = 1 over a lane solver) */
static void solve_lane(const job_t *j, size_t l)
{
    const orc_ckt *c = j->c;
    const uint32_t n = c->n, nnz = c->nnz;
    const orc_tol *tol = j->tol;
    lane_dev ld;
    lane_dev_init(c, j->values + l * c->d.n_live, &ld);
    double *x = calloc(n + 1, 8), *x_old = calloc(n + 1, 8), *dx = calloc(n + 1, 8);
    double *g = calloc(nnz + 1, 8), *rhs = calloc(n + 1, 8), *sdiag = calloc(n + 1, 8);
    double *rr = calloc(n + 1, 8);
    float *b32 = calloc(n + 1, 4), *x32 = calloc(n + 1, 4);
    orc_lu32 *f32 = orc_lu32_new(n);
    int lane64 = j->prec.precision == 0;
    int fresh_iter = -1;
    int init_fix;
    if (j->start_x) {
        memcpy(x, j->start_x + l * n, (size_t)n * 8);
        seed_at(c, &ld, x);
        init_fix = 0;
    } else {
        seed_cold(c, &ld);
        init_fix = 1;
    }
    /* this lane's tape: a private copy of the lead's pattern, re-pivoted on
     * a failed replay */
    orc_lu *lu = orc_lu_new(n, nnz);
    {
        const orc_lu *s = j->lead;
        memcpy(lu->q, s->q, n * 4); memcpy(lu->pinv, s->pinv, n * 4);
        memcpy(lu->prow, s->prow, nnz * 4); memcpy(lu->lp, s->lp, (n + 1) * 4);
        memcpy(lu->up, s->up, (n + 1) * 4);
        memcpy(lu->void_col, s->void_col, n); memcpy(lu->scaled, s->scaled, n);
        memcpy(lu->void_slots, s->void_slots, (size_t)s->n_void_slots * 4);
        lu->n_void_slots = s->n_void_slots;
        lu->cap_l = s->nnz_l + 16; lu->cap_u = s->nnz_u + 16;
        lu->li = realloc(lu->li, lu->cap_l * 4); lu->lx = realloc(lu->lx, lu->cap_l * 8);
        lu->ui = realloc(lu->ui, lu->cap_u * 4); lu->ux = realloc(lu->ux, lu->cap_u * 8);
        memcpy(lu->li, s->li, s->nnz_l * 4); memcpy(lu->ui, s->ui, s->nnz_u * 4);
        lu->nnz_l = s->nnz_l; lu->nnz_u = s->nnz_u;
        lu->factored = 1;
    }
    uint32_t repiv = 0;
    uint32_t reason = 1, iters = tol->max_iter;
    double limit = 1e9;
    const double wide = 1.0 / 1e-3;
    /* learned start: warm 1 while active, 2 after the restart from nominal */
    uint32_t warm = (j->start_x && j->prec.nominal_x) ? 1 : 0, k0 = 0;
    double f0 = 0;
    uint32_t iter = 0;
    for (;; iter++) {
        if (iter - k0 >= tol->max_iter) {
            if (warm != 1) break;
            memcpy(x, j->prec.nominal_x, (size_t)n * 8);
            seed_at(c, &ld, x);
            k0 = iter;
            warm = 2;
        }
        assemble(c, &ld, x, g, rhs);
        if (warm == 1) {
            uint64_t fb = 0;
            for (uint32_t i = 0; i < n; i++) {
                const uint64_t b = db(fabs(rhs[i]));
                if (b > fb) fb = b;
            }
            if (iter == k0) {
                f0 = bd(fb);
            } else if (iter == k0 + 1 && !(bd(fb) <= f0)) {
                memcpy(x, j->prec.nominal_x, (size_t)n * 8);
                seed_at(c, &ld, x);
                k0 = iter;
                warm = 2;
                iter--;
                continue;
            }
        }
        if (iter == k0) limit = plausible_limit(c, x, rhs, g);
        for (uint32_t i = 0; i < n; i++) sdiag[i] = g[c->diag[i]];
        int singular = 0, done = 0;
        if (!lane64) {
            /* FP32 factors: a failed replay re-pivots (from the FP64
             * matrix); a replay that fails on the tape just built from this
             * very matrix, or a refinement that stalls, sends the lane to
             * FP64 for this iteration and the rest. */
            int fail = orc_lu_lane_refactor32(lu, f32, c->col_ptr, g, 1e-12, wide);
            if (fail && fresh_iter != (int)iter) {
                repiv++;
                if (orc_lu_factor(lu, c->col_ptr, c->row_idx, g, j->q, 1e-3) != 0) {
                    singular = 1;
                } else {
                    fresh_iter = (int)iter;
                    fail = orc_lu_lane_refactor32(lu, f32, c->col_ptr, g, 1e-12, wide);
                    if (fail) {
                        lane64 = 1;   /* FP64 replay of the fresh tape == the factor just made */
                        orc_lu_solve_neg(lu, rhs, dx);
                        done = 1;
                    }
                }
            } else if (fail) {
                lane64 = 1;
            }
            if (!singular && !done && !lane64) {
                if (refine_solve(c, lu, f32, g, rhs, &j->prec, dx, rr, b32, x32) == 0) done = 1;
                else lane64 = 1;
            }
        }
        if (!singular && !done) {
            if (orc_lu_lane_refactor(lu, c->col_ptr, g, 1e-12, wide)) {
                repiv++;
                if (orc_lu_factor(lu, c->col_ptr, c->row_idx, g, j->q, 1e-3) != 0) singular = 1;
            }
            if (!singular) orc_lu_solve_neg(lu, rhs, dx);
        }
        if (singular) {
            reason = 3;
            iters = iter + 1;
            break;
        }
        /* updateAndNorm */
        uint64_t worst = 0;
        for (uint32_t i = 0; i < n; i++) {
            const double xo = x[i], d = dx[i], xn = xo + d;
            x_old[i] = xo;
            x[i] = xn;
            const double t = tol->reltol * fmax(fabs(xn), fabs(xo)) + (c->current_row[i] ? tol->abstol : tol->vntol);
            const int finite = fabs(xn) < INFINITY && fabs(d) < INFINITY;
            const double qq = fabs(d) / t;
            const double score = !finite ? INFINITY : (qq == qq ? qq : 0.0);
            uint64_t sb = db(score);
            if (sb > worst) worst = sb;
        }
        const double scaled = bd(worst);
        const int limited = apply_limits(c, &ld, x);
        int ok = !limited;
        if (ok && iter == k0) ok = 0;
        if (ok && !(scaled < 1.0)) ok = 0;
        if (ok)
            for (uint32_t i = 0; i < n; i++)
                if (!(fabs(x[i]) <= limit)) { ok = 0; break; }
        if (ok)
            for (uint32_t i = 0; i < n; i++) {
                const double s0 = fabs(sdiag[i]);
                const double s = s0 == 0 ? 1.0 : s0;
                const double rt = fmax(tol->residual_tol, 10.0 * s * (tol->reltol * fabs(x[i]) + tol->vntol));
                if (!(fabs(rhs[i]) <= rt)) { ok = 0; break; }
            }
        if (ok && init_fix) {
            init_fix = 0;
        } else if (ok) {
            memcpy(x, x_old, (size_t)n * 8);
            reason = 0;
            iters = iter + 1;
            break;
        }
    }
    if (reason == 1) iters = iter;
    const int ladder = (tol->reserved & ORC_TOL_LADDER) && (reason != 0 || (tol->reserved & ORC_TOL_LADDER_ALWAYS));
    if (ladder && j->prec.precision == 0 && !j->prec.nominal_x) {
        newton_ws w = {x_old, dx, g, rhs, sdiag, calloc(n + 1, 8)};
        reason = op_ladder(c, &ld, lu, j->q, tol, x, &w, iters, &iters, &repiv);
        free(w.x_good);
    }
    memcpy(j->x + l * n, x, (size_t)n * 8);
    j->iterations[l] = iters;
    if (j->fell_back) j->fell_back[l] = warm == 2;
    j->reasons[l] = reason;
    if (j->repivots) j->repivots[l] = repiv;
    if (j->fp64_lanes) j->fp64_lanes[l] = (uint32_t)(lane64 && j->prec.precision != 0);
    orc_lu_free(lu);
    orc_lu32_free(f32);
    free(x); free(x_old); free(dx); free(g); free(rhs); free(sdiag); free(rr); free(b32); free(x32);
    lane_dev_free(&ld);
}

static void *run_job(void *arg)
{
    job_t *j = arg;
    for (size_t l = j->first; l < j->first + j->count; l++) solve_lane(j, l);
    return NULL;
}

int orc_op_solve(orc_ckt *c, const double *values, size_t count, const orc_tol *tol,
                 const uint32_t *q, const double *start_x, int threads, double *x,
                 uint32_t *iterations, uint32_t *reasons, uint32_t *repivots)
{
    const orc_prec p64 = {0, 0, 0.0, NULL};
    return orc_op_solve2(c, values, count, tol, &p64, q, start_x, threads, x, iterations, reasons, repivots, NULL,
                         NULL);
}

int orc_op_solve2(orc_ckt *c, const double *values, size_t count, const orc_tol *tol, const orc_prec *prec,
                  const uint32_t *q, const double *start_x, int threads, double *x, uint32_t *iterations,
                  uint32_t *reasons, uint32_t *repivots, uint32_t *fp64_lanes, uint32_t *fell_back)
{
    if (count == 0) return 0;
    const uint32_t n = c->n, nnz = c->nnz;
    /* lead tape: point 0's first Jacobian */
    double *g = calloc(nnz + 1, 8), *rhs = calloc(n + 1, 8);
    if (start_x) orc_ckt_assemble_at(c, values, start_x, g, rhs);
    else orc_ckt_assemble_cold(c, values, g, rhs);
    orc_lu *lead = orc_lu_new(n, nnz);
    int rc = orc_lu_factor(lead, c->col_ptr, c->row_idx, g, q, 1e-3);
    free(g); free(rhs);
    if (rc != 0) {
        for (size_t l = 0; l < count; l++) { reasons[l] = 3; iterations[l] = 0; }
        orc_lu_free(lead);
        return -1;
    }
    if (threads < 1) threads = 1;
    if ((size_t)threads > count) threads = (int)count;
    job_t *jobs = calloc(threads, sizeof(job_t));
    pthread_t *th = calloc(threads, sizeof(pthread_t));
    size_t per = count / threads, rem = count % threads, at = 0;
    for (int t = 0; t < threads; t++) {
        size_t k = per + ((size_t)t < rem ? 1 : 0);
        jobs[t] = (job_t){c, values, at, k, tol, lead, q, start_x, x, iterations, reasons, repivots, *prec, fp64_lanes,
                          fell_back};
        at += k;
    }
    if (threads == 1) {
        run_job(&jobs[0]);
    } else {
        for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, run_job, &jobs[t]);
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    }
    free(jobs); free(th);
    orc_lu_free(lead);
    return 0;
}

/* ---- transient lanes (fixed step; the GPU lanes' esl_tran_solve) ---- */
typedef struct {
    orc_ckt *c;
    const double *values;
    size_t first, count;
    const orc_tol *tol;
    double dt;
    uint32_t steps, n_probes;
    const uint32_t *probes;
    const orc_lu *lead;            /* the transient lead tape */
    const uint32_t *q;
    const uint8_t *has_c;          /* [nnz] slots with charge-Jacobian entries */
    const double *x_op;            /* [count][n] operating points */
    const uint32_t *op_reason;
    double *x, *probe_out;
    uint32_t *iterations, *reasons;
    double *first_x, *first_lim;   /* traces (tests): [count][steps][n], [count][steps][n_inst*4] */
} tjob_t;

/* A = G + ag0 C on the charge slots, F += ag0 (q - q1) - ag1 i_prev */
static void companion(const orc_ckt *c, const uint8_t *has_c, double ag0, double ag1, double *g, const double *cp,
                      double *rhs, const double *q, const double *q1, const double *ip)
{
    for (uint32_t p = 0; p < c->nnz; p++)
        if (has_c[p]) g[p] = g[p] + ag0 * cp[p];
    for (uint32_t i = 0; i < c->n; i++) {
        const double d = ag0 * (q[i] - q1[i]);
        const double base = rhs[i] + d;
        rhs[i] = base - ag1 * ip[i];
    }
}

static orc_lu *lu_copy(const orc_lu *s, uint32_t n, uint32_t nnz)
{
    orc_lu *lu = orc_lu_new(n, nnz);
    memcpy(lu->q, s->q, n * 4); memcpy(lu->pinv, s->pinv, n * 4);
    memcpy(lu->prow, s->prow, nnz * 4); memcpy(lu->lp, s->lp, (n + 1) * 4);
    memcpy(lu->up, s->up, (n + 1) * 4);
    memcpy(lu->void_col, s->void_col, n); memcpy(lu->scaled, s->scaled, n);
    memcpy(lu->void_slots, s->void_slots, (size_t)s->n_void_slots * 4);
    lu->n_void_slots = s->n_void_slots;
    lu->cap_l = s->nnz_l + 16; lu->cap_u = s->nnz_u + 16;
    lu->li = realloc(lu->li, lu->cap_l * 4); lu->lx = realloc(lu->lx, lu->cap_l * 8);
    lu->ui = realloc(lu->ui, lu->cap_u * 4); lu->ux = realloc(lu->ux, lu->cap_u * 8);
    memcpy(lu->li, s->li, s->nnz_l * 4); memcpy(lu->ui, s->ui, s->nnz_u * 4);
    lu->nnz_l = s->nnz_l; lu->nnz_u = s->nnz_u;
    lu->factored = 1;
    return lu;
}

/* One transient step's Newton (the .op gate order with the residual gate
 * off: only a non-finite residual refuses) at source time t with companion
 * coefficients ag0, ag1: the replay, a re-pivot on a failed peel test,
 * dx = -A^-1 F, the delta and plausible gates and the limit pass. On
 * return x holds the published point (x_old of the converging iteration)
 * and q its charges; *iters counts the iterations. Returns 0 converged, 1 no
 * convergence within tol->max_iter, 3 singular. */
static uint32_t tran_newton(const orc_ckt *c, lane_dev *ld, orc_lu *lu, const uint32_t *qcol, const orc_tol *tol,
                            const uint8_t *has_c, double t, double ag0, double ag1, double *x, double *x_old,
                            double *dx, double *g, double *cp, double *rhs, double *q, const double *q1,
                            const double *ip, uint32_t *iters)
{
    const uint32_t n = c->n;
    const double wide = 1.0 / 1e-3;
    double limit = 1e9;
    uint32_t iter = 0;
    for (; iter < tol->max_iter; iter++) {
        assemble_t(c, ld, x, t, 1, g, rhs, cp, q);
        companion(c, has_c, ag0, ag1, g, cp, rhs, q, q1, ip);
        if (iter == 0) limit = plausible_limit(c, x, rhs, g);
        if (orc_lu_lane_refactor(lu, c->col_ptr, g, 1e-12, wide)) {
            if (orc_lu_factor(lu, c->col_ptr, c->row_idx, g, qcol, 1e-3) != 0) {
                *iters = iter;
                return 3;
            }
        }
        orc_lu_solve_neg(lu, rhs, dx);
        uint64_t worst = 0;
        for (uint32_t i = 0; i < n; i++) {
            const double xo = x[i], d = dx[i], xn = xo + d;
            x_old[i] = xo;
            x[i] = xn;
            const double tl = tol->reltol * fmax(fabs(xn), fabs(xo)) + (c->current_row[i] ? tol->abstol : tol->vntol);
            const int finite = fabs(xn) < INFINITY && fabs(d) < INFINITY;
            const double qq = fabs(d) / tl;
            const double score = !finite ? INFINITY : (qq == qq ? qq : 0.0);
            const uint64_t sb = db(score);
            if (sb > worst) worst = sb;
        }
        const double scaled = bd(worst);
        const int limited = apply_limits(c, ld, x);
        int ok = !limited && iter != 0 && scaled < 1.0;
        if (ok)
            for (uint32_t i = 0; i < n; i++)
                if (!(fabs(x[i]) <= limit) || !(fabs(rhs[i]) <= INFINITY)) { ok = 0; break; }
        if (ok) {
            memcpy(x, x_old, (size_t)n * 8);
            *iters = iter + 1;
            return 0;
        }
    }
    *iters = iter;
    return 1;
}

static void tran_lane(const tjob_t *j, size_t l)
{
    const orc_ckt *c = j->c;
    const uint32_t n = c->n, nnz = c->nnz;
    if (j->op_reason[l] != 0) {
        memcpy(j->x + l * n, j->x_op + l * n, (size_t)n * 8);
        j->iterations[l] = 0;
        j->reasons[l] = j->op_reason[l];
        return;
    }
    const orc_tol *tol = j->tol;
    lane_dev ld;
    lane_dev_init(c, j->values + l * c->d.n_live, &ld);
    double *x = calloc(n + 1, 8), *x_old = calloc(n + 1, 8), *dx = calloc(n + 1, 8);
    double *g = calloc(nnz + 1, 8), *cp = calloc(nnz + 1, 8), *rhs = calloc(n + 1, 8);
    double *q = calloc(n + 1, 8), *q1 = calloc(n + 1, 8), *ip = calloc(n + 1, 8), *xp = calloc(n + 1, 8);
    memcpy(x, j->x_op + l * n, (size_t)n * 8);
    memcpy(xp, x, (size_t)n * 8);
    seed_at(c, &ld, x);
    assemble_t(c, &ld, x, 0.0, 1, g, rhs, cp, q);
    memcpy(q1, q, (size_t)n * 8);
    orc_lu *lu = lu_copy(j->lead, n, nnz);
    uint32_t reason = 0, total = 0;
    for (uint32_t s = 0; s < j->steps && reason == 0; s++) {
        const double t = (double)(s + 1) * j->dt;
        const double ag0 = s == 0 ? 1.0 / j->dt : 2.0 / j->dt;
        const double ag1 = s == 0 ? 0.0 : 1.0;
        uint32_t iter = 0;
        reason = tran_newton(c, &ld, lu, j->q, tol, j->has_c, t, ag0, ag1, x, x_old, dx, g, cp, rhs, q, q1, ip,
                             &iter);
        total += iter;
        if (reason) break;
        for (uint32_t i = 0; i < n; i++) {
            const double d = ag0 * (q[i] - q1[i]);
            ip[i] = d - ag1 * ip[i];
            q1[i] = q[i];
        }
        for (uint32_t k = 0; k < j->n_probes; k++)
            j->probe_out[((size_t)l * j->steps + s) * j->n_probes + k] = x[j->probes[k]];
        /* the limiting state at the accepted point, MOS fetlim thresholds
         * from its last evaluation */
        seed_at_von(c, &ld, x, 1);
        /* the next step's first iterate: the last two accepted points
         * extrapolated by dt/dt_prev = 1 (MODEINITPRED, tran.zig:1131-1137,
         * predict at tran.zig:37-41), then one limit pass of the trial
         * against the accepted point (applyLimits(trial, cur),
         * tran.zig:1138; lane_device.zig:141-153 keeps x_old = cur for it) */
        if (s + 1 < j->steps) {
            if (!(tol->reserved & ORC_TOL_NO_PREDICT))
                for (uint32_t i = 0; i < n; i++) {
                    const double cu = x[i];
                    x[i] = cu + (cu - xp[i]);
                    xp[i] = cu;
                }
            apply_limits(c, &ld, x);
            if (j->first_x) {
                memcpy(j->first_x + ((size_t)l * j->steps + s + 1) * n, x, (size_t)n * 8);
                memcpy(j->first_lim + ((size_t)l * j->steps + s + 1) * c->n_inst * 4, ld.lim,
                       (size_t)c->n_inst * 4 * 8);
            }
        }
    }
    memcpy(j->x + l * n, x, (size_t)n * 8);
    j->iterations[l] = total;
    j->reasons[l] = reason;
    orc_lu_free(lu);
    free(x); free(x_old); free(dx); free(g); free(cp); free(rhs); free(q); free(q1); free(ip); free(xp);
    lane_dev_free(&ld);
}

/* ---- adaptive transient lanes: per-lane step control (tran.zig Sim.begin
 * and Sim.end, 728-1086, for the trapezoid method: orders 1 (backward
 * Euler) and 2 (trapezoid), the LTE bound of integrator.stepBound /
 * lane_kernels.zig:230-269 over the row plane, accept/reject, order
 * promotion, breakpoints of step sources; no device state flips, VPI or
 * delay lines, which the lane decks do not have) ---- */

/* stepBound (lane_kernels.zig:230-269): method 0 BE, 1 trap; cur_method the
 * method of the step being judged (its dynamic current) */
static double step_bound(int method, int cur_method, const double *q0, const double *q1, const double *q2,
                         const double *q3, const double *ip, uint32_t len, const orc_tran_adapt *o,
                         const orc_tol *tol, double dt, double dt1, double dt2, double ag0, double ag1)
{
    const double coeff = method == 0 ? 0.5 : 0.08333333333;
    const double sum01 = dt + dt1, sum12 = dt1 + dt2, sum012 = dt + (dt1 + dt2);
    double vmin = INFINITY;
    int poison = 0;
    for (uint32_t j = 0; j < len; j++) {
        const double qc = q0[j], qp = q1[j], qp2 = q2[j], qp3 = q3[j], i_p = ip[j];
        const double i_new = cur_method == 1 ? ag0 * (qc - qp) - ag1 * i_p : ag0 * (qc - qp);
        const double volttol = tol->abstol + tol->reltol * fmax(fabs(i_new), fabs(i_p));
        const double chargetol = tol->reltol * fmax(fmax(fabs(qc), fabs(qp)), o->chgtol) / dt;
        const double tl = fmax(volttol, chargetol);
        poison |= !isfinite(qc) || !isfinite(qp) || !isfinite(qp2) || !isfinite(i_p);
        const double f12 = (qp - qp2) / dt1;
        double dd = ((qc - qp) / dt - f12) / sum01;
        if (method != 0) {
            poison |= !isfinite(qp3);
            const double f123 = (f12 - (qp2 - qp3) / dt2) / sum12;
            dd = (dd - f123) / sum012;
        }
        const double del = o->trtol * tl / fmax(tol->abstol, coeff * fabs(dd));
        poison |= del != del;
        vmin = fmin(vmin, del);
    }
    if (poison) return NAN;
    return method == 0 ? vmin : sqrt(vmin);
}

/* the first step-source edge at or after t_from (ESL_V_WAVE 1 at td > 0),
 * or INFINITY */
static double next_break(const orc_ckt *c, const lane_dev *ld, double t_from)
{
    double bp = INFINITY;
    for (uint32_t i = 0; i < c->n_inst; i++) {
        const orc_inst *in = &c->inst[i];
        if (in->kind != K_V) continue;
        const double *p = ld->p + (size_t)in->dev * NPARAM;
        if (p[1] == 1.0 && p[3] > 0 && p[3] >= t_from && p[3] < bp) bp = p[3];
    }
    return bp;
}

static int almost_equal_ulps(double a, double b, int64_t ulps)
{
    if (a == b) return 1;
    if ((a < 0) != (b < 0)) return 0;
    const int64_t ia = (int64_t)db(fabs(a)), ib = (int64_t)db(fabs(b));
    const int64_t d = ia > ib ? ia - ib : ib - ia;
    return d <= ulps;
}

typedef struct {
    orc_ckt *c;
    const double *values;
    size_t first, count;
    const orc_tol *tol;
    const orc_tran_adapt *o;
    uint32_t n_probes;
    const uint32_t *probes;
    const orc_lu *lead;
    const uint32_t *q;
    const uint8_t *has_c;
    const double *x_op;
    const uint32_t *op_reason;
    int any_charge;
    double *x, *times, *probe_out;
    uint32_t *accepted, *rejected, *iterations, *reasons;
} tajob_t;

static void tran_adapt_lane(const tajob_t *j, size_t l)
{
    const orc_ckt *c = j->c;
    const orc_tran_adapt *o = j->o;
    const orc_tol *tol = j->tol;
    const uint32_t n = c->n, nnz = c->nnz;
    j->accepted[l] = 0;
    j->rejected[l] = 0;
    if (j->op_reason[l] != 0) {
        memcpy(j->x + l * n, j->x_op + l * n, (size_t)n * 8);
        j->iterations[l] = 0;
        j->reasons[l] = j->op_reason[l];
        return;
    }
    lane_dev ld;
    lane_dev_init(c, j->values + l * c->d.n_live, &ld);
    double *x = calloc(n + 1, 8), *x_old = calloc(n + 1, 8), *dx = calloc(n + 1, 8);
    double *g = calloc(nnz + 1, 8), *cp = calloc(nnz + 1, 8), *rhs = calloc(n + 1, 8);
    double *q = calloc(n + 1, 8), *ip = calloc(n + 1, 8);
    double *q1 = calloc(n + 1, 8), *q2 = calloc(n + 1, 8), *q3 = calloc(n + 1, 8);
    double *xc = calloc(n + 1, 8), *xp = calloc(n + 1, 8);
    memcpy(xc, j->x_op + l * n, (size_t)n * 8);
    memcpy(xp, xc, (size_t)n * 8);
    memcpy(x, xc, (size_t)n * 8);
    seed_at(c, &ld, x);
    /* the charge rings start full of the operating point's charges */
    assemble_t(c, &ld, x, 0.0, 1, g, rhs, cp, q);
    memcpy(q1, q, (size_t)n * 8);
    memcpy(q2, q, (size_t)n * 8);
    memcpy(q3, q, (size_t)n * 8);
    orc_lu *lu = lu_copy(j->lead, n, nnz);
    const double dt_max = o->dt_max, delmin = 1e-11 * dt_max, min_break = 10.0 * delmin;
    /* the first step (dctran.c:134, 578-586, firsttime /10) */
    double dt = fmin(fmin(o->t_step, o->t_stop / 100.0) / 10.0, dt_max);
    const double bp0 = next_break(c, &ld, min_break);
    if (bp0 < INFINITY) dt = fmin(dt, 0.1 * bp0);
    dt /= 10.0;
    double dt_old0 = dt_max, dt_old1 = dt_max, t = 0.0;
    double bp_next = INFINITY, bp_save_dt = 0.0;
    int bp_clamped = 0, order = 1;
    uint32_t steps = 0, rejects = 0, total = 0, reason = 0;
    for (;;) {
        if (!(t < o->t_stop && !almost_equal_ulps(t, o->t_stop, 100) && steps < o->max_steps)) break;
        const int method = order == 1 ? 0 : 1;
        const double ag0 = method == 0 ? 1.0 / dt : 1.0 / dt / (1.0 - 0.5);
        const double ag1 = method == 0 ? 0.0 : 0.5 / (1.0 - 0.5);
        /* the first iterate: the accepted points extrapolated by dt/dt_prev,
         * one limit pass against the accepted point */
        const double xfact = dt / dt_old0;
        seed_at_von(c, &ld, xc, 1);
        for (uint32_t i = 0; i < n; i++) x[i] = xc[i] + xfact * (xc[i] - xp[i]);
        apply_limits(c, &ld, x);
        uint32_t iters = 0;
        const uint32_t nr = tran_newton(c, &ld, lu, j->q, tol, j->has_c, t + dt, ag0, ag1, x, x_old, dx, g, cp, rhs,
                                        q, q1, ip, &iters);
        total += iters;
        if (nr == 3) { reason = 3; break; }
        if (nr != 0) {
            /* Newton failed: dt / 8 at order 1 (dctran.c:815, 823) */
            rejects++;
            order = 1;
            dt = dt / 8.0;
            if (dt < o->dt_min) { reason = 1; break; }
            continue;
        }
        double dt_next = steps == 0 ? dt : fmin(dt * 2.0, dt_max);
        if (j->any_charge && steps > 0) {
            const double del = step_bound(method, method, q, q1, q2, q3, ip, n, o, tol, dt, dt_old0, dt_old1, ag0, ag1);
            if (!(del >= 0.9 * dt)) {
                rejects++;
                if (del != del) {
                    order = 1;
                    dt = dt / 8.0;
                } else {
                    dt = del;
                }
                if (dt < o->dt_min) { reason = 1; break; }
                continue;
            }
            dt_next = fmin(fmin(fmax(del, o->dt_min), 2.0 * dt), dt_max);
        }
        /* accepted: order promotion (dctran.c:901-913) */
        if (j->any_charge && steps > 0 && order == 1) {
            const double trial = step_bound(1, method, q, q1, q2, q3, ip, n, o, tol, dt, dt_old0, dt_old1, ag0, ag1);
            if (trial == trial) {
                const double nd2 = fmin(2.0 * dt, trial);
                if (nd2 > 1.05 * dt) order = 2;
                dt_next = fmin(fmax(nd2, o->dt_min), dt_max);
            }
        }
        /* the dynamic current under the method used, then the ring */
        for (uint32_t i = 0; i < n; i++) {
            const double d = ag0 * (q[i] - q1[i]);
            ip[i] = d - ag1 * ip[i];
        }
        memcpy(q3, q2, (size_t)n * 8);
        memcpy(q2, q1, (size_t)n * 8);
        memcpy(q1, q, (size_t)n * 8);
        dt_old1 = dt_old0;
        dt_old0 = dt;
        memcpy(xp, xc, (size_t)n * 8);
        memcpy(xc, x, (size_t)n * 8);
        t = t + dt;
        if (j->times) j->times[(size_t)l * o->max_steps + steps] = t;
        for (uint32_t k = 0; k < j->n_probes; k++)
            j->probe_out[((size_t)l * o->max_steps + steps) * j->n_probes + k] = xc[j->probes[k]];
        steps++;
        /* landed on a breakpoint: back to order 1, resume at a tenth of the
         * gap (spice3 dctran's rule) */
        if (bp_next < INFINITY) {
            const int landed = bp_clamped ? fabs(t - bp_next) <= min_break
                                          : (bp_next - t <= delmin || almost_equal_ulps(t, bp_next, 100));
            if (landed) {
                order = 1;
                double shrink = bp_save_dt;
                const double nb = next_break(c, &ld, t + min_break);
                if (nb < INFINITY) shrink = fmin(shrink, nb - t);
                dt_next = fmin(dt_next, 0.1 * shrink);
            }
            bp_next = INFINITY;
        }
        bp_next = next_break(c, &ld, t + min_break);
        bp_clamped = 0;
        if (bp_next < INFINITY) {
            const double to_bp = bp_next - t;
            if (to_bp < dt_next) {
                bp_save_dt = dt_next;
                dt_next = to_bp;
                bp_clamped = 1;
            }
        }
        dt = dt_next;
        if (t + dt > o->t_stop) dt = o->t_stop - t;
    }
    if (reason == 0 && !(t >= o->t_stop || almost_equal_ulps(t, o->t_stop, 100))) reason = 1;   /* max_steps */
    memcpy(j->x + l * n, xc, (size_t)n * 8);
    j->iterations[l] = total;
    j->reasons[l] = reason;
    j->accepted[l] = steps;
    j->rejected[l] = rejects;
    orc_lu_free(lu);
    free(x); free(x_old); free(dx); free(g); free(cp); free(rhs); free(q); free(ip);
    free(q1); free(q2); free(q3); free(xc); free(xp);
    lane_dev_free(&ld);
}

static void *run_tajob(void *arg)
{
    tajob_t *j = arg;
    for (size_t l = j->first; l < j->first + j->count; l++) tran_adapt_lane(j, l);
    return NULL;
}

static void *run_tjob(void *arg)
{
    tjob_t *j = arg;
    for (size_t l = j->first; l < j->first + j->count; l++) tran_lane(j, l);
    return NULL;
}

int orc_tran_solve(orc_ckt *c, const double *values, size_t count, const orc_tol *tol_op, const orc_tol *tol,
                   double dt, uint32_t steps, const uint32_t *probes, uint32_t n_probes, const uint32_t *q,
                   int threads, double *x, double *probe_out, uint32_t *iterations, uint32_t *reasons)
{
    return orc_tran_solve_traced(c, values, count, tol_op, tol, dt, steps, probes, n_probes, q, threads, x,
                                 probe_out, iterations, reasons, NULL, NULL);
}

int orc_tran_solve_traced(orc_ckt *c, const double *values, size_t count, const orc_tol *tol_op,
                          const orc_tol *tol, double dt, uint32_t steps, const uint32_t *probes,
                          uint32_t n_probes, const uint32_t *q, int threads, double *x, double *probe_out,
                          uint32_t *iterations, uint32_t *reasons, double *first_x, double *first_lim)
{
    if (count == 0) return 0;
    const uint32_t n = c->n, nnz = c->nnz;
    double *x_op = calloc(count * n + 1, 8);
    uint32_t *op_it = calloc(count + 1, 4), *op_rs = calloc(count + 1, 4);
    orc_op_solve(c, values, count, tol_op, q, NULL, threads, x_op, op_it, op_rs, NULL);
    size_t l0 = count;
    for (size_t l = 0; l < count; l++)
        if (op_rs[l] == 0) { l0 = l; break; }
    uint8_t *has_c = calloc(nnz + 1, 1);
    for (uint32_t i = 0; i < c->n_inst; i++) {
        const orc_inst *in = &c->inst[i];
        if (in->kind != K_C && in->kind != K_D) continue;
        for (int e = 0; e < 4; e++)
            if (in->slot[e] != TRASH) has_c[in->slot[e]] = 1;
    }
    int rc = 0;
    orc_lu *lead = NULL;
    if (l0 < count) {
        /* the transient lead tape: point l0's first transient Jacobian */
        lane_dev ld;
        lane_dev_init(c, values + l0 * c->d.n_live, &ld);
        double *xx = malloc((size_t)(n + 1) * 8), *g = calloc(nnz + 1, 8), *cp = calloc(nnz + 1, 8);
        double *rhs = calloc(n + 1, 8), *qq = calloc(n + 1, 8), *q1 = calloc(n + 1, 8), *ip = calloc(n + 1, 8);
        memcpy(xx, x_op + l0 * n, (size_t)n * 8);
        seed_at(c, &ld, xx);
        assemble_t(c, &ld, xx, 0.0, 1, g, rhs, cp, qq);
        memcpy(q1, qq, (size_t)n * 8);
        assemble_t(c, &ld, xx, dt, 1, g, rhs, cp, qq);
        companion(c, has_c, 1.0 / dt, 0.0, g, cp, rhs, qq, q1, ip);
        lead = orc_lu_new(n, nnz);
        if (orc_lu_factor(lead, c->col_ptr, c->row_idx, g, q, 1e-3) != 0) {
            orc_lu_free(lead);
            lead = NULL;
        }
        free(xx); free(g); free(cp); free(rhs); free(qq); free(q1); free(ip);
        lane_dev_free(&ld);
    }
    if (!lead) {
        for (size_t l = 0; l < count; l++) {
            memcpy(x + l * n, x_op + l * n, (size_t)n * 8);
            iterations[l] = 0;
            reasons[l] = op_rs[l] != 0 ? op_rs[l] : 3;
        }
        rc = -1;
    } else {
        if (threads < 1) threads = 1;
        if ((size_t)threads > count) threads = (int)count;
        tjob_t *jobs = calloc(threads, sizeof(tjob_t));
        pthread_t *th = calloc(threads, sizeof(pthread_t));
        size_t per = count / threads, rem = count % threads, at = 0;
        for (int t = 0; t < threads; t++) {
            size_t k = per + ((size_t)t < rem ? 1 : 0);
            jobs[t] = (tjob_t){c, values, at, k, tol, dt, steps, n_probes, probes, lead, q, has_c, x_op, op_rs,
                               x, probe_out, iterations, reasons, first_x, first_lim};
            at += k;
        }
        if (threads == 1) run_tjob(&jobs[0]);
        else {
            for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, run_tjob, &jobs[t]);
            for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
        }
        free(jobs); free(th);
        orc_lu_free(lead);
    }
    free(x_op); free(op_it); free(op_rs); free(has_c);
    return rc;
}

int orc_tran_adaptive(orc_ckt *c, const double *values, size_t count, const orc_tol *tol_op, const orc_tol *tol,
                      const orc_tran_adapt *o, const uint32_t *probes, uint32_t n_probes, const uint32_t *q,
                      int threads, double *x, double *times, double *probe_out, uint32_t *accepted,
                      uint32_t *rejected, uint32_t *iterations, uint32_t *reasons)
{
    if (count == 0) return 0;
    const uint32_t n = c->n, nnz = c->nnz;
    double *x_op = calloc(count * n + 1, 8);
    uint32_t *op_it = calloc(count + 1, 4), *op_rs = calloc(count + 1, 4);
    orc_op_solve(c, values, count, tol_op, q, NULL, threads, x_op, op_it, op_rs, NULL);
    size_t l0 = count;
    for (size_t l = 0; l < count; l++)
        if (op_rs[l] == 0) { l0 = l; break; }
    uint8_t *has_c = calloc(nnz + 1, 1);
    int any_charge = 0;
    for (uint32_t i = 0; i < c->n_inst; i++) {
        const orc_inst *in = &c->inst[i];
        if (in->kind != K_C && in->kind != K_D) continue;
        for (int e = 0; e < 4; e++)
            if (in->slot[e] != TRASH) has_c[in->slot[e]] = 1;
    }
    for (uint32_t p = 0; p < nnz; p++) any_charge |= has_c[p];
    int rc = 0;
    orc_lu *lead = NULL;
    if (l0 < count) {
        /* the transient lead tape: point l0's first transient Jacobian (the
         * first step's dt, backward Euler) */
        lane_dev ld;
        lane_dev_init(c, values + l0 * c->d.n_live, &ld);
        const double dt_max = o->dt_max, min_break = 10.0 * 1e-11 * dt_max;
        double dt = fmin(fmin(o->t_step, o->t_stop / 100.0) / 10.0, dt_max);
        const double bp0 = next_break(c, &ld, min_break);
        if (bp0 < INFINITY) dt = fmin(dt, 0.1 * bp0);
        dt /= 10.0;
        double *xx = malloc((size_t)(n + 1) * 8), *g = calloc(nnz + 1, 8), *cp = calloc(nnz + 1, 8);
        double *rhs = calloc(n + 1, 8), *qq = calloc(n + 1, 8), *q1 = calloc(n + 1, 8), *ip = calloc(n + 1, 8);
        memcpy(xx, x_op + l0 * n, (size_t)n * 8);
        seed_at(c, &ld, xx);
        assemble_t(c, &ld, xx, 0.0, 1, g, rhs, cp, qq);
        memcpy(q1, qq, (size_t)n * 8);
        assemble_t(c, &ld, xx, dt, 1, g, rhs, cp, qq);
        companion(c, has_c, 1.0 / dt, 0.0, g, cp, rhs, qq, q1, ip);
        lead = orc_lu_new(n, nnz);
        if (orc_lu_factor(lead, c->col_ptr, c->row_idx, g, q, 1e-3) != 0) {
            orc_lu_free(lead);
            lead = NULL;
        }
        free(xx); free(g); free(cp); free(rhs); free(qq); free(q1); free(ip);
        lane_dev_free(&ld);
    }
    if (!lead) {
        for (size_t l = 0; l < count; l++) {
            memcpy(x + l * n, x_op + l * n, (size_t)n * 8);
            iterations[l] = 0;
            accepted[l] = rejected[l] = 0;
            reasons[l] = op_rs[l] != 0 ? op_rs[l] : 3;
        }
        rc = -1;
    } else {
        if (threads < 1) threads = 1;
        if ((size_t)threads > count) threads = (int)count;
        tajob_t *jobs = calloc(threads, sizeof(tajob_t));
        pthread_t *th = calloc(threads, sizeof(pthread_t));
        size_t per = count / threads, rem = count % threads, at = 0;
        for (int t = 0; t < threads; t++) {
            size_t k = per + ((size_t)t < rem ? 1 : 0);
            jobs[t] = (tajob_t){c, values, at, k, tol, o, n_probes, probes, lead, q, has_c, x_op, op_rs, any_charge,
                                x, times, probe_out, accepted, rejected, iterations, reasons};
            at += k;
        }
        if (threads == 1) run_tajob(&jobs[0]);
        else {
            for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, run_tajob, &jobs[t]);
            for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
        }
        free(jobs); free(th);
        orc_lu_free(lead);
    }
    free(x_op); free(op_it); free(op_rs); free(has_c);
    return rc;
}
