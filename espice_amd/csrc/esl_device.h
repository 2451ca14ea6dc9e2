// Device physics shared by the lane kernels (one body for every tiling):
// the ngspice limiters and the MOS1 DC evaluation of models/mos1.va.
// Compiled without floating-point contraction, like the oracle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "esl_internal.h"
#include "eslmath.h"

namespace esl {

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

__device__ inline Mos mos_load(const DeckProg &dp, uint32_t j, const double *live, uint32_t live_stride = 1)
{
    const uint32_t nm = dp.n_mos;
    const double *f = dp.m_par;
    Mos m;
    m.tp = f[MF_TYPE * nm + j];
    const int bv = dp.m_bind_vto[j], bk = dp.m_bind_kp[j];
    const double vtox = bv >= 0 ? live[(size_t)bv * live_stride] : f[MF_VTO * nm + j];
    const double kpx = bk >= 0 ? live[(size_t)bk * live_stride] : f[MF_KP * nm + j];
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

// Limit step after an update: (vgs, vds, vbs) are the last limited branch
// voltages on entry and the new ones on return; true when any value moved.
__device__ inline bool mos_limit_v(const Mos &m, double xd, double xg, double xs, double xb, double &lgs,
                                   double &lds, double &lbs, double von)
{
    const double tp = m.tp;
    const double cgs = tp * (xg - xs), cds = tp * (xd - xs), cbs = tp * (xb - xs);
    const double ogs = lgs, ods = lds, obs = lbs;
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
    lgs = vgs;
    lds = vds;
    lbs = vbs;
    return vgs != cgs || vds != cds || vbs != cbs;
}

// DC evaluation at the limited voltages (vgs, vds, vbs) plus the residual
// correction toward the probe x; out[0..10] are the Jacobian entries in the
// order (d,d) (d,g) (d,s) (d,b) (s,d) (s,g) (s,s) (s,b) (b,d) (b,s) (b,b) and
// out[11..13] the residuals of rows d, s, b. Returns von for the next limit.
__device__ inline double mos_eval_v(const Mos &m, double xd, double xg, double xs, double xb, double vgs, double vds,
                                    double vbs, double *out)
{
    const double tp = m.tp;
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
    out[0] = cd + gbd;
    out[1] = cg;
    out[2] = cs;
    out[3] = cb - gbd;
    out[4] = -cd;
    out[5] = -cg;
    out[6] = -cs + gbs;
    out[7] = -cb - gbs;
    out[8] = -gbd;
    out[9] = -gbs;
    out[10] = gbd + gbs;
    out[11] = ich - ib_d;
    out[12] = -ich - ib_s;
    out[13] = ib_d + ib_s;
    return von;
}


// ---- junction diode (models/diode.va, the RS = 0 subset without breakdown,
// recombination, knee currents or sidewall) ----
struct Dio {
    double is, vte, vcrit, gmin, cjo, vj, mt, qexp, vfc, f1, inv_f2, f3, tt;
    double bv, xbv, vtebrk;       // breakdown, armed when bv > 0
    double ikf;                   // high-injection knee IKF * area, 0 off
    double isr, vterec;           // recombination ISR(T) and NR vt, isr 0 off
};

__device__ inline Dio dio_load(const DeckProg &dp, uint32_t j, const double *live, uint32_t live_stride = 1)
{
    const uint32_t nd = dp.n_dio;
    const double *f = dp.d_par;
    Dio d;
    const int b = dp.d_bind_is[j];
    const double isx = b >= 0 ? live[(size_t)b * live_stride] : f[DF_IS * nd + j];
    d.is = isx * f[DF_IS_SCALE * nd + j];
    d.vte = f[DF_VTE * nd + j];
    d.vcrit = d.vte * esl::log(d.vte / (1.4142135623730951 * d.is));
    d.gmin = f[DF_GMIN * nd + j];
    d.cjo = f[DF_CJO * nd + j];
    d.vj = f[DF_VJ * nd + j];
    d.mt = f[DF_MT * nd + j];
    d.qexp = f[DF_QEXP * nd + j];
    d.vfc = f[DF_VFC * nd + j];
    d.f1 = f[DF_F1 * nd + j];
    d.inv_f2 = f[DF_INVF2 * nd + j];
    d.f3 = f[DF_F3 * nd + j];
    d.tt = f[DF_TT * nd + j];
    // the breakdown knee (diotemp.c): where -is(T) exp(-(xbv + vd) / vtebrk)
    // carries IBV at vd = -BV; it moves with a per-point IS
    d.bv = f[DF_BV * nd + j];
    d.vtebrk = f[DF_VTEBRK * nd + j];
    d.xbv = d.bv;
    const double ibv = f[DF_IBV * nd + j], vt = f[DF_VT * nd + j];
    if (d.bv > 0.0 && ibv >= d.is * d.bv / vt) {
        double xbv = d.bv - d.vtebrk * esl::log(1.0 + ibv / d.is);
        xbv = d.bv - d.vtebrk * esl::log(ibv / d.is + 1.0 - xbv / vt);
        d.xbv = xbv;
    }
    d.ikf = f[DF_IKF * nd + j];
    d.isr = f[DF_ISR * nd + j];
    d.vterec = f[DF_VTEREC * nd + j];
    return d;
}

// I = ihi + gmin vd (diode.va:264-314): the junction current with the 80
// vte tangent and the cubic reverse tail, the ISR/NR recombination current
// in forward bias, the breakdown branch past -xbv, the IKF knee roll-off;
// g = dI/dvd.
__device__ inline double dio_current(const Dio &d, double vd, double &g)
{
    double id, gd;
    if (vd >= -3.0 * d.vte) {
        const double u = vd / d.vte;
        const double ul = fmin(u, 80.0);
        const double e = esl::exp(ul);
        id = d.is * (e * (1.0 + (u - ul)) - 1.0);
        gd = d.is * e / d.vte;
    } else {
        const double w = 3.0 * d.vte / (vd * 2.718281828459045);
        const double w3 = w * w * w;
        id = -d.is * (1.0 + w3);
        gd = 3.0 * d.is * w3 / vd;
    }
    double irec = 0.0, grec = 0.0;
    if (d.isr > 0.0 && vd >= -3.0 * d.vte) {
        const double ur = vd / d.vterec;
        const double url = fmin(ur, 80.0);
        const double er = esl::exp(url);
        const double a = 1.0 - vd / d.vj;
        const double gg = a * a + 0.005;
        const double genfac = esl::exp((d.mt / 2.0) * esl::log(gg));
        const double base = er * (1.0 + (ur - url)) - 1.0;
        irec = d.isr * base * genfac;
        const double dgen = genfac * (d.mt / 2.0) * (-2.0 * a / d.vj) / gg;
        grec = d.isr * (er / d.vterec * genfac + base * dgen);
    }
    if (d.bv > 0.0 && vd < -3.0 * d.vte && vd < -d.xbv) {
        const double wb = -(d.xbv + vd) / d.vtebrk;
        const double wl = fmin(wb, 80.0);
        const double e = esl::exp(wl);
        id = -d.is * e * (1.0 + (wb - wl));
        gd = d.is * e / d.vtebrk;
    }
    double ihi = id + irec, ghi = gd + grec;
    if (d.ikf > 0.0 && ihi > 0.0) {
        // ihi / (1 + s), s = sqrt(ihi / ikf): d/dihi = (1 + s/2) / (1 + s)^2
        const double sq = sqrt(ihi / d.ikf);
        const double den = 1.0 + sq;
        ghi = ghi * ((1.0 + 0.5 * sq) / (den * den));
        ihi = ihi / den;
    }
    g = ghi + d.gmin;
    return ihi + d.gmin * vd;
}

// Junction charge (depletion, two-branch SPICE form) plus tt * I; c = dq/dvd.
__device__ inline double dio_charge(const Dio &d, double vd, double i, double g, double &c)
{
    double q = 0.0, cap = 0.0;
    if (d.cjo > 0) {
        if (vd < d.vfc) {
            const double a = 1.0 - vd / d.vj;
            const double p = esl::exp(d.qexp * esl::log(a));
            q = d.cjo * d.vj * (1.0 - p) / d.qexp;
            cap = d.cjo * p / a;
        } else {
            q = d.cjo * (d.f1 + d.inv_f2 * (d.f3 * (vd - d.vfc) + (d.mt / (2.0 * d.vj)) * (vd * vd - d.vfc * d.vfc)));
            cap = d.cjo * d.inv_f2 * (d.f3 + (d.mt / d.vj) * vd);
        }
    }
    if (d.tt > 0) {
        q = q + d.tt * i;
        cap = cap + d.tt * g;
    }
    c = cap;
    return q;
}

// pnjlim of the junction voltage against the previous limited value, with
// the same snap to the probe as the MOS limiter; true when it moved.
__device__ inline bool dio_limit_v(const Dio &d, double xa, double xk, double &lvd)
{
    const double vd = xa - xk;
    double v = pnjlim(vd, lvd, d.vte, d.vcrit);
    const double tol = 16.0 * 2.220446049250313e-16 * fmax(fabs(xa), fabs(xk));
    if (fabs(v - vd) <= tol) v = vd;
    lvd = v;
    return v != vd;
}

// Stamps at the limited voltage vl with the residual correction toward the
// probe: out[0..3] Jacobian (a,a) (a,k) (k,a) (k,k), out[4..5] residual rows
// a, k; with charges, out[6..9] the charge Jacobian and out[10..11] the row
// charges, corrected the same way.
__device__ inline void dio_eval_v(const Dio &d, double xa, double xk, double vl, bool charges, double *out)
{
    double g;
    double i = dio_current(d, vl, g);
    const double dv = (xa - xk) - vl;
    double q = 0.0, c = 0.0;
    if (charges) q = dio_charge(d, vl, i, g, c);
    if (dv != 0) {
        i = i + g * dv;
        if (charges) q = q + c * dv;
    }
    out[0] = g;
    out[1] = -g;
    out[2] = -g;
    out[3] = g;
    out[4] = i;
    out[5] = -i;
    if (charges) {
        out[6] = c;
        out[7] = -c;
        out[8] = -c;
        out[9] = c;
        out[10] = q;
        out[11] = -q;
    }
}

// ---- Gummel-Poon BJT (models/gummel_poon.va, DC part with RB = RC = RE = 0,
// no excess phase; vertical substrate) ----
struct Bjt {
    double tp, is, bf, br, vt, nfvt, nrvt, ise, nevt, isc, ncvt, iss, nsvt, inv_vaf, inv_var, inv_ikf, inv_ikr, nkf,
        mgmin, vert, vcrit;
};

__device__ inline Bjt bjt_load(const DeckProg &dp, uint32_t j, const double *live, uint32_t live_stride = 1)
{
    const uint32_t nq = dp.n_bjt;
    const double *f = dp.q_par;
    Bjt b;
    b.tp = f[QF_TYPE * nq + j];
    const int bi = dp.q_bind_is[j], bb = dp.q_bind_bf[j];
    const double isx = bi >= 0 ? live[(size_t)bi * live_stride] : f[QF_IS * nq + j];
    const double bfx = bb >= 0 ? live[(size_t)bb * live_stride] : f[QF_BF * nq + j];
    b.is = isx * f[QF_IS_SCALE * nq + j];
    b.bf = bfx * f[QF_BFAC * nq + j];
    b.br = f[QF_BR_T * nq + j];
    b.vt = f[QF_VT * nq + j];
    b.nfvt = f[QF_NFVT * nq + j];
    b.nrvt = f[QF_NRVT * nq + j];
    b.ise = f[QF_ISE * nq + j];
    b.nevt = f[QF_NEVT * nq + j];
    b.isc = f[QF_ISC * nq + j];
    b.ncvt = f[QF_NCVT * nq + j];
    b.iss = f[QF_ISS * nq + j];
    b.nsvt = f[QF_NSVT * nq + j];
    b.inv_vaf = f[QF_INV_VAF * nq + j];
    b.inv_var = f[QF_INV_VAR * nq + j];
    b.inv_ikf = f[QF_INV_IKF * nq + j];
    b.inv_ikr = f[QF_INV_IKR * nq + j];
    b.nkf = f[QF_NKF * nq + j];
    b.mgmin = f[QF_MGMIN * nq + j];
    b.vert = f[QF_VERT * nq + j];
    b.vcrit = b.vt * esl::log(b.vt / (1.4142135623730951 * b.is));
    return b;
}

// Junction law with the 80 vtn tangent and the cubic reverse tail (the
// model's pnj); g = di/dv.
__device__ inline double bjt_pnj(double v, double isat, double vtn, double &g)
{
    if (v >= -3.0 * vtn) {
        const double u = fmin(v / vtn, 80.0);
        const double e = esl::exp(u);
        g = isat * e / vtn;
        return isat * (e * (1.0 + (v / vtn - u)) - 1.0);
    }
    const double u = 3.0 * vtn / (2.718281828459045 * v);
    const double u3 = u * u * u;
    g = 3.0 * isat * u3 / v;
    return -isat * (1.0 + u3);
}

// pnjlim of vbe and vbc (npn frame) against the previous limited values,
// with the snap to the probe; true when either moved.
__device__ inline bool bjt_limit_v(const Bjt &b, double xc, double xb, double xe, double &lbe, double &lbc)
{
    const double vbe = b.tp * (xb - xe), vbc = b.tp * (xb - xc);
    double ve = pnjlim(vbe, lbe, b.vt, b.vcrit);
    double vc = pnjlim(vbc, lbc, b.vt, b.vcrit);
    const double tol = 16.0 * 2.220446049250313e-16 * fmax(fmax(fabs(xc), fabs(xb)), fabs(xe));
    if (fabs(ve - vbe) <= tol) ve = vbe;
    if (fabs(vc - vbc) <= tol) vc = vbc;
    lbe = ve;
    lbc = vc;
    return ve != vbe || vc != vbc;
}

// DC stamps at the limited (vbe, vbc) with the residual correction toward
// the probe. out[0..11]: Jacobian (c,c) (c,b) (c,e) (c,s) (b,c) (b,b) (b,e)
// (e,c) (e,b) (e,e) (s,c) (s,s); out[12..15]: residual rows c, b, e, s.
__device__ inline void bjt_eval_v(const Bjt &b, double xc, double xb, double xe, double xs, double vbe, double vbc,
                                  double *out)
{
    const double tp = b.tp;
    const double dbe = tp * (xb - xe) - vbe, dbc = tp * (xb - xc) - vbc;
    const double vsub = tp * (xs - xc);
    double gbe, gben, gbc, gbcn, gsub;
    const double cbe = bjt_pnj(vbe, b.is, b.nfvt, gbe);
    const double cben = bjt_pnj(vbe, b.ise, b.nevt, gben) + b.mgmin * vbe;
    gben = gben + b.mgmin;
    const double cbc = bjt_pnj(vbc, b.is, b.nrvt, gbc);
    const double cbcn = bjt_pnj(vbc, b.isc, b.ncvt, gbcn) + b.mgmin * vbc;
    gbcn = gbcn + b.mgmin;
    const double cdsub = bjt_pnj(vsub, b.iss, b.nsvt, gsub) + b.mgmin * vsub;
    gsub = gsub + b.mgmin;
    const double q1 = 1.0 / (1.0 - vbc * b.inv_vaf - vbe * b.inv_var);
    const double q2 = cbe * b.inv_ikf + cbc * b.inv_ikr;
    const double arg = fmax(1e-30, 1.0 + 4.0 * q2);
    const double p = b.nkf == 0.5 ? sqrt(arg) : esl::exp(b.nkf * esl::log(arg));
    const double qb = 0.5 * q1 * (1.0 + p);
    const double dp_dq2 = (1.0 + 4.0 * q2 > 1e-30) ? 4.0 * b.nkf * p / arg : 0.0;
    const double dqb_be = 0.5 * (q1 * q1 * b.inv_var * (1.0 + p) + q1 * dp_dq2 * (gbe * b.inv_ikf));
    const double dqb_bc = 0.5 * (q1 * q1 * b.inv_vaf * (1.0 + p) + q1 * dp_dq2 * (gbc * b.inv_ikr));
    double icc = cbe / qb - cbc / qb;
    const double dcc = (cbe - cbc) / (qb * qb);
    const double dicc_be = gbe / qb - dcc * dqb_be;
    const double dicc_bc = -gbc / qb - dcc * dqb_bc;
    double ibe = cbe / b.bf + cben;
    const double gbe_t = gbe / b.bf + gben;
    double ibc = cbc / b.br + cbcn;
    const double gbc_t = gbc / b.br + gbcn;
    const double isc_ = b.vert * cdsub;
    const double gs = b.vert * gsub;
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

}  // namespace esl
