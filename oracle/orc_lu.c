/* Oracle sparse LU: a restatement of ESPice's SparseLu (left-looking
 * Gilbert-Peierls with threshold partial pivoting), its numeric replay and
 * its solves, and the lane replay with its peel tests.
 *
 *   orc_lu_factor        sparse_lu.zig:260-288 (factor), 313-555 (factorPass,
 *                        the plain column loop; the supernodal copy is bitwise
 *                        the same per sparse_lu.zig:772-776), 847-871
 *                        (voidUnknown)
 *   orc_lu_refactor      sparse_lu.zig:916-1024 (refactorColumns)
 *   orc_lu_lane_refactor lane_kernels.zig:49-118 (Bodies.refactor + the
 *                        wideMultipliers test)
 *   orc_lu_solve         sparse_lu.zig:1140-1184 (solve, column path)
 *   orc_lu_solve_neg     lane_kernels.zig:123-143 (Bodies.solveNeg)
 *
 * Parity: pinned by the closed forms of direct.zig:655-724, dense Gaussian
 * elimination on random MNA-like systems (solver/tests.zig:37-62,
 * sparse_lu.zig:1234-1311), replay == factor bitwise on the values that chose
 * the pivots, the w-all-zero invariant, and power-of-two scaling exactness
 * (sparse_lu.zig:1468), all in tests/test_oracle_lu.py. The reference's exact
 * pivot sequence on a given matrix depends on its AMD ordering, which the
 * caller here supplies instead: that sequence is parity unpinned.
 *
 * Compiled with -ffp-contract=off: every multiply-subtract rounds twice, as
 * the reference's bodies do (lu_kernels.zig:1-6).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

static uint64_t dbits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }
static double bitsd(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }

orc_lu *orc_lu_new(uint32_t n, uint32_t nnz)
{
    orc_lu *lu = calloc(1, sizeof(orc_lu));
    lu->n = n;
    lu->nnz = nnz;
    lu->q = calloc(n + 1, 4);
    lu->pinv = calloc(n + 1, 4);
    lu->prow = calloc(nnz + 1, 4);
    lu->lp = calloc(n + 1, 4);
    lu->up = calloc(n + 1, 4);
    lu->udiag = calloc(n + 1, 8);
    lu->void_col = calloc(n + 1, 1);
    lu->scaled = calloc(n + 1, 1);
    lu->void_slots = calloc(nnz + 1, 4);
    lu->w = calloc(n + 1, 8);
    lu->y = calloc(n + 1, 8);
    lu->cap_l = lu->cap_u = (nnz > 2 * n ? nnz : 2 * n) + 16;
    lu->li = malloc((size_t)lu->cap_l * 4);
    lu->lx = malloc((size_t)lu->cap_l * 8);
    lu->ui = malloc((size_t)lu->cap_u * 4);
    lu->ux = malloc((size_t)lu->cap_u * 8);
    return lu;
}

void orc_lu_free(orc_lu *lu)
{
    if (!lu) return;
    free(lu->q); free(lu->pinv); free(lu->prow); free(lu->lp); free(lu->up);
    free(lu->udiag); free(lu->void_col); free(lu->scaled); free(lu->void_slots);
    free(lu->w); free(lu->y); free(lu->li); free(lu->lx); free(lu->ui); free(lu->ux);
    free(lu);
}

static void grow_l(orc_lu *lu, uint32_t need)
{
    if (need <= lu->cap_l) return;
    uint32_t cap = lu->cap_l;
    while (cap < need) cap = cap * 2;
    lu->li = realloc(lu->li, (size_t)cap * 4);
    lu->lx = realloc(lu->lx, (size_t)cap * 8);
    lu->cap_l = cap;
}

static void grow_u(orc_lu *lu, uint32_t need)
{
    if (need <= lu->cap_u) return;
    uint32_t cap = lu->cap_u;
    while (cap < need) cap = cap * 2;
    lu->ui = realloc(lu->ui, (size_t)cap * 4);
    lu->ux = realloc(lu->ux, (size_t)cap * 8);
    lu->cap_u = cap;
}

/* voidUnknown (sparse_lu.zig:847-871). */
static int void_unknown(const orc_lu *lu, const uint32_t *col_ptr, const uint32_t *row_idx,
                        const double *vals, uint32_t c)
{
    if (lu->pinv[c] != ORC_NONE) return 0;
    for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++)
        if (vals[p] != 0) return 0;
    for (uint32_t j = 0; j < lu->n; j++) {
        if (j == c) continue;
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++)
            if (row_idx[p] == c && vals[p] != 0) return 0;
    }
    return 1;
}

int orc_lu_factor(orc_lu *lu, const uint32_t *col_ptr, const uint32_t *row_idx,
                  const double *vals, const uint32_t *q, double pivot_tol)
{
    const uint32_t n = lu->n, nnz = col_ptr[n];
    memcpy(lu->q, q, (size_t)n * 4);
    lu->factored = 0;
    for (uint32_t i = 0; i < n; i++) lu->pinv[i] = ORC_NONE;
    memset(lu->void_col, 0, n);
    memset(lu->scaled, 0, n);
    lu->nnz_l = lu->nnz_u = lu->n_void_slots = 0;
    int has_void = 0;

    uint32_t *flag = calloc(n, 4), *topo = malloc((size_t)n * 4);
    uint32_t *stack = malloc((size_t)n * 4), *pstack = malloc((size_t)n * 4);
    uint32_t *want = malloc((size_t)n * 4), *want_of = malloc((size_t)n * 4);
    double *rscale = calloc(n, 8);
    double *w = lu->w;
    int rc = 0;

    /* A non-finite entry fails the factor up front (normInf test). */
    uint64_t amaxb = 0;
    for (uint32_t p = 0; p < nnz; p++) {
        uint64_t bb = dbits(fabs(vals[p]));
        if (bb > amaxb) amaxb = bb;
    }
    if (!(bitsd(amaxb) <= 1.7976931348623157e308)) { rc = -1; goto done; }
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++)
            rscale[row_idx[p]] = fmax(rscale[row_idx[p]], fabs(vals[p]));
    for (uint32_t i = 0; i < n; i++)
        rscale[i] = (rscale[i] > 0 && isfinite(rscale[i])) ? 1.0 / rscale[i] : 1.0;

    uint64_t nonfinite = 0;
    for (uint32_t i = 0; i < n; i++) { want[i] = i; want_of[i] = i; }

    for (uint32_t k = 0; k < n; k++) {
        const uint32_t c = q[k];
        lu->lp[k] = lu->nnz_l;
        lu->up[k] = lu->nnz_u;
        const uint32_t mark = k + 1;
        /* Reach of A[:,c] through the graph of L (rows still original). */
        uint32_t nt = 0;
        for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) {
            uint32_t r = row_idx[p];
            if (flag[r] == mark) continue;
            uint32_t sp = 0;
            stack[0] = r;
            pstack[0] = lu->pinv[r] == ORC_NONE ? 0 : lu->lp[lu->pinv[r]];
            flag[r] = mark;
            for (;;) {
                r = stack[sp];
                uint32_t kc = lu->pinv[r];
                uint32_t end = kc == ORC_NONE ? 0 : lu->lp[kc + 1];
                uint32_t pos = pstack[sp];
                int descended = 0;
                while (pos < end) {
                    uint32_t child = lu->li[pos];
                    pos++;
                    if (flag[child] != mark) {
                        pstack[sp] = pos;
                        flag[child] = mark;
                        sp++;
                        stack[sp] = child;
                        pstack[sp] = lu->pinv[child] == ORC_NONE ? 0 : lu->lp[lu->pinv[child]];
                        descended = 1;
                        break;
                    }
                }
                if (descended) continue;
                topo[nt++] = r;
                if (sp == 0) break;
                sp--;
            }
        }
        grow_u(lu, lu->nnz_u + nt);
        grow_l(lu, lu->nnz_l + nt);

        for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) w[row_idx[p]] = vals[p];
        /* Triangular solve in reverse finish order. */
        for (uint32_t idx = nt; idx > 0;) {
            idx--;
            uint32_t r = topo[idx], kc = lu->pinv[r];
            if (kc == ORC_NONE) continue;
            double ukr = w[r];
            lu->ui[lu->nnz_u] = kc;
            lu->ux[lu->nnz_u] = ukr;
            lu->nnz_u++;
            for (uint32_t pl = lu->lp[kc]; pl < lu->lp[kc + 1]; pl++)
                w[lu->li[pl]] -= lu->lx[pl] * ukr;
        }
        double amax = 0, smax = 0;
        uint32_t piv = ORC_NONE;
        for (uint32_t t = 0; t < nt; t++) {
            uint32_t r = topo[t];
            if (lu->pinv[r] != ORC_NONE) continue;
            double a = fabs(w[r]);
            if (a > amax) { amax = a; piv = r; }
            smax = fmax(smax, a * rscale[r]);
        }
        if (piv == ORC_NONE || amax == 0 || !isfinite(amax)) {
            if (!void_unknown(lu, col_ptr, row_idx, vals, c)) {
                for (uint32_t t = 0; t < nt; t++) w[topo[t]] = 0;
                rc = -1;
                goto done;
            }
            lu->void_col[k] = 1;
            has_void = 1;
            lu->udiag[k] = 1;
            lu->pinv[c] = k;
            for (uint32_t t = 0; t < nt; t++) w[topo[t]] = 0;
            continue;
        }
        const uint32_t pr = want[c];
        if (lu->pinv[pr] == ORC_NONE) {
            double dmag = fabs(w[pr]);
            if (dmag > 0 && dmag >= pivot_tol * amax) {
                piv = pr;
            } else if (dmag > 0 && dmag * rscale[pr] >= pivot_tol * smax) {
                piv = pr;
                lu->scaled[k] = 1;
            }
        }
        if (piv != pr) {
            uint32_t c2 = want_of[piv];
            want[c2] = pr;
            want_of[pr] = c2;
        }
        const double d = w[piv];
        lu->udiag[k] = d;
        lu->pinv[piv] = k;
        for (uint32_t t = 0; t < nt; t++) {
            uint32_t r = topo[t];
            double v = w[r];
            nonfinite |= dbits(v - v);
            if (lu->pinv[r] == ORC_NONE) {
                double l = v / d;
                nonfinite |= dbits(l - l);
                lu->li[lu->nnz_l] = r;
                lu->lx[lu->nnz_l] = l;
                lu->nnz_l++;
            }
            w[r] = 0;
        }
    }
    lu->lp[n] = lu->nnz_l;
    lu->up[n] = lu->nnz_u;
    if (nonfinite != 0) { rc = -1; goto done; }
    for (uint32_t p = 0; p < lu->nnz_l; p++) lu->li[p] = lu->pinv[lu->li[p]];
    for (uint32_t p = 0; p < nnz; p++) lu->prow[p] = lu->pinv[row_idx[p]];
    if (has_void) {
        for (uint32_t j = 0; j < n; j++) {
            uint32_t c = q[j];
            for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++)
                if (lu->void_col[j] || lu->void_col[lu->prow[p]])
                    lu->void_slots[lu->n_void_slots++] = p;
        }
    }
    lu->factored = 1;
done:
    free(flag); free(topo); free(stack); free(pstack); free(want); free(want_of); free(rscale);
    return rc;
}

int orc_lu_refactor(orc_lu *lu, const uint32_t *col_ptr, const double *vals, double growth_limit)
{
    const uint32_t n = lu->n;
    double *w = lu->w;
    for (uint32_t v = 0; v < lu->n_void_slots; v++)
        if (vals[lu->void_slots[v]] != 0) return -1;
    uint64_t nonfinite = 0;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t c = lu->q[k];
        const uint32_t lk0 = lu->lp[k], lk1 = lu->lp[k + 1];
        for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) w[lu->prow[p]] = vals[p];
        for (uint32_t p = lu->up[k]; p < lu->up[k + 1]; p++) {
            uint32_t i = lu->ui[p];
            double uki = w[i];
            w[i] = 0;
            lu->ux[p] = uki;
            nonfinite |= dbits(uki - uki);
            for (uint32_t pl = lu->lp[i]; pl < lu->lp[i + 1]; pl++)
                w[lu->li[pl]] -= lu->lx[pl] * uki;
        }
        if (lu->void_col[k]) {
            lu->udiag[k] = 1;
            for (uint32_t p = lk0; p < lk1; p++) { lu->lx[p] = 0; w[lu->li[p]] = 0; }
            w[k] = 0;
            for (uint32_t p = lu->up[k]; p < lu->up[k + 1]; p++) {
                uint32_t i = lu->ui[p];
                for (uint32_t pl = lu->lp[i]; pl < lu->lp[i + 1]; pl++) w[lu->li[pl]] = 0;
            }
            continue;
        }
        const double d = w[k];
        w[k] = 0;
        if (d == 0 || !isfinite(d)) {
            for (uint32_t p = lk0; p < lk1; p++) w[lu->li[p]] = 0;
            return -1;
        }
        lu->udiag[k] = d;
        if (growth_limit > 0 && !lu->scaled[k]) {
            int64_t cbits = (int64_t)dbits(fabs(d));
            for (uint32_t p = lk0; p < lk1; p++) {
                uint32_t r = lu->li[p];
                double v = w[r];
                w[r] = 0;
                int64_t vb = (int64_t)dbits(fabs(v));
                if (vb > cbits) cbits = vb;
                lu->lx[p] = v / d;
            }
            double cmax = bitsd((uint64_t)cbits);
            if (!(fabs(d) >= growth_limit * cmax)) {
                /* w is zero here: every L row of this column was cleared
                 * and later rows are untouched yet. */
                return -1;
            }
        } else {
            for (uint32_t p = lk0; p < lk1; p++) {
                uint32_t r = lu->li[p];
                double l = w[r] / d;
                nonfinite |= dbits(l - l);
                lu->lx[p] = l;
                w[r] = 0;
            }
        }
    }
    return nonfinite != 0 ? -1 : 0;
}

int orc_lu_lane_refactor(orc_lu *lu, const uint32_t *col_ptr, const double *vals,
                         double growth_limit, double wide)
{
    const uint32_t n = lu->n;
    double *w = lu->w;
    int bad = 0;
    uint64_t nonfinite = 0;
    for (uint32_t v = 0; v < lu->n_void_slots; v++)
        if (vals[lu->void_slots[v]] != 0) return 1;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t c = lu->q[k];
        for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) w[lu->prow[p]] = vals[p];
        for (uint32_t p = lu->up[k]; p < lu->up[k + 1]; p++) {
            uint32_t i = lu->ui[p];
            double uki = w[i];
            w[i] = 0;
            lu->ux[p] = uki;
            nonfinite |= dbits(uki - uki);
            for (uint32_t pl = lu->lp[i]; pl < lu->lp[i + 1]; pl++)
                w[lu->li[pl]] -= lu->lx[pl] * uki;
        }
        if (lu->void_col[k]) {
            lu->udiag[k] = 1;
            for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) { lu->lx[p] = 0; w[lu->li[p]] = 0; }
            w[k] = 0;
            for (uint32_t p = lu->up[k]; p < lu->up[k + 1]; p++) {
                uint32_t i = lu->ui[p];
                for (uint32_t pl = lu->lp[i]; pl < lu->lp[i + 1]; pl++) w[lu->li[pl]] = 0;
            }
            continue;
        }
        double d = w[k];
        w[k] = 0;
        int dead = d == 0 || !(fabs(d) < INFINITY);
        bad = bad || dead;
        if (dead) d = 1;
        lu->udiag[k] = d;
        if (growth_limit > 0 && !lu->scaled[k]) {
            int64_t cbits = (int64_t)dbits(fabs(d));
            for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) {
                uint32_t r = lu->li[p];
                double v = w[r];
                w[r] = 0;
                int64_t vb = (int64_t)dbits(fabs(v));
                if (vb > cbits) cbits = vb;
                lu->lx[p] = v / d;
            }
            double cmax = bitsd((uint64_t)cbits);
            bad = bad || !(fabs(d) >= growth_limit * cmax);
        } else {
            for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) {
                uint32_t r = lu->li[p];
                double lv = w[r] / d;
                nonfinite |= dbits(lv - lv);
                lu->lx[p] = lv;
                w[r] = 0;
            }
        }
    }
    if (nonfinite != 0) return 1;
    uint64_t worst = 0;
    for (uint32_t k = 0; k < n; k++) {
        if (lu->scaled[k]) continue;
        for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) {
            uint64_t b = dbits(fabs(lu->lx[p]));
            if (b > worst) worst = b;
        }
    }
    return bad || !(bitsd(worst) <= wide);
}

void orc_lu_solve(orc_lu *lu, const double *b, double *x)
{
    const uint32_t n = lu->n;
    double *y = lu->y;
    for (uint32_t r = 0; r < n; r++) y[lu->pinv[r]] = b[r];
    for (uint32_t k = 0; k < n; k++) {
        double yk = y[k];
        if (yk == 0) continue;
        for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) y[lu->li[p]] -= lu->lx[p] * yk;
    }
    for (uint32_t k = n; k > 0;) {
        k--;
        double zk = y[k] / lu->udiag[k];
        y[k] = zk;
        if (zk == 0) continue;
        for (uint32_t p = lu->up[k]; p < lu->up[k + 1]; p++) y[lu->ui[p]] -= lu->ux[p] * zk;
    }
    for (uint32_t j = 0; j < n; j++) x[lu->q[j]] = y[j];
}

void orc_lu_solve_neg(orc_lu *lu, const double *rhs, double *dx)
{
    const uint32_t n = lu->n;
    double *y = lu->y;
    for (uint32_t r = 0; r < n; r++) y[lu->pinv[r]] = -rhs[r];
    for (uint32_t k = 0; k < n; k++) {
        double yk = y[k];
        if (yk != 0)
            for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) y[lu->li[p]] -= lu->lx[p] * yk;
    }
    for (uint32_t k = n; k > 0;) {
        k--;
        double zk = y[k] / lu->udiag[k];
        y[k] = zk;
        if (zk != 0)
            for (uint32_t p = lu->up[k]; p < lu->up[k + 1]; p++) y[lu->ui[p]] -= lu->ux[p] * zk;
    }
    for (uint32_t j = 0; j < n; j++) dx[lu->q[j]] = y[j];
}

uint64_t orc_lu_flops(const orc_lu *lu)
{
    uint64_t f = 0;
    for (uint32_t p = 0; p < lu->nnz_u; p++) {
        uint32_t i = lu->ui[p];
        f += lu->lp[i + 1] - lu->lp[i];
    }
    return f;
}

int orc_lu_factor_export(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx,
                         const double *vals, const uint32_t *q, double pivot_tol,
                         uint32_t *sizes, uint32_t *pinv, uint32_t *lp, uint32_t *li,
                         uint32_t *up, uint32_t *ui, uint32_t *flags, uint32_t *prow,
                         double *lx, double *ux, double *udiag, size_t cap_l, size_t cap_u)
{
    orc_lu *lu = orc_lu_new(n, col_ptr[n]);
    int rc = orc_lu_factor(lu, col_ptr, row_idx, vals, q, pivot_tol);
    if (rc == 0) {
        sizes[0] = lu->nnz_l;
        sizes[1] = lu->nnz_u;
        sizes[2] = lu->n_void_slots;
        if (lu->nnz_l <= cap_l && lu->nnz_u <= cap_u) {
            if (pinv) memcpy(pinv, lu->pinv, (size_t)n * 4);
            if (lp) memcpy(lp, lu->lp, (size_t)(n + 1) * 4);
            if (up) memcpy(up, lu->up, (size_t)(n + 1) * 4);
            if (li) memcpy(li, lu->li, (size_t)lu->nnz_l * 4);
            if (ui) memcpy(ui, lu->ui, (size_t)lu->nnz_u * 4);
            if (lx) memcpy(lx, lu->lx, (size_t)lu->nnz_l * 8);
            if (ux) memcpy(ux, lu->ux, (size_t)lu->nnz_u * 8);
            if (udiag) memcpy(udiag, lu->udiag, (size_t)n * 8);
            if (prow) memcpy(prow, lu->prow, (size_t)col_ptr[n] * 4);
            if (flags)
                for (uint32_t k = 0; k < n; k++)
                    flags[k] = (lu->void_col[k] ? 1u : 0u) | (lu->scaled[k] ? 2u : 0u);
        } else {
            rc = 1;
        }
    }
    orc_lu_free(lu);
    return rc;
}

int orc_lu_replay_export(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx,
                         const double *lead_vals, const uint32_t *q, double pivot_tol,
                         const double *a, const double *rhs, size_t lanes, double growth,
                         double *lx, double *ux, double *udiag, double *dx, uint32_t *fail)
{
    const uint32_t nnz = col_ptr[n];
    orc_lu *lu = orc_lu_new(n, nnz);
    if (orc_lu_factor(lu, col_ptr, row_idx, lead_vals, q, pivot_tol) != 0) {
        orc_lu_free(lu);
        return -1;
    }
    const double wide = 1.0 / pivot_tol;
    for (size_t l = 0; l < lanes; l++) {
        int bad = orc_lu_lane_refactor(lu, col_ptr, a + l * nnz, growth, wide);
        if (fail) fail[l] = (uint32_t)bad;
        if (lx) memcpy(lx + l * lu->nnz_l, lu->lx, (size_t)lu->nnz_l * 8);
        if (ux) memcpy(ux + l * lu->nnz_u, lu->ux, (size_t)lu->nnz_u * 8);
        if (udiag) memcpy(udiag + l * n, lu->udiag, (size_t)n * 8);
        if (dx && rhs) orc_lu_solve_neg(lu, rhs + l * n, dx + l * n);
    }
    orc_lu_free(lu);
    return 0;
}

int orc_lu_solve_export(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx,
                        const double *vals, const uint32_t *q, double pivot_tol,
                        const double *b, double *x)
{
    orc_lu *lu = orc_lu_new(n, col_ptr[n]);
    int rc = orc_lu_factor(lu, col_ptr, row_idx, vals, q, pivot_tol);
    if (rc == 0) orc_lu_solve(lu, b, x);
    orc_lu_free(lu);
    return rc;
}

/* ---- FP32 factors on an FP64 tape ---- */
static uint64_t fbits64(float f) { return dbits((double)f); }

orc_lu32 *orc_lu32_new(uint32_t n)
{
    orc_lu32 *f = calloc(1, sizeof(orc_lu32));
    f->n = n;
    f->udiag = calloc(n + 1, 4);
    f->w = calloc(n + 1, 4);
    f->y = calloc(n + 1, 4);
    return f;
}

void orc_lu32_free(orc_lu32 *f)
{
    if (!f) return;
    free(f->lx); free(f->ux); free(f->udiag); free(f->w); free(f->y);
    free(f);
}

static void lu32_fit(const orc_lu *lu, orc_lu32 *f)
{
    if (f->cap_l < lu->nnz_l + 1) {
        f->cap_l = lu->nnz_l + 1;
        f->lx = realloc(f->lx, (size_t)f->cap_l * 4);
    }
    if (f->cap_u < lu->nnz_u + 1) {
        f->cap_u = lu->nnz_u + 1;
        f->ux = realloc(f->ux, (size_t)f->cap_u * 4);
    }
}

/* orc_lu_lane_refactor in FP32: the same column loop with a float
 * accumulator; growth (cmax as the bit-pattern max) and wide in FP64. */
int orc_lu_lane_refactor32(const orc_lu *lu, orc_lu32 *f, const uint32_t *col_ptr,
                           const double *vals, double growth_limit, double wide)
{
    const uint32_t n = lu->n;
    lu32_fit(lu, f);
    float *w = f->w;
    int bad = 0;
    uint64_t nonfinite = 0;
    for (uint32_t v = 0; v < lu->n_void_slots; v++)
        if ((float)vals[lu->void_slots[v]] != 0) return 1;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t c = lu->q[k];
        for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) w[lu->prow[p]] = (float)vals[p];
        for (uint32_t p = lu->up[k]; p < lu->up[k + 1]; p++) {
            uint32_t i = lu->ui[p];
            float uki = w[i];
            w[i] = 0;
            f->ux[p] = uki;
            nonfinite |= fbits64(uki - uki);
            for (uint32_t pl = lu->lp[i]; pl < lu->lp[i + 1]; pl++)
                w[lu->li[pl]] = w[lu->li[pl]] - f->lx[pl] * uki;
        }
        if (lu->void_col[k]) {
            f->udiag[k] = 1;
            for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) { f->lx[p] = 0; w[lu->li[p]] = 0; }
            w[k] = 0;
            for (uint32_t p = lu->up[k]; p < lu->up[k + 1]; p++) {
                uint32_t i = lu->ui[p];
                for (uint32_t pl = lu->lp[i]; pl < lu->lp[i + 1]; pl++) w[lu->li[pl]] = 0;
            }
            continue;
        }
        float d = w[k];
        w[k] = 0;
        int dead = d == 0 || !(fabs((double)d) < INFINITY);
        bad = bad || dead;
        if (dead) d = 1;
        f->udiag[k] = d;
        if (growth_limit > 0 && !lu->scaled[k]) {
            int64_t cbits = (int64_t)fbits64(fabsf(d));
            for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) {
                uint32_t r = lu->li[p];
                float v = w[r];
                w[r] = 0;
                int64_t vb = (int64_t)fbits64(fabsf(v));
                if (vb > cbits) cbits = vb;
                f->lx[p] = v / d;
            }
            double cmax = bitsd((uint64_t)cbits);
            bad = bad || !(fabs((double)d) >= growth_limit * cmax);
        } else {
            for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) {
                uint32_t r = lu->li[p];
                float lv = w[r] / d;
                nonfinite |= fbits64(lv - lv);
                f->lx[p] = lv;
                w[r] = 0;
            }
        }
    }
    if (nonfinite != 0) return 1;
    uint64_t worst = 0;
    for (uint32_t k = 0; k < n; k++) {
        if (lu->scaled[k]) continue;
        for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) {
            uint64_t b = fbits64(fabsf(f->lx[p]));
            if (b > worst) worst = b;
        }
    }
    return bad || !(bitsd(worst) <= wide);
}

void orc_lu_solve32(const orc_lu *lu, orc_lu32 *f, const float *b, float *x)
{
    const uint32_t n = lu->n;
    float *y = f->y;
    for (uint32_t r = 0; r < n; r++) y[lu->pinv[r]] = b[r];
    for (uint32_t k = 0; k < n; k++) {
        float yk = y[k];
        if (yk != 0)
            for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) y[lu->li[p]] = y[lu->li[p]] - f->lx[p] * yk;
    }
    for (uint32_t k = n; k > 0;) {
        k--;
        float zk = y[k] / f->udiag[k];
        y[k] = zk;
        if (zk != 0)
            for (uint32_t p = lu->up[k]; p < lu->up[k + 1]; p++) y[lu->ui[p]] = y[lu->ui[p]] - f->ux[p] * zk;
    }
    for (uint32_t j = 0; j < n; j++) x[lu->q[j]] = y[j];
}

int orc_lu_replay32_export(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx,
                           const double *lead_vals, const uint32_t *q, double pivot_tol,
                           const double *a, const double *b, size_t lanes, double growth,
                           float *lx, float *ux, float *udiag, float *x, uint32_t *fail)
{
    const uint32_t nnz = col_ptr[n];
    orc_lu *lu = orc_lu_new(n, nnz);
    if (orc_lu_factor(lu, col_ptr, row_idx, lead_vals, q, pivot_tol) != 0) {
        orc_lu_free(lu);
        return -1;
    }
    orc_lu32 *f = orc_lu32_new(n);
    float *b32 = calloc(n + 1, 4);
    const double wide = 1.0 / pivot_tol;
    for (size_t l = 0; l < lanes; l++) {
        int bad = orc_lu_lane_refactor32(lu, f, col_ptr, a + l * nnz, growth, wide);
        if (fail) fail[l] = (uint32_t)bad;
        if (lx) memcpy(lx + l * lu->nnz_l, f->lx, (size_t)lu->nnz_l * 4);
        if (ux) memcpy(ux + l * lu->nnz_u, f->ux, (size_t)lu->nnz_u * 4);
        if (udiag) memcpy(udiag + l * n, f->udiag, (size_t)n * 4);
        if (x && b) {
            for (uint32_t i = 0; i < n; i++) b32[i] = (float)b[l * n + i];
            orc_lu_solve32(lu, f, b32, x + l * n);
        }
    }
    free(b32);
    orc_lu32_free(f);
    orc_lu_free(lu);
    return 0;
}
