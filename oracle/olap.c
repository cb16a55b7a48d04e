/*
 * olap.c -- oracle restatement of the Laplace/Cauchy FMM (test
 * infrastructure; see oracle.h).  OpenMP parallel over boxes inside each of
 * the eight FMM steps, the way the Fortran original schedules them
 * (src/laplace/cfmm2d.f:763-1000, PARALLEL DO SCHEDULE(DYNAMIC)).
 *
 * Follows:
 *   P2P kernels c2d_direct{c,d,cd}{p,g,h}   c_translation/src/cauchykernels2d.c:31-442
 *   form mp / local, eval, M2M, L2L, M2L    c_translation/src/laprouts2d.c:41-990
 *   binomial table                          c_translation/src/fmmcommon2d.c:110-129
 *   8-step engine                           c_translation/src/cfmm2d.c:390-1047
 *   drivers                                 c_translation/src/cfmm2d.c:1053-1401,
 *                                           c_translation/src/cfmm2d_ndiv.c:94-439,
 *                                           c_translation/src/lfmm2d_ndiv.c:57-253,
 *                                           src/laplace/lfmm2d.f:29-200
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define LO(l) (2 * (l))
#define HI(l) (2 * (l) + 1)

/* ------------------------------------------------------------------ */
/* near field                                                          */
/* ------------------------------------------------------------------ */

/* One source->target pair block.  The skip test differs per routine in the
 * reference (cauchykernels2d.c:10-14): charge-only potential and
 * charge+dipole potential skip on rr <= thresh^2, charge gradients on
 * rr < thresh^2, dipole-only routines on |z| <= thresh (potential) or
 * |z| < thresh (gradient/Hessian). */
void oracle_c2d_direct(int nd, const double *src, int ns, const cplx *charge,
                       const cplx *dipstr, const double *targ, int nt, int ifpgh,
                       cplx *pot, cplx *grad, cplx *hess, double thresh)
{
    double th2 = thresh * thresh;
    int dip_only = (charge == NULL);
    for (int j = 0; j < nt; j++) {
        for (int i = 0; i < ns; i++) {
            double dx = targ[2 * j] - src[2 * i], dy = targ[2 * j + 1] - src[2 * i + 1];
            double rr = dx * dx + dy * dy;
            cplx z = dx + dy * I;
            int skip;
            if (dip_only) skip = ifpgh == 1 ? (cabs(z) <= thresh) : (cabs(z) < thresh);
            else skip = ifpgh == 1 ? (rr <= th2) : (rr < th2);
            if (skip) continue;
            cplx zi = 1.0 / z, z2 = -zi * zi, z3 = -2.0 * z2 * zi;
            double lr = log(rr) / 2.0;
            for (int d = 0; d < nd; d++) {
                if (charge) {
                    cplx c = charge[nd * i + d];
                    pot[nd * j + d] += lr * c;
                    if (ifpgh >= 2) grad[nd * j + d] += zi * c;
                    if (ifpgh >= 3) hess[nd * j + d] += z2 * c;
                }
                if (dipstr) {
                    cplx q = dipstr[nd * i + d];
                    pot[nd * j + d] += zi * q;
                    if (ifpgh >= 2) grad[nd * j + d] += z2 * q;
                    if (ifpgh >= 3) hess[nd * j + d] += z3 * q;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* expansions (laprouts2d.c)                                           */
/* ------------------------------------------------------------------ */

/* mp_0 += c;  mp_n += -c (z/r)^n / n + q z^(n-1) / r^n  (laprouts2d.c:41-182) */
static void form_mp(int nd, double r, const double *src, int ns, const cplx *c,
                    const cplx *q, const double *ctr, int p, cplx *mp, cplx *w)
{
    for (int j = 0; j < ns; j++) {
        cplx z = (src[2 * j] - ctr[0]) + (src[2 * j + 1] - ctr[1]) * I;
        cplx zr = z / r, zc = -1.0, zd = 1.0 / r;
        for (int n = 1; n <= p; n++) {
            zc *= zr;
            w[n] = zc / (double)n;         /* charge weight */
            w[p + 1 + n] = zd;              /* dipole weight */
            zd *= zr;
        }
        for (int d = 0; d < nd; d++) {
            if (c) mp[d] += c[nd * j + d];
            for (int n = 1; n <= p; n++) {
                if (q) mp[nd * n + d] += q[nd * j + d] * w[p + 1 + n];
                if (c) mp[nd * n + d] += c[nd * j + d] * w[n];
            }
        }
    }
}

/* local from far sources: L_0 += c log|z0|; L_n += -c (r/z0)^n / n
 * - q r^n / z0^(n+1)  (laprouts2d.c:358-500) */
static void form_ta(int nd, double r, const double *src, int ns, const cplx *c,
                    const cplx *q, const double *ctr, int p, cplx *loc, cplx *w)
{
    for (int j = 0; j < ns; j++) {
        cplx z0 = (src[2 * j] - ctr[0]) + (src[2 * j + 1] - ctr[1]) * I;
        cplx t = r / z0, zc = -1.0, zd = -1.0 / z0;
        w[0] = log(cabs(z0));
        w[p + 1] = zd;
        for (int n = 1; n <= p; n++) {
            zc *= t;
            w[n] = zc / (double)n;
            zd *= t;
            w[p + 1 + n] = zd;
        }
        for (int d = 0; d < nd; d++)
            for (int n = 0; n <= p; n++) {
                if (c) loc[nd * n + d] += c[nd * j + d] * w[n];
                if (q) loc[nd * n + d] += q[nd * j + d] * w[p + 1 + n];
            }
    }
}

/* pot += mp_0 log|z| + sum mp_n (r/z)^n, grad/hess = d/dz, d2/dz2
 * (laprouts2d.c:193-344) */
static void eval_mp(int nd, double r, const double *ctr, const cplx *mp, int p,
                    const double *t, int nt, int ifpgh, cplx *pot, cplx *grad,
                    cplx *hess, cplx *w)
{
    double ri = 1.0 / r;
    for (int k = 0; k < nt; k++) {
        cplx z = (t[2 * k] - ctr[0]) + (t[2 * k + 1] - ctr[1]) * I;
        cplx zt = r / z;
        w[0] = log(cabs(z));
        w[1] = zt;
        for (int n = 2; n <= p + 2; n++) w[n] = w[n - 1] * zt;
        for (int d = 0; d < nd; d++) {
            cplx P = 0, G = 0, H = 0;
            P = mp[d] * w[0];
            G = mp[d] * (w[1] * ri);
            H = mp[d] * (-w[2] * ri * ri);
            for (int n = 1; n <= p; n++) {
                cplx m = mp[nd * n + d];
                P += m * w[n];
                G += m * (-w[n + 1] * ri * n);
                H += m * (w[n + 2] * (double)(n * (n + 1)) * ri * ri);
            }
            pot[nd * k + d] += P;
            if (ifpgh >= 2) grad[nd * k + d] += G;
            if (ifpgh >= 3) hess[nd * k + d] += H;
        }
    }
}

/* pot += sum L_n ((z-c)/r)^n and derivatives (laprouts2d.c:508-642) */
static void eval_loc(int nd, double r, const double *ctr, const cplx *loc, int p,
                     const double *t, int nt, int ifpgh, cplx *pot, cplx *grad,
                     cplx *hess, cplx *w)
{
    double ri = 1.0 / r;
    for (int k = 0; k < nt; k++) {
        cplx z = ((t[2 * k] - ctr[0]) + (t[2 * k + 1] - ctr[1]) * I) / r;
        w[0] = 1.0;
        for (int n = 1; n <= p; n++) w[n] = w[n - 1] * z;
        for (int d = 0; d < nd; d++) {
            cplx P = 0, G = 0, H = 0;
            for (int n = 0; n <= p; n++) {
                cplx L = loc[nd * n + d];
                P += L * w[n];
                if (n >= 1) G += L * (w[n - 1] * (double)n * ri);
                if (n >= 2) H += L * (w[n - 2] * (double)(n * (n - 1)) * ri * ri);
            }
            pot[nd * k + d] += P;
            if (ifpgh >= 2) grad[nd * k + d] += G;
            if (ifpgh >= 3) hess[nd * k + d] += H;
        }
    }
}

static double *binom_table(void)
{
    double *C = calloc(101 * 101, sizeof(double)); /* C[l + 101*m] = binom(l, m) */
    for (int l = 0; l <= 100; l++) {
        C[l] = 1.0;
        for (int m = 1; m <= l; m++)
            C[l + 101 * m] = C[(l - 1) + 101 * (m - 1)] + (m <= l - 1 ? C[(l - 1) + 101 * m] : 0.0);
    }
    return C;
}
#define BIN(C, l, m) ((C)[(l) + 101 * (m)])

/* child multipole (centre c1, scale r1, order p1) -> parent (c2, r2, p2)
 * (laprouts2d.c:666-766) */
static void mpmp(int nd, double r1, const double *c1, const cplx *h1, int p1,
                 double r2, const double *c2, cplx *h2, int p2, const double *C,
                 cplx *tmp)
{
    cplx z0 = (c1[0] - c2[0]) + (c1[1] - c2[1]) * I;
    cplx a = r1 / z0, b = z0 / r2;
    for (int d = 0; d < nd; d++) {
        cplx *x = tmp; /* x_i = h1_i (r1/z0)^i */
        cplx pw = 1.0;
        for (int i = 0; i <= p1; i++) { x[i] = h1[nd * i + d] * pw; pw *= a; }
        h2[d] += h1[d];
        cplx pb = 1.0;
        for (int i = 1; i <= p2; i++) {
            pb *= b;
            cplx s = -x[0] / (double)i;
            int jm = i < p1 ? i : p1;
            for (int j = 1; j <= jm; j++) s += x[j] * BIN(C, i - 1, j - 1);
            h2[nd * i + d] += s * pb;
        }
    }
}

/* parent local (c1, r1, p1) -> child local (c2, r2, p2) (laprouts2d.c:777-867) */
static void locloc(int nd, double r1, const double *c1, const cplx *j1, int p1,
                   double r2, const double *c2, cplx *j2, int p2, const double *C,
                   cplx *tmp)
{
    cplx z0 = (c2[0] - c1[0]) + (c2[1] - c1[1]) * I;
    cplx a = z0 / r1, b = r2 / z0;
    for (int d = 0; d < nd; d++) {
        cplx *x = tmp;
        cplx pw = 1.0;
        for (int i = 0; i <= p1; i++) { x[i] = j1[nd * i + d] * pw; pw *= a; }
        cplx pb = 1.0;
        for (int i = 0; i <= p2; i++) {
            cplx s = 0;
            for (int j = i; j <= p1; j++) s += x[j] * BIN(C, j, i);
            j2[nd * i + d] += s * pb;
            pb *= b;
        }
    }
}

/* multipole (c1, r1, p1) -> local (c2, r2, p2) (laprouts2d.c:886-990) */
static void mploc(int nd, double r1, const double *c1, const cplx *h1, int p1,
                  double r2, const double *c2, cplx *j2, int p2, const double *C,
                  cplx *tmp)
{
    cplx z0 = (c1[0] - c2[0]) + (c1[1] - c2[1]) * I;
    cplx zi = 1.0 / z0;
    double lg = log(cabs(z0));
    for (int d = 0; d < nd; d++) {
        cplx *x = tmp; /* x_j = h1_j (-r1/z0)^j */
        cplx pw = 1.0, a = -r1 * zi;
        for (int j = 0; j <= p1; j++) { x[j] = h1[nd * j + d] * pw; pw *= a; }
        cplx s0 = x[0] * lg;
        for (int j = 1; j <= p1; j++) s0 += x[j];
        j2[d] += s0;
        cplx pb = 1.0, b = r2 * zi;
        for (int i = 1; i <= p2; i++) {
            pb *= b;
            cplx s = -x[0] / (double)i;
            for (int j = 1; j <= p1; j++) s += x[j] * BIN(C, i + j - 1, j - 1);
            j2[nd * i + d] += s * pb;
        }
    }
}

/* ------------------------------------------------------------------ */
/* 8-step engine (cfmm2d.c:390-1047), nexpc = 0                        */
/* ------------------------------------------------------------------ */
typedef struct {
    int nd, ifcharge, ifdipole, ifpgh, ifpghtarg, ifnear;
    const double *src, *targ;   /* sorted */
    const cplx *chg, *dip;      /* sorted */
    cplx *pot, *grad, *hess, *pott, *gradt, *hesst; /* sorted, zeroed */
} lap_job;

static int npts_tgt(const otree *t, const lap_job *J, int ib)
{
    int n = 0;
    if (J->ifpghtarg > 0) n += t->itargse[2 * ib + 1] - t->itargse[2 * ib] + 1;
    if (J->ifpgh > 0) n += t->isrcse[2 * ib + 1] - t->isrcse[2 * ib] + 1;
    return n;
}

static void lap_main(const otree *t, const olists *L, const lap_job *J, double eps)
{
    int nd = J->nd, nl = t->nlevels, nb = t->nboxes;
    int p = o_l2dterms(eps);
    double *C = binom_table();
    size_t per = (size_t)(p + 1) * nd;
    cplx *mp = calloc(per * (nb + 1), sizeof(cplx));
    cplx *lo = calloc(per * (nb + 1), sizeof(cplx));
    const double *bs = t->boxsize;
    const cplx *c = J->ifcharge ? J->chg : NULL, *q = J->ifdipole ? J->dip : NULL;
#define MP(ib) (mp + per * (ib))
#define LC(ib) (lo + per * (ib))
#define S0(ib) (t->isrcse[2 * (ib)])
#define NS(ib) (t->isrcse[2 * (ib) + 1] - t->isrcse[2 * (ib)] + 1)
#define T0(ib) (t->itargse[2 * (ib)])
#define NT(ib) (t->itargse[2 * (ib) + 1] - t->itargse[2 * (ib)] + 1)

    /* step 1: leaf multipoles */
    for (int l = 2; l <= nl; l++) {
#pragma omp parallel
        {
            cplx *w = malloc(sizeof(cplx) * (2 * p + 4));
#pragma omp for schedule(dynamic, 16)
            for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++)
                if (t->nchild[ib] == 0 && NS(ib) > 0)
                    form_mp(nd, bs[l], J->src + 2 * (S0(ib) - 1), NS(ib),
                            c ? c + nd * (S0(ib) - 1) : NULL, q ? q + nd * (S0(ib) - 1) : NULL,
                            t->centers + 2 * ib, p, MP(ib), w);
            free(w);
        }
    }
    /* step 2: list 4 */
    for (int l = 2; l <= nl; l++) {
#pragma omp parallel
        {
            cplx *w = malloc(sizeof(cplx) * (2 * p + 4));
#pragma omp for schedule(dynamic, 16)
            for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++) {
                if (npts_tgt(t, J, ib) <= 0) continue;
                for (int i = 0; i < L->n4[ib]; i++) {
                    int jb = L->l4[L->mn4 * ib + i];
                    form_ta(nd, bs[l], J->src + 2 * (S0(jb) - 1), NS(jb),
                            c ? c + nd * (S0(jb) - 1) : NULL, q ? q + nd * (S0(jb) - 1) : NULL,
                            t->centers + 2 * ib, p, LC(ib), w);
                }
            }
            free(w);
        }
    }
    /* step 3: M2M */
    for (int l = nl - 1; l >= 1; l--) {
#pragma omp parallel
        {
            cplx *w = malloc(sizeof(cplx) * (p + 2));
#pragma omp for schedule(dynamic, 16)
            for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++)
                for (int i = 0; i < t->nchild[ib]; i++) {
                    int jb = t->ichild[4 * ib + i];
                    if (NS(jb) > 0)
                        mpmp(nd, bs[l + 1], t->centers + 2 * jb, MP(jb), p, bs[l],
                             t->centers + 2 * ib, MP(ib), p, C, w);
                }
            free(w);
        }
    }
    /* step 4: M2L */
    for (int l = 2; l <= nl; l++) {
#pragma omp parallel
        {
            cplx *w = malloc(sizeof(cplx) * (p + 2));
#pragma omp for schedule(dynamic, 16)
            for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++) {
                if (npts_tgt(t, J, ib) <= 0) continue;
                for (int i = 0; i < L->n2[ib]; i++) {
                    int jb = L->l2[L->mn2 * ib + i];
                    mploc(nd, bs[l], t->centers + 2 * jb, MP(jb), p, bs[l],
                          t->centers + 2 * ib, LC(ib), p, C, w);
                }
            }
            free(w);
        }
    }
    /* step 5: L2L */
    for (int l = 1; l <= nl - 1; l++) {
#pragma omp parallel
        {
            cplx *w = malloc(sizeof(cplx) * (p + 2));
#pragma omp for schedule(dynamic, 16)
            for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++) {
                if (npts_tgt(t, J, ib) <= 0) continue;
                for (int i = 0; i < t->nchild[ib]; i++) {
                    int jb = t->ichild[4 * ib + i];
                    locloc(nd, bs[l], t->centers + 2 * ib, LC(ib), p, bs[l + 1],
                           t->centers + 2 * jb, LC(jb), p, C, w);
                }
            }
            free(w);
        }
    }
    /* step 6: list 3 */
    for (int l = 1; l <= nl - 1; l++) {
#pragma omp parallel
        {
            cplx *w = malloc(sizeof(cplx) * (p + 4));
#pragma omp for schedule(dynamic, 16)
            for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++)
                for (int i = 0; i < L->n3[ib]; i++) {
                    int jb = L->l3[L->mn3 * ib + i];
                    if (J->ifpghtarg > 0)
                        eval_mp(nd, bs[l + 1], t->centers + 2 * jb, MP(jb), p,
                                J->targ + 2 * (T0(ib) - 1), NT(ib), J->ifpghtarg,
                                J->pott + nd * (T0(ib) - 1), J->gradt + nd * (T0(ib) - 1),
                                J->hesst + nd * (T0(ib) - 1), w);
                    if (J->ifpgh > 0)
                        eval_mp(nd, bs[l + 1], t->centers + 2 * jb, MP(jb), p,
                                J->src + 2 * (S0(ib) - 1), NS(ib), J->ifpgh,
                                J->pot + nd * (S0(ib) - 1), J->grad + nd * (S0(ib) - 1),
                                J->hess + nd * (S0(ib) - 1), w);
                }
            free(w);
        }
    }
    /* step 7: local evaluation at leaves */
    for (int l = 0; l <= nl; l++) {
#pragma omp parallel
        {
            cplx *w = malloc(sizeof(cplx) * (p + 2));
#pragma omp for schedule(dynamic, 16)
            for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++) {
                if (t->nchild[ib] != 0) continue;
                if (J->ifpghtarg > 0)
                    eval_loc(nd, bs[l], t->centers + 2 * ib, LC(ib), p,
                             J->targ + 2 * (T0(ib) - 1), NT(ib), J->ifpghtarg,
                             J->pott + nd * (T0(ib) - 1), J->gradt + nd * (T0(ib) - 1),
                             J->hesst + nd * (T0(ib) - 1), w);
                if (J->ifpgh > 0)
                    eval_loc(nd, bs[l], t->centers + 2 * ib, LC(ib), p,
                             J->src + 2 * (S0(ib) - 1), NS(ib), J->ifpgh,
                             J->pot + nd * (S0(ib) - 1), J->grad + nd * (S0(ib) - 1),
                             J->hess + nd * (S0(ib) - 1), w);
            }
            free(w);
        }
    }
    /* step 8: list 1 direct */
    double thresh = bs[0] * 0x1p-51;
    if (J->ifnear) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int ib = 1; ib <= nb; ib++)
            for (int i = 0; i < L->n1[ib]; i++) {
                int jb = L->l1[L->mn1 * ib + i];
                const double *s = J->src + 2 * (S0(jb) - 1);
                const cplx *cc = c ? c + nd * (S0(jb) - 1) : NULL;
                const cplx *qq = q ? q + nd * (S0(jb) - 1) : NULL;
                if (J->ifpghtarg > 0)
                    oracle_c2d_direct(nd, s, NS(jb), cc, qq, J->targ + 2 * (T0(ib) - 1), NT(ib),
                                      J->ifpghtarg, J->pott + nd * (T0(ib) - 1),
                                      J->gradt + nd * (T0(ib) - 1), J->hesst + nd * (T0(ib) - 1), thresh);
                if (J->ifpgh > 0)
                    oracle_c2d_direct(nd, s, NS(jb), cc, qq, J->src + 2 * (S0(ib) - 1), NS(ib),
                                      J->ifpgh, J->pot + nd * (S0(ib) - 1),
                                      J->grad + nd * (S0(ib) - 1), J->hess + nd * (S0(ib) - 1), thresh);
            }
    }
    free(C); free(mp); free(lo);
#undef MP
#undef LC
}

/* gather rows of a (w, n) array by a 1-based permutation */
static void *gather(const void *a, size_t w, int n, const int *perm)
{
    char *o = malloc(w * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) memcpy(o + w * i, (const char *)a + w * (perm[i] - 1), w);
    return o;
}
static void scatter(void *dst, const void *a, size_t w, int n, const int *perm)
{
    for (int i = 0; i < n; i++) memcpy((char *)dst + w * (perm[i] - 1), (const char *)a + w * i, w);
}

/* cfmm2d_ndiv semantics (cfmm2d_ndiv.c:94-439) */
void orc_cfmm2d_ndiv_(const int *nd, const double *eps, const int *ns, const double *sources,
                      const int *ifcharge, const cplx *charge, const int *ifdipole,
                      const cplx *dipstr, int *iper, const int *ifpgh, cplx *pot, cplx *grad,
                      cplx *hess, const int *nt, const double *targ, const int *ifpghtarg,
                      cplx *pottarg, cplx *gradtarg, cplx *hesstarg, const int *ndiv,
                      const int *idivflag, const int *ifnear, double *timeinfo, int *ier)
{
    otree t;
    olists L;
    *iper = 0;
    for (int i = 0; i < 8; i++) timeinfo[i] = 0.0;
    otree_build(sources, *ns, targ, *nt, *idivflag, *ndiv, 0, 51, 0, &t);
    olists_build(&t, &L);
    size_t cw = sizeof(cplx) * (*nd);
    lap_job J;
    memset(&J, 0, sizeof(J));
    J.nd = *nd; J.ifcharge = *ifcharge; J.ifdipole = *ifdipole;
    J.ifpgh = *ifpgh; J.ifpghtarg = *ifpghtarg; J.ifnear = *ifnear;
    J.src = gather(sources, 2 * sizeof(double), *ns, t.isrc);
    J.targ = gather(targ, 2 * sizeof(double), *nt, t.itarg);
    J.chg = *ifcharge ? gather(charge, cw, *ns, t.isrc) : NULL;
    J.dip = *ifdipole ? gather(dipstr, cw, *ns, t.isrc) : NULL;
    int ns1 = *ns > 0 ? *ns : 1, nt1 = *nt > 0 ? *nt : 1;
    J.pot = calloc((size_t)ns1 * (*nd), sizeof(cplx));
    J.grad = calloc((size_t)ns1 * (*nd), sizeof(cplx));
    J.hess = calloc((size_t)ns1 * (*nd), sizeof(cplx));
    J.pott = calloc((size_t)nt1 * (*nd), sizeof(cplx));
    J.gradt = calloc((size_t)nt1 * (*nd), sizeof(cplx));
    J.hesst = calloc((size_t)nt1 * (*nd), sizeof(cplx));
    lap_main(&t, &L, &J, *eps);
    if (*ifpgh >= 1) scatter(pot, J.pot, cw, *ns, t.isrc);
    if (*ifpgh >= 2) scatter(grad, J.grad, cw, *ns, t.isrc);
    if (*ifpgh >= 3) scatter(hess, J.hess, cw, *ns, t.isrc);
    if (*ifpghtarg >= 1) scatter(pottarg, J.pott, cw, *nt, t.itarg);
    if (*ifpghtarg >= 2) scatter(gradtarg, J.gradt, cw, *nt, t.itarg);
    if (*ifpghtarg >= 3) scatter(hesstarg, J.hesst, cw, *nt, t.itarg);
    free((void *)J.src); free((void *)J.targ); free((void *)J.chg); free((void *)J.dip);
    free(J.pot); free(J.grad); free(J.hess); free(J.pott); free(J.gradt); free(J.hesst);
    olists_free(&L);
    otree_free(&t);
    *ier = 0;
}

/* cfmm2d: lndiv2d picks ndiv, idivflag 0, near field on (cfmm2d.c:1115-1124) */
void orc_cfmm2d_(const int *nd, const double *eps, const int *ns, const double *sources,
                 const int *ifcharge, const cplx *charge, const int *ifdipole,
                 const cplx *dipstr, int *iper, const int *ifpgh, cplx *pot, cplx *grad,
                 cplx *hess, const int *nt, const double *targ, const int *ifpghtarg,
                 cplx *pottarg, cplx *gradtarg, cplx *hesstarg, int *ier)
{
    int ndiv = o_lndiv(*eps, *ns, *nt), idiv = 0, ifnear = 1;
    double ti[8];
    orc_cfmm2d_ndiv_(nd, eps, ns, sources, ifcharge, charge, ifdipole, dipstr, iper, ifpgh,
                     pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg,
                     &ndiv, &idiv, &ifnear, ti, ier);
}

/* Laplace with complex densities: split re/im into 2*nd Cauchy channels,
 * dipstr1 = -(v1 + i v2) Re/Im(d), unpack pot = Re(ch1) + i Re(ch2),
 * grad = (Re, -Im) of d/dz, hess_yy = -hess_xx (lfmm2d_ndiv.c:139-243). */
static void lfmm_common(int use_ndiv, const int *nd, const double *eps, const int *ns,
                        const double *sources, const int *ifcharge, const cplx *charge,
                        const int *ifdipole, const cplx *dipstr, const double *dipvec,
                        int *iper, const int *ifpgh, cplx *pot, cplx *grad, cplx *hess,
                        const int *nt, const double *targ, const int *ifpghtarg,
                        cplx *pottarg, cplx *gradtarg, cplx *hesstarg, const int *ndiv,
                        const int *idivflag, const int *ifnear, double *timeinfo, int *ier)
{
    int n = *nd, S = *ns, T = *nt, nd2 = 2 * n;
    int S1 = S > 0 ? S : 1, T1 = T > 0 ? T : 1;
    cplx *c1 = calloc((size_t)2 * n * S1, sizeof(cplx)), *d1 = calloc((size_t)2 * n * S1, sizeof(cplx));
    if (*ifcharge == 1)
        for (int i = 0; i < S; i++)
            for (int j = 0; j < n; j++) {
                c1[2 * (n * i + j)] = creal(charge[n * i + j]);
                c1[2 * (n * i + j) + 1] = cimag(charge[n * i + j]);
            }
    if (*ifdipole == 1)
        for (int i = 0; i < S; i++)
            for (int j = 0; j < n; j++) {
                cplx z = -(dipvec[j + n * (0 + 2 * i)] + dipvec[j + n * (1 + 2 * i)] * I);
                d1[2 * (n * i + j)] = creal(dipstr[n * i + j]) * z;
                d1[2 * (n * i + j) + 1] = cimag(dipstr[n * i + j]) * z;
            }
    cplx *p1 = calloc((size_t)2 * n * S1, sizeof(cplx)), *g1 = calloc((size_t)2 * n * S1, sizeof(cplx));
    cplx *h1 = calloc((size_t)2 * n * S1, sizeof(cplx));
    cplx *pt1 = calloc((size_t)2 * n * T1, sizeof(cplx)), *gt1 = calloc((size_t)2 * n * T1, sizeof(cplx));
    cplx *ht1 = calloc((size_t)2 * n * T1, sizeof(cplx));
    double ti[8];
    if (use_ndiv)
        orc_cfmm2d_ndiv_(&nd2, eps, ns, sources, ifcharge, c1, ifdipole, d1, iper, ifpgh, p1, g1, h1,
                         nt, targ, ifpghtarg, pt1, gt1, ht1, ndiv, idivflag, ifnear, timeinfo, ier);
    else
        orc_cfmm2d_(&nd2, eps, ns, sources, ifcharge, c1, ifdipole, d1, iper, ifpgh, p1, g1, h1,
                    nt, targ, ifpghtarg, pt1, gt1, ht1, ier);
    (void)ti;
    for (int pass = 0; pass < 2; pass++) {
        int N = pass ? T : S, f = pass ? *ifpghtarg : *ifpgh;
        cplx *P = pass ? pottarg : pot, *G = pass ? gradtarg : grad, *H = pass ? hesstarg : hess;
        cplx *P1 = pass ? pt1 : p1, *G1 = pass ? gt1 : g1, *H1 = pass ? ht1 : h1;
        for (int i = 0; i < N; i++)
            for (int j = 0; j < n; j++) {
                int a = 2 * (n * i + j);
                if (f >= 1) P[n * i + j] = creal(P1[a]) + I * creal(P1[a + 1]);
                if (f >= 2) {
                    G[j + n * (0 + 2 * i)] = creal(G1[a]) + I * creal(G1[a + 1]);
                    G[j + n * (1 + 2 * i)] = -cimag(G1[a]) - I * cimag(G1[a + 1]);
                }
                if (f >= 3) {
                    H[j + n * (0 + 3 * i)] = creal(H1[a]) + I * creal(H1[a + 1]);
                    H[j + n * (1 + 3 * i)] = -cimag(H1[a]) - I * cimag(H1[a + 1]);
                    H[j + n * (2 + 3 * i)] = -H[j + n * (0 + 3 * i)];
                }
            }
    }
    free(c1); free(d1); free(p1); free(g1); free(h1); free(pt1); free(gt1); free(ht1);
}

void orc_lfmm2d_ndiv_(const int *nd, const double *eps, const int *ns, const double *sources,
                      const int *ifcharge, const cplx *charge, const int *ifdipole,
                      const cplx *dipstr, const double *dipvec, int *iper, const int *ifpgh,
                      cplx *pot, cplx *grad, cplx *hess, const int *nt, const double *targ,
                      const int *ifpghtarg, cplx *pottarg, cplx *gradtarg, cplx *hesstarg,
                      const int *ndiv, const int *idivflag, const int *ifnear, double *timeinfo,
                      int *ier)
{
    lfmm_common(1, nd, eps, ns, sources, ifcharge, charge, ifdipole, dipstr, dipvec, iper, ifpgh,
                pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg, ndiv, idivflag,
                ifnear, timeinfo, ier);
}

void orc_lfmm2d_(const int *nd, const double *eps, const int *ns, const double *sources,
                 const int *ifcharge, const cplx *charge, const int *ifdipole, const cplx *dipstr,
                 const double *dipvec, int *iper, const int *ifpgh, cplx *pot, cplx *grad,
                 cplx *hess, const int *nt, const double *targ, const int *ifpghtarg,
                 cplx *pottarg, cplx *gradtarg, cplx *hesstarg, int *ier)
{
    double ti[8];
    lfmm_common(0, nd, eps, ns, sources, ifcharge, charge, ifdipole, dipstr, dipvec, iper, ifpgh,
                pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg, NULL, NULL,
                NULL, ti, ier);
}
