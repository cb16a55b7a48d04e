/*
 * ohelm.c -- oracle restatement of the Helmholtz FMM, low-frequency path
 * (test infrastructure; see oracle.h).  OpenMP over boxes per step.
 *
 * Follows:
 *   P2P h2d_direct{c,d,cd}{p,g,h}          c_translation/src/helmkernels2d.c:21-520
 *   charge/dipole -> expansion             c_translation/src/helmrouts2d.c:43-105 (ctompole,
 *                                          dtompole), formation :294-391, :513-612
 *   expansion evaluation, derivatives      helmrouts2d.c:107-287 (mpole_eval*, mk_mpole*),
 *                                          :394-510, :615-733
 *   translations mpmp / locloc / mploc     helmrouts2d.c:736-1042, written here as one
 *                                          sum out_i = sum_j in_j B(i-j) F(i,j) with the
 *                                          scale factors F implied by the scaled
 *                                          conventions (mp_n = M_n / rscale^|n|,
 *                                          lo_n = L_n rscale^|n|)
 *   8-step engine                          c_translation/src/hfmm2d.c:619-1523
 *   drivers                                hfmm2d.c:1529-1907, hfmm2d_ndiv.c:68-351
 * Boxes larger than 16 wavelengths: the reference translates them in the
 * diagonal form (wideband2d.c:33-290, hfmm2d.c:946-1254), signatures of
 * nsig = next235(2 P + 1) points multiplied pointwise by the transformed
 * kernel.  With nsig >= 2 P + 1 the circular convolution has no aliasing,
 * so the result is the dense translation sum_n a_n B(j - n) exactly; the
 * oracle computes that dense sum at every level (pinned against the
 * compiled reference's diagonal form at zk = 500 in test_oracle_helm.py).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define LO(l) (2 * (l))
#define HI(l) (2 * (l) + 1)
static const double PI = 3.14159265358979323846;
#define IMA4 (I / 4.0)

/* ------------------------------------------------------------------ */
/* near field (helmkernels2d.c): charge and dipole parts of one pair     */
/* ------------------------------------------------------------------ */
void oracle_h2d_direct(int nd, cplx zk, const double *src, int ns, const cplx *charge,
                       const cplx *dipstr, const double *dipvec, const double *targ, int nt,
                       int ifpgh, cplx *pot, cplx *grad, cplx *hess, double thresh)
{
    int pot_only_charge = (charge && !dipstr && ifpgh == 1);
    for (int j = 0; j < nt; j++) {
        for (int i = 0; i < ns; i++) {
            double dx = targ[2 * j] - src[2 * i], dy = targ[2 * j + 1] - src[2 * i + 1];
            double rr = dx * dx + dy * dy, r = sqrt(rr);
            cplx z = zk * r;
            /* h2d_directcp skips on |z| <= thresh, every other routine on < */
            if (pot_only_charge ? cabs(z) <= thresh : cabs(z) < thresh) continue;
            cplx h0, h1;
            o_hank01(z, &h0, &h1);
            double ct = dx / r, st = dy / r;
            cplx h2 = 2.0 * h1 / z - h0, h3 = 4.0 * h2 / z - h1;
            for (int d = 0; d < nd; d++) {
                cplx *P = pot + nd * j + d;
                cplx *G = grad + d + nd * 2 * j, *H = hess + d + nd * 3 * j;
                if (charge) {
                    cplx c = charge[nd * i + d];
                    *P += h0 * c * IMA4;
                    if (ifpgh >= 2) {
                        cplx cd = -h1 * (zk * IMA4 / r) * c;
                        G[0] += cd * dx;
                        G[nd] += cd * dy;
                    }
                    if (ifpgh >= 3) {
                        cplx cd = (zk * IMA4 / r) / rr * c, h2z = -z * h0 + 2.0 * h1;
                        H[0] += cd * (h2z * dx * dx - rr * h1);
                        H[nd] += cd * (h2z * dx * dy);
                        H[2 * nd] += cd * (h2z * dy * dy - rr * h1);
                    }
                }
                if (dipstr) {
                    cplx q = dipstr[nd * i + d];
                    double v1 = dipvec[d + nd * (0 + 2 * i)], v2 = dipvec[d + nd * (1 + 2 * i)];
                    *P += (h1 / r * zk * IMA4) * q * (dx * v1 + dy * v2);
                    if (ifpgh >= 2) {
                        cplx cd = -zk * zk * IMA4 * q;
                        cplx f1 = h2 * (ct * ct - 0.5) - h0 / 2.0, f2 = h2 * ct * st;
                        cplx f3 = h2 * (st * st - 0.5) - h0 / 2.0;
                        G[0] += cd * f1 * v1 + cd * f2 * v2;
                        G[nd] += cd * f2 * v1 + cd * f3 * v2;
                    }
                    if (ifpgh >= 3) {
                        cplx cd = -zk * zk * zk * q * IMA4;
                        cplx a = (-h1 / 2.0 * (-1.5) - h3 / 2.0 * (ct * ct - 0.5 - st * st)) * ct;
                        cplx b = (-h1 / 2.0 * (-0.5) - h3 / 2.0 * (1.5 * ct * ct - 0.5 * st * st)) * st;
                        cplx c2 = (-h1 / 2.0 * (-0.5) - h3 / 2.0 * (1.5 * st * st - 0.5 * ct * ct)) * ct;
                        cplx e = (-h1 / 2.0 * (-1.5) - h3 / 2.0 * (st * st - 0.5 - ct * ct)) * st;
                        H[0] += cd * a * v1 + cd * b * v2;
                        H[nd] += cd * b * v1 + cd * c2 * v2;
                        H[2 * nd] += cd * c2 * v1 + cd * e * v2;
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* expansions; layout (nd, -p:p): e[d + nd*(n + p)]                      */
/* ------------------------------------------------------------------ */
#define EX(e, d, n, p, nd) ((e)[(d) + (nd) * ((n) + (p))])

/* charge: e_0 += c v_0, e_n += c v_n zmul^n, e_-n += c v_n (-zinv)^n (helmrouts2d.c:43-64) */
static void to_pole_c(int nd, cplx zmul, cplx zinv, cplx *e, const cplx *v, const cplx *c, int p)
{
    for (int d = 0; d < nd; d++) EX(e, d, 0, p, nd) += c[d] * v[0];
    cplx a = zmul, b = -zinv;
    for (int n = 1; n <= p; n++) {
        for (int d = 0; d < nd; d++) {
            EX(e, d, n, p, nd) += v[n] * a * c[d];
            EX(e, d, -n, p, nd) += v[n] * b * c[d];
        }
        a *= zmul;
        b = -b * zinv;
    }
}

/* dipole (helmrouts2d.c:67-105); dv points at dipvec(1,1,j) */
static void to_pole_d(int nd, cplx zk, double rs, cplx zmul, cplx zinv, cplx *e, const cplx *v,
                      const cplx *q, const double *dv, int p)
{
    cplx t3 = zinv / rs, t4 = zmul * rs, t5 = zmul / rs, t6 = zinv * rs;
    for (int d = 0; d < nd; d++) {
        double v1 = dv[d], v2 = dv[d + nd];
        EX(e, d, 0, p, nd) += -q[d] * v[1] * zk / 2.0 * rs *
                              ((zmul + zinv) * v1 + (zmul - zinv) * I * v2);
    }
    cplx t1 = -zmul * zk / 2.0, t2 = zinv * zk / 2.0;
    for (int n = 1; n <= p; n++) {
        for (int d = 0; d < nd; d++) {
            double v1 = dv[d], v2 = dv[d + nd];
            cplx a7 = t3 * (-v1 + I * v2), a8 = t4 * (v1 + I * v2);
            cplx a9 = t5 * (-v1 - I * v2), a10 = t6 * (v1 - I * v2);
            EX(e, d, n, p, nd) += (v[n - 1] * a7 + v[n + 1] * a8) * t1 * q[d];
            EX(e, d, -n, p, nd) += (v[n - 1] * a9 + v[n + 1] * a10) * t2 * q[d];
        }
        t1 *= zmul;
        t2 = -t2 * zinv;
    }
}

/* derivative expansions (helmrouts2d.c:187-287): x/y partials of an expansion
 * of order p as expansions of order p+1 */
static void d_expansion(int nd, cplx zk, double rs, const cplx *e, int p, cplx *ex, cplx *ey)
{
    int q = p + 1;
    memset(ex, 0, sizeof(cplx) * nd * (2 * q + 1));
    memset(ey, 0, sizeof(cplx) * nd * (2 * q + 1));
    cplx z1 = zk / 2.0 / rs, z2 = zk / 2.0 * rs, z3 = z1 * I, z4 = z2 * I;
    for (int d = 0; d < nd; d++) {
        cplx m0 = EX(e, d, 0, p, nd);
        EX(ex, d, -1, q, nd) += z1 * m0; EX(ey, d, -1, q, nd) += z3 * m0;
        EX(ex, d, 1, q, nd) += -z1 * m0; EX(ey, d, 1, q, nd) += z3 * m0;
        for (int i = 1; i <= p; i++) {
            cplx m = EX(e, d, i, p, nd), mm = EX(e, d, -i, p, nd);
            EX(ex, d, i - 1, q, nd) += z2 * m; EX(ey, d, i - 1, q, nd) += z4 * m;
            EX(ex, d, i + 1, q, nd) += -z1 * m; EX(ey, d, i + 1, q, nd) += z3 * m;
            EX(ex, d, -i - 1, q, nd) += z1 * mm; EX(ey, d, -i - 1, q, nd) += z3 * mm;
            EX(ex, d, -i + 1, q, nd) += -z2 * mm; EX(ey, d, -i + 1, q, nd) += z4 * mm;
        }
    }
}

/* sum_n e_n v_|n| phase_n (i/4) with phase_n = zmul^n, (-zinv)^|n|, n=0 term
 * (i/4) v_0 e_0 (helmrouts2d.c:107-185) */
static cplx eval_sum(const cplx *e, int d, int nd, int p, const cplx *mt)
{
    cplx s = mt[0] * EX(e, d, 0, p, nd) * IMA4;
    for (int n = 1; n <= p; n++) s += EX(e, d, n, p, nd) * mt[n] + EX(e, d, -n, p, nd) * mt[-n];
    return s;
}

/* evaluate an expansion (multipole: v = scaled H_n; local: v = scaled J_n)
 * and its gradient / Hessian at one point */
static void eval_at(int nd, cplx zk, double rs_deriv, const cplx *e, int p, const cplx *v,
                    double th, int pg, cplx *pot, cplx *grad, cplx *hess, cplx *wk)
{
    cplx *mt0 = wk, *mt = wk + (p + 2);   /* mt[-p-2..p+2] */
    cplx zmul = cexp(I * th), zinv = conj(zmul);
    mt[0] = v[0];
    cplx a = zmul * IMA4, b = -zinv * IMA4;
    for (int j = 1; j <= p + 2; j++) {
        mt[j] = a * v[j];
        mt[-j] = b * v[j];
        a *= zmul;
        b = -b * zinv;
    }
    (void)mt0;
    cplx *ex = wk + 2 * (p + 3), *ey = ex + nd * (2 * p + 3);
    cplx *exx = ey + nd * (2 * p + 3), *exy = exx + nd * (2 * p + 5), *eyy = exy + nd * (2 * p + 5);
    if (pg >= 2) d_expansion(nd, zk, rs_deriv, e, p, ex, ey);
    if (pg >= 3) {
        cplx *tmp = eyy + nd * (2 * p + 5);
        d_expansion(nd, zk, rs_deriv, ex, p + 1, exx, exy);
        d_expansion(nd, zk, rs_deriv, ey, p + 1, tmp, eyy);
    }
    for (int d = 0; d < nd; d++) {
        pot[d] += eval_sum(e, d, nd, p, mt);
        if (pg >= 2) {
            grad[d] += eval_sum(ex, d, nd, p + 1, mt);
            grad[d + nd] += eval_sum(ey, d, nd, p + 1, mt);
        }
        if (pg >= 3) {
            hess[d] += eval_sum(exx, d, nd, p + 2, mt);
            hess[d + nd] += eval_sum(exy, d, nd, p + 2, mt);
            hess[d + 2 * nd] += eval_sum(eyy, d, nd, p + 2, mt);
        }
    }
}

/* ------------------------------------------------------------------ */
/* translations: out_i += sum_j in_j B(i-j) F(i,j), |i| <= p2, |j| <= p1  */
/* B(m) = v_|m| * (zmul^m or (-zinv)^|m|), v = scaled J (shifts) or H (M2L) */
/* evaluated at zk |c2 - c1|, angle arg(c2 - c1) - pi (helmrouts2d.c:736-1042) */
/* ------------------------------------------------------------------ */
enum { T_MPMP, T_LOCLOC, T_MPLOC };

static void translate(int kind, int nd, cplx zk, double r1, const double *c1, const cplx *in,
                      int p1, double r2, const double *c2, cplx *out, int p2)
{
    int P = p1 + p2;
    double dx = c2[0] - c1[0], dy = c2[1] - c1[1];
    double r = sqrt(dx * dx + dy * dy);
    double th = (dx == 0 && dy == 0) ? 0.0 : atan2(dy, dx);
    th -= PI;
    cplx *v = malloc(sizeof(cplx) * (P + 12));
    cplx *B = malloc(sizeof(cplx) * (2 * P + 1));
    if (kind == T_MPLOC) o_h2dall(P + 1, zk * r, r1, v);
    else o_jbessel(P + 3, zk * r, r1, v);
    cplx zmul = cexp(-I * th), zinv = conj(zmul);
    B[P] = v[0];
    cplx a = zmul, b = -zinv;
    for (int m = 1; m <= P; m++) {
        B[P + m] = a * v[m];
        B[P - m] = b * v[m];
        a *= zmul;
        b = -b * zinv;
    }
    /* The scale factors F, written with non-negative exponents only (the
     * orders satisfy |i| <= |j| + |i-j| and its permutations), so that no
     * power of a tiny scale (deep boxes: rscale ~ |zk| bs) is inverted:
     *   M2M    r1^(|j|+|m|-|i|) (r1/r2)^|i|
     *   L2L    r1^(|i|+|m|-|j|) (r2/r1)^|i|
     *   M2L    r1^(|i|+|j|-|m|) (r2/r1)^|i|       (m = i - j)          */
    double *pr = malloc(sizeof(double) * (2 * P + 2)), *pq = malloc(sizeof(double) * (2 * P + 2));
    double q = kind == T_MPMP ? r1 / r2 : r2 / r1;
    pr[0] = pq[0] = 1.0;
    for (int k = 1; k <= 2 * P + 1; k++) pr[k] = pr[k - 1] * r1, pq[k] = pq[k - 1] * q;
    for (int i = -p2; i <= p2; i++) {
        int ai = abs(i);
        for (int j = -p1; j <= p1; j++) {
            int aj = abs(j), am = abs(i - j);
            double F;
            if (kind == T_MPMP) F = pr[aj + am - ai] * pq[ai];
            else if (kind == T_LOCLOC) F = pr[ai + am - aj] * pq[ai];
            else F = pr[ai + aj - am] * pq[ai];
            cplx t = B[P + i - j] * F;
            for (int d = 0; d < nd; d++) EX(out, d, i, p2, nd) += EX(in, d, j, p1, nd) * t;
        }
    }
    free(v); free(B); free(pr); free(pq);
}

/* ------------------------------------------------------------------ */
/* engine (hfmm2d.c:619-1523), nexpc = 0                                 */
/* ------------------------------------------------------------------ */
typedef struct {
    int nd, ifcharge, ifdipole, ifpgh, ifpghtarg, ifnear;
    cplx zk;
    const double *src, *targ, *dv;
    const cplx *chg, *dip;
    cplx *pot, *grad, *hess, *pott, *gradt, *hesst;
} helm_job;

static void polar(double dx, double dy, double *r, double *th)
{
    *r = sqrt(dx * dx + dy * dy);
    *th = (dx == 0 && dy == 0) ? 0.0 : atan2(dy, dx);
}

static int helm_main(const otree *t, const olists *L, const helm_job *J, const int *nterms,
                     const double *rsc)
{
    int nd = J->nd, nl = t->nlevels, nb = t->nboxes;
    cplx zk = J->zk;
    double zi = cimag(zk);
    const double *bs = t->boxsize;
    int pmax = 0;
    for (int l = 0; l <= nl; l++) pmax = nterms[l] > pmax ? nterms[l] : pmax;
    /* Boxes of more than 16 wavelengths use the diagonal form in the
     * reference (wideband2d.c).  Its signatures of nsig >= 2 P + 1 points
     * make the circular convolution exact, so it equals the dense
     * translation below in exact arithmetic: the oracle translates every
     * level densely. */
    size_t per = (size_t)nd * (2 * pmax + 1);
    cplx *mp = calloc(per * (nb + 1), sizeof(cplx)), *lo = calloc(per * (nb + 1), sizeof(cplx));
    const cplx *c = J->ifcharge ? J->chg : NULL, *q = J->ifdipole ? J->dip : NULL;
#define MP(ib) (mp + per * (ib))
#define LC(ib) (lo + per * (ib))
#define S0(ib) (t->isrcse[2 * (ib)])
#define NS(ib) (t->isrcse[2 * (ib) + 1] - t->isrcse[2 * (ib)] + 1)
#define T0(ib) (t->itargse[2 * (ib)])
#define NT(ib) (t->itargse[2 * (ib) + 1] - t->itargse[2 * (ib)] + 1)
#define GATE(l) (zi * bs[(l)] < 40.0)
    size_t wksz = (size_t)(pmax + 16) * 8 + (size_t)nd * (2 * pmax + 8) * 8;

    /* step 1 + 2: formation (leaves) and list 4 */
    for (int l = 2; l <= nl; l++) {
        if (!GATE(l)) continue;
        int p = nterms[l];
        double rs = rsc[l];
#pragma omp parallel
        {
            cplx *v = malloc(sizeof(cplx) * (p + 12));
#pragma omp for schedule(dynamic, 16)
            for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++) {
                const double *ctr = t->centers + 2 * ib;
                if (t->nchild[ib] == 0 && NS(ib) > 0) {
                    for (int s = S0(ib) - 1; s < S0(ib) - 1 + NS(ib); s++) {
                        double r, th;
                        polar(J->src[2 * s] - ctr[0], J->src[2 * s + 1] - ctr[1], &r, &th);
                        o_jbessel(p + 1, zk * r, rs, v);
                        cplx zm = cexp(-I * th), zv = conj(zm);
                        if (c) to_pole_c(nd, zm, zv, MP(ib), v, c + nd * s, p);
                        if (q) to_pole_d(nd, zk, rs, zm, zv, MP(ib), v, q + nd * s, J->dv + 2 * nd * s, p);
                    }
                }
                int np = (J->ifpghtarg > 0 ? NT(ib) : 0) + (J->ifpgh > 0 ? NS(ib) : 0);
                if (np <= 0) continue;
                for (int k = 0; k < L->n4[ib]; k++) {
                    int jb = L->l4[L->mn4 * ib + k];
                    for (int s = S0(jb) - 1; s < S0(jb) - 1 + NS(jb); s++) {
                        double r, th;
                        polar(J->src[2 * s] - ctr[0], J->src[2 * s + 1] - ctr[1], &r, &th);
                        o_h2dall(p + 2, zk * r, rs, v);
                        cplx zm = cexp(-I * th), zv = conj(zm);
                        if (c) to_pole_c(nd, zm, zv, LC(ib), v, c + nd * s, p);
                        if (q) to_pole_d(nd, zk, 1.0 / rs, zm, zv, LC(ib), v, q + nd * s, J->dv + 2 * nd * s, p);
                    }
                }
            }
            free(v);
        }
    }
    /* step 3: M2M */
    for (int l = nl - 1; l >= 1; l--) {
        if (!GATE(l)) continue;
#pragma omp parallel for schedule(dynamic, 16)
        for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++)
            for (int k = 0; k < t->nchild[ib]; k++) {
                int jb = t->ichild[4 * ib + k];
                if (NS(jb) > 0)
                    translate(T_MPMP, nd, zk, rsc[l + 1], t->centers + 2 * jb, MP(jb), nterms[l + 1],
                              rsc[l], t->centers + 2 * ib, MP(ib), nterms[l]);
            }
    }
    /* step 4: M2L */
    for (int l = 2; l <= nl; l++) {
        if (!GATE(l)) continue;
#pragma omp parallel for schedule(dynamic, 16)
        for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++) {
            int np = (J->ifpghtarg > 0 ? NT(ib) : 0) + (J->ifpgh > 0 ? NS(ib) : 0);
            if (np <= 0) continue;
            for (int k = 0; k < L->n2[ib]; k++) {
                int jb = L->l2[L->mn2 * ib + k];
                translate(T_MPLOC, nd, zk, rsc[l], t->centers + 2 * jb, MP(jb), nterms[l], rsc[l],
                          t->centers + 2 * ib, LC(ib), nterms[l]);
            }
        }
    }
    /* step 5: L2L */
    for (int l = 1; l <= nl - 1; l++) {
        if (!GATE(l)) continue;
#pragma omp parallel for schedule(dynamic, 16)
        for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++) {
            int np = (J->ifpghtarg > 0 ? NT(ib) : 0) + (J->ifpgh > 0 ? NS(ib) : 0);
            if (np <= 0) continue;
            for (int k = 0; k < t->nchild[ib]; k++) {
                int jb = t->ichild[4 * ib + k];
                translate(T_LOCLOC, nd, zk, rsc[l], t->centers + 2 * ib, LC(ib), nterms[l], rsc[l + 1],
                          t->centers + 2 * jb, LC(jb), nterms[l + 1]);
            }
        }
    }
    /* steps 6 + 7: list 3 and local evaluation */
    for (int l = 0; l <= nl; l++) {
#pragma omp parallel
        {
            cplx *v = malloc(sizeof(cplx) * (pmax + 16));
            cplx *wk = malloc(sizeof(cplx) * wksz);
#pragma omp for schedule(dynamic, 16)
            for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++) {
                for (int pass = 0; pass < 2; pass++) {
                    int pg = pass ? J->ifpghtarg : J->ifpgh;
                    if (pg <= 0) continue;
                    int n0 = pass ? T0(ib) - 1 : S0(ib) - 1, n = pass ? NT(ib) : NS(ib);
                    const double *X = pass ? J->targ : J->src;
                    cplx *Po = pass ? J->pott : J->pot, *Gr = pass ? J->gradt : J->grad;
                    cplx *He = pass ? J->hesst : J->hess;
                    if (l <= nl - 1 && GATE(l + 1))
                        for (int k = 0; k < L->n3[ib]; k++) {
                            int jb = L->l3[L->mn3 * ib + k], p = nterms[l + 1];
                            const double *cj = t->centers + 2 * jb;
                            for (int s = n0; s < n0 + n; s++) {
                                double r, th;
                                polar(X[2 * s] - cj[0], X[2 * s + 1] - cj[1], &r, &th);
                                o_h2dall(p + 3, zk * r, rsc[l + 1], v);
                                eval_at(nd, zk, rsc[l + 1], MP(jb), p, v, th, pg, Po + nd * s,
                                        Gr + 2 * nd * s, He + 3 * nd * s, wk);
                            }
                        }
                    if (t->nchild[ib] == 0 && GATE(l)) {
                        int p = nterms[l];
                        const double *cb = t->centers + 2 * ib;
                        for (int s = n0; s < n0 + n; s++) {
                            double r, th;
                            polar(X[2 * s] - cb[0], X[2 * s + 1] - cb[1], &r, &th);
                            o_jbessel(p + 3, zk * r, rsc[l], v);
                            eval_at(nd, zk, 1.0 / rsc[l], LC(ib), p, v, th, pg, Po + nd * s,
                                    Gr + 2 * nd * s, He + 3 * nd * s, wk);
                        }
                    }
                }
            }
            free(v);
            free(wk);
        }
    }
    /* step 8: list 1 */
    double thresh = cabs(zk) * bs[0] * 0x1p-51;
    if (J->ifnear) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int ib = 1; ib <= nb; ib++)
            for (int k = 0; k < L->n1[ib]; k++) {
                int jb = L->l1[L->mn1 * ib + k], s0 = S0(jb) - 1;
                const cplx *cc = c ? c + nd * s0 : NULL, *qq = q ? q + nd * s0 : NULL;
                const double *dv = J->dv + 2 * nd * s0;
                if (J->ifpghtarg > 0)
                    oracle_h2d_direct(nd, zk, J->src + 2 * s0, NS(jb), cc, qq, dv,
                                      J->targ + 2 * (T0(ib) - 1), NT(ib), J->ifpghtarg,
                                      J->pott + nd * (T0(ib) - 1), J->gradt + 2 * nd * (T0(ib) - 1),
                                      J->hesst + 3 * nd * (T0(ib) - 1), thresh);
                if (J->ifpgh > 0)
                    oracle_h2d_direct(nd, zk, J->src + 2 * s0, NS(jb), cc, qq, dv,
                                      J->src + 2 * (S0(ib) - 1), NS(ib), J->ifpgh,
                                      J->pot + nd * (S0(ib) - 1), J->grad + 2 * nd * (S0(ib) - 1),
                                      J->hess + 3 * nd * (S0(ib) - 1), thresh);
            }
    }
    free(mp);
    free(lo);
    return 0;
}

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

/* hfmm2d_ndiv semantics (hfmm2d_ndiv.c:68-351) */
void orc_hfmm2d_ndiv_(const int *nd, const double *eps, const cplx *zk, const int *ns,
                      const double *sources, const int *ifcharge, const cplx *charge,
                      const int *ifdipole, const cplx *dipstr, const double *dipvec, int *iper,
                      const int *ifpgh, cplx *pot, cplx *grad, cplx *hess, const int *nt,
                      const double *targ, const int *ifpghtarg, cplx *pottarg, cplx *gradtarg,
                      cplx *hesstarg, const int *ndiv, const int *idivflag, const int *ifnear,
                      double *timeinfo, int *ier)
{
    otree t;
    olists L;
    *iper = 0;
    for (int i = 0; i < 8; i++) timeinfo[i] = 0.0;
    otree_build(sources, *ns, targ, *nt, *idivflag, *ndiv, 0, 51, 0, &t);
    olists_build(&t, &L);
    int n = *nd, nl = t.nlevels;
    int *nterms = malloc(sizeof(int) * (nl + 1));
    double *rsc = malloc(sizeof(double) * (nl + 1));
    for (int l = 0; l <= nl; l++) {
        cplx zt = (*zk) * t.boxsize[l] / (2.0 * PI);
        rsc[l] = cabs(zt) > 1.0 ? 1.0 : cabs(zt);
        int e;
        nterms[l] = o_h2dterms(t.boxsize[l], *zk, *eps, &e);
    }
    size_t cw = sizeof(cplx) * n;
    helm_job J;
    memset(&J, 0, sizeof(J));
    J.nd = n; J.ifcharge = *ifcharge; J.ifdipole = *ifdipole; J.ifpgh = *ifpgh;
    J.ifpghtarg = *ifpghtarg; J.ifnear = *ifnear; J.zk = *zk;
    J.src = gather(sources, 2 * sizeof(double), *ns, t.isrc);
    J.targ = gather(targ, 2 * sizeof(double), *nt, t.itarg);
    J.chg = *ifcharge ? gather(charge, cw, *ns, t.isrc) : NULL;
    J.dip = *ifdipole ? gather(dipstr, cw, *ns, t.isrc) : NULL;
    J.dv = *ifdipole ? gather(dipvec, 2 * sizeof(double) * n, *ns, t.isrc) : calloc(2 * n, sizeof(double));
    int S1 = *ns > 0 ? *ns : 1, T1 = *nt > 0 ? *nt : 1;
    J.pot = calloc((size_t)S1 * n, sizeof(cplx));
    J.grad = calloc((size_t)S1 * n * 2, sizeof(cplx));
    J.hess = calloc((size_t)S1 * n * 3, sizeof(cplx));
    J.pott = calloc((size_t)T1 * n, sizeof(cplx));
    J.gradt = calloc((size_t)T1 * n * 2, sizeof(cplx));
    J.hesst = calloc((size_t)T1 * n * 3, sizeof(cplx));
    int e = helm_main(&t, &L, &J, nterms, rsc);
    if (e == 0) {
        if (*ifpgh >= 1) scatter(pot, J.pot, cw, *ns, t.isrc);
        if (*ifpgh >= 2) scatter(grad, J.grad, 2 * cw, *ns, t.isrc);
        if (*ifpgh >= 3) scatter(hess, J.hess, 3 * cw, *ns, t.isrc);
        if (*ifpghtarg >= 1) scatter(pottarg, J.pott, cw, *nt, t.itarg);
        if (*ifpghtarg >= 2) scatter(gradtarg, J.gradt, 2 * cw, *nt, t.itarg);
        if (*ifpghtarg >= 3) scatter(hesstarg, J.hesst, 3 * cw, *nt, t.itarg);
    }
    free((void *)J.src); free((void *)J.targ); free((void *)J.chg); free((void *)J.dip);
    free((void *)J.dv); free(J.pot); free(J.grad); free(J.hess); free(J.pott); free(J.gradt);
    free(J.hesst); free(nterms); free(rsc);
    olists_free(&L);
    otree_free(&t);
    *ier = e;
}

void orc_hfmm2d_(const int *nd, const double *eps, const cplx *zk, const int *ns,
                 const double *sources, const int *ifcharge, const cplx *charge, const int *ifdipole,
                 const cplx *dipstr, const double *dipvec, int *iper, const int *ifpgh, cplx *pot,
                 cplx *grad, cplx *hess, const int *nt, const double *targ, const int *ifpghtarg,
                 cplx *pottarg, cplx *gradtarg, cplx *hesstarg, int *ier)
{
    int ndiv = o_lndiv(*eps, *ns, *nt), idiv = 0, ifnear = 1; /* hndiv2d: same table */
    double ti[8];
    orc_hfmm2d_ndiv_(nd, eps, zk, ns, sources, ifcharge, charge, ifdipole, dipstr, dipvec, iper,
                     ifpgh, pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg, hesstarg, &ndiv,
                     &idiv, &ifnear, ti, ier);
}
