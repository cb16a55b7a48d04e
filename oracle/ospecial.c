/*
 * ospecial.c -- oracle special functions (test infrastructure; see oracle.h).
 *
 * H0^(1), H1^(1) of complex argument: our own evaluation, NOT a restatement
 * of c_translation/src/hank103.c (whose fitted tables we do not reproduce).
 * Published methods, by region of z (upper half plane; the lower half plane
 * uses H^(1)(z) = 2 J(z) - conj(H^(1)(conj z))):
 *   |z| <= 4          ascending series (Abramowitz & Stegun 9.1.10-9.1.11)
 *   4 < |z| < 20      Laplace-type integral
 *                     H_nu(z) = sqrt(2/(pi z)) e^{i(z - nu pi/2 - pi/4)} / Gamma(nu+1/2)
 *                               * int_0^inf e^{-t} t^{nu-1/2} (1 + i t/(2z))^{nu-1/2} dt
 *                     by a 20-point Gauss-Laguerre rule (gl20.h, weight t^-1/2 e^-t)
 *   |z| >= 20         Hankel's asymptotic expansion (A&S 9.2.7, 9.2.9)
 * Pinned in tests/test_oracle_helm.py against mpmath at 50 digits across the
 * regions (parity pinned by mathematics, not by the reference's tables).
 *
 * jbessel2d (c_translation/src/cdjseval2d.c:11-164): scaled J_n(z)/rscale^n by
 * the same Miller recurrence with the same start-index rule and
 * normalisation; h2dall (c_translation/src/h2dcommon.c:35-90): scaled
 * H_n(z) rscale^n by upward recurrence from H0, H1; h2dterms
 * (c_translation/src/h2dterms.c:22-78): expansion order rule.
 */
#include "oracle.h"
#include "gl20.h"
#include <math.h>
#include <stdlib.h>

static const double PI = 3.14159265358979323846;
static const double EULER = 0.57721566490153286061;

static void series01(cplx z, cplx *h0, cplx *h1)
{
    cplx w = z * z / 4.0, lz = clog(z / 2.0);
    cplx t0 = 1.0, t1 = 1.0;          /* w^k/(k!)^2, w^k/(k!(k+1)!) */
    cplx j0 = 0, j1 = 0, s0 = 0, s1 = 0;
    double hk = 0.0, psi1 = -EULER, psi2 = 1.0 - EULER; /* H_k, psi(k+1), psi(k+2) */
    for (int k = 0; k < 60; k++) {
        double sg = (k & 1) ? -1.0 : 1.0;
        j0 += sg * t0;
        j1 += sg * t1;
        s0 += -sg * hk * t0;                       /* sum (-1)^{k+1} H_k w^k/(k!)^2 */
        s1 += sg * (psi1 + psi2) * t1;
        if (cabs(t0) < 1e-18 * (cabs(j0) + 1e-300) && k > 2) break;
        t0 *= w / ((double)(k + 1) * (k + 1));
        t1 *= w / ((double)(k + 1) * (k + 2));
        hk += 1.0 / (k + 1);
        psi1 += 1.0 / (k + 1);
        psi2 += 1.0 / (k + 2);
    }
    j1 *= z / 2.0;
    cplx y0 = (2.0 / PI) * ((lz + EULER) * j0 + s0);
    cplx y1 = -2.0 / (PI * z) + (2.0 / PI) * lz * j1 - (1.0 / PI) * (z / 2.0) * s1;
    *h0 = j0 + I * y0;
    *h1 = j1 + I * y1;
}

/* Miller backward recurrence; J[n] for n = 0..nmax (unscaled) */
static void miller(cplx z, int nmax, cplx *J)
{
    int top = nmax + 20 + (int)(cabs(z) * 1.5);
    cplx *f = malloc(sizeof(cplx) * (top + 2));
    cplx zi = 1.0 / z;
    f[top + 1] = 0;
    f[top] = 1e-300;
    for (int n = top; n >= 1; n--) {
        f[n - 1] = 2.0 * n * zi * f[n] - f[n + 1];
        if (cabs(f[n - 1]) > 1e250) {
            for (int m = n - 1; m <= top + 1; m++) f[m] *= 1e-250;
        }
    }
    /* e^{-iz} = J0 + 2 sum (-i)^n J_n for Im z >= 0, e^{iz} = J0 + 2 sum i^n J_n
     * otherwise: the side where the sum does not cancel (cdjseval2d.c:128-141) */
    cplx u = cimag(z) >= 0 ? -I : I;
    cplx s = 0, ph = 1.0;
    for (int n = 1; n <= top; n++) {
        ph *= u;
        s += ph * f[n];
    }
    cplx lam = (f[0] + 2.0 * s) * cexp(-u * z);
    for (int n = 0; n <= nmax; n++) J[n] = f[n] / lam;
    free(f);
}

static void laguerre01(cplx z, cplx *h0, cplx *h1)
{
    /* Gamma(1/2) = sqrt(pi), Gamma(3/2) = sqrt(pi)/2; nu = 1 uses t * (1+..)^(1/2) */
    cplx i0 = 0, i1 = 0, a = I / (2.0 * z);
    for (int k = 0; k < GL20_N; k++) {
        cplx s = csqrt(1.0 + a * GL20_X[k]);
        i0 += GL20_W[k] / s;
        i1 += GL20_W[k] * GL20_X[k] * s;
    }
    cplx pre = csqrt(2.0 / (PI * z)) / sqrt(PI);
    *h0 = pre * cexp(I * (z - PI / 4)) * i0;
    *h1 = pre * cexp(I * (z - 3 * PI / 4)) * i1 * 2.0;
}

static void asym01(cplx z, cplx *h0, cplx *h1)
{
    cplx pre = csqrt(2.0 / (PI * z));
    cplx s0 = 0, s1 = 0, t0 = 1.0, t1 = 1.0, iz = I / z;
    for (int k = 0; k < 40; k++) {
        s0 += t0;
        s1 += t1;
        double m0 = (4.0 * 0 - (2.0 * k + 1) * (2.0 * k + 1)) / (8.0 * (k + 1));
        double m1 = (4.0 * 1 - (2.0 * k + 1) * (2.0 * k + 1)) / (8.0 * (k + 1));
        cplx n0 = t0 * m0 * iz, n1 = t1 * m1 * iz;
        if (cabs(n0) > cabs(t0) && k > 4) break;    /* stop at the smallest term */
        t0 = n0;
        t1 = n1;
        if (cabs(t0) < 1e-18 && cabs(t1) < 1e-18) { s0 += t0; s1 += t1; break; }
    }
    *h0 = pre * cexp(I * (z - PI / 4)) * s0;
    *h1 = pre * cexp(I * (z - 3 * PI / 4)) * s1;
}

static void hank_upper(cplx z, cplx *h0, cplx *h1)
{
    double a = cabs(z);
    if (a <= 4.0) series01(z, h0, h1);
    else if (a < 20.0) laguerre01(z, h0, h1);
    else asym01(z, h0, h1);
}

void o_hank01(cplx z, cplx *h0, cplx *h1)
{
    if (cimag(z) >= 0) {
        hank_upper(z, h0, h1);
        return;
    }
    /* H^(1)(z) = 2 J(z) - conj(H^(1)(conj z)) for real order */
    cplx g0, g1, J[3];
    hank_upper(conj(z), &g0, &g1);
    miller(z, 2, J);
    *h0 = 2.0 * J[0] - conj(g0);
    *h1 = 2.0 * J[1] - conj(g1);
}

void oracle_hank01(const double *z, double *out)
{
    cplx h0, h1;
    o_hank01(z[0] + I * z[1], &h0, &h1);
    out[0] = creal(h0); out[1] = cimag(h0); out[2] = creal(h1); out[3] = cimag(h1);
}

/* scaled H_n(z) rscale^n, n = 0..nterms (h2dcommon.c:35-90) */
void o_h2dall(int nterms, cplx z, double rscale, cplx *hvec)
{
    if (cabs(z) < 1e-200) {
        for (int i = 0; i <= nterms; i++) hvec[i] = 0;
        return;
    }
    cplx h0, h1;
    o_hank01(z, &h0, &h1);
    hvec[0] = h0;
    if (nterms >= 1) hvec[1] = h1 * rscale;
    double s2 = rscale * rscale;
    cplx zi = rscale / z;
    for (int i = 1; i <= nterms - 1; i++) hvec[i + 1] = (zi * (2.0 * i)) * hvec[i] - s2 * hvec[i - 1];
}

/* scaled J_n(z)/rscale^n, n = 0..nterms (cdjseval2d.c:11-164): upward
 * recurrence from nterms to find the start index where |f| exceeds 1e20,
 * downward recurrence with rescaling by 1e-40, uniform scaling by rscale,
 * normalisation by the generating function. */
void o_jbessel(int nterms, cplx z, double rscale, cplx *fjs)
{
    if (cabs(z) < 1e-200) {
        fjs[0] = 1.0;
        for (int i = 1; i <= nterms; i++) fjs[i] = 0;
        return;
    }
    const double ub2 = 1e40, ub2i = 1e-40;
    cplx zi = 1.0 / z, fn = 1.0, fnm1 = 0.0;
    int ntop = nterms + 10000;
    for (int i = nterms; i <= nterms + 10000; i++) {
        cplx fnp1 = (2.0 * i) * zi * fn - fnm1;
        if (creal(fnp1) * creal(fnp1) + cimag(fnp1) * cimag(fnp1) > ub2) { ntop = i + 1; break; }
        fnm1 = fn;
        fn = fnp1;
    }
    cplx *f = malloc(sizeof(cplx) * (ntop + 2));
    char *sc = calloc(ntop + 2, 1);
    f[ntop] = 0;
    f[ntop - 1] = 1.0;
    for (int i = ntop - 1; i >= 1; i--) {
        cplx t = (2.0 * i) * zi * f[i] - f[i + 1];
        f[i - 1] = t;
        if (creal(t) * creal(t) + cimag(t) * cimag(t) > ub2) {
            f[i] *= ub2i;
            f[i - 1] *= ub2i;
            sc[i] = 1;
        }
    }
    double sinv = 1.0 / rscale, tot = 1.0;
    for (int i = 1; i <= ntop; i++) {
        tot *= sinv;
        if (sc[i - 1]) tot *= ub2i;
        f[i] *= tot;
    }
    cplx zm = cimag(z) < 0 ? I : -I, zs = 1.0;
    for (int i = 0; i < ntop % 4; i++) zs *= zm;
    cplx zmi = 1.0 / zm, psi = 0;
    for (int i = ntop; i >= 1; i--) {
        psi = rscale * psi + f[i] * zs;
        zs *= zmi;
    }
    psi = 2.0 * psi * rscale + f[0];
    cplx scale = (cimag(z) < 0 ? cexp(I * z) : cexp(-I * z)) / psi;
    for (int i = 0; i <= nterms; i++) fjs[i] = f[i] * scale;
    free(f);
    free(sc);
}

/* expansion order for a box of size bsize (h2dterms.c:22-78) */
int o_h2dterms(double bsize, cplx zk, double eps, int *ier)
{
    const int ntmax = 50000;
    double pi2 = 8 * atan(1.0), rscale = 1.0;
    cplx zb = zk * bsize;
    if (cabs(zb) < pi2) rscale = cabs(zb);
    cplx *hf = malloc(sizeof(cplx) * (ntmax + 101)), *jf = malloc(sizeof(cplx) * (ntmax + 101));
    o_h2dall(ntmax, zb * 1.5, rscale, hf);
    o_jbessel(ntmax, zb * sqrt(2.0) / 2.0, rscale, jf);
    double x0 = cabs(jf[0] * hf[0]) + cabs(jf[1] * hf[1]);
    int nt = 10001;
    *ier = 13;
    for (int j = 2; j <= ntmax; j++) {
        double x = (cabs(jf[j] * hf[j]) + cabs(jf[j - 1] * hf[j - 1])) * cabs(hf[0]);
        if (x < eps * x0) { nt = j + 1; *ier = 0; break; }
    }
    free(hf);
    free(jf);
    return nt;
}

int oracle_h2dterms(double bsize, const double *zk, double eps)
{
    int ier;
    return o_h2dterms(bsize, zk[0] + I * zk[1], eps, &ier);
}

void oracle_jbessel(int nterms, const double *z, double rscale, double *out)
{
    o_jbessel(nterms, z[0] + I * z[1], rscale, (cplx *)out);
}

void oracle_h2dall(int nterms, const double *z, double rscale, double *out)
{
    o_h2dall(nterms, z[0] + I * z[1], rscale, (cplx *)out);
}
