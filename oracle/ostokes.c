/*
 * ostokes.c -- Stokes direct sums (test infrastructure, see oracle.h).
 *
 * Restates st2ddirectstokg / st2ddirectstokstrsg
 * (c_translation/src/stokkernels2d.c:11-173): the velocity u (nd,2,nt),
 * pressure p (nd,nt) and velocity gradient g (nd,2,2,nt): g(a,b) =
 * d u_b / d x_a at ((i*2+b)*2+a)*nd + d) of Stokeslets f and stresslets
 * (mu, nu), each
 * output incremented, pairs with r^2 < thresh^2 skipped.  Per pair, with
 * x = t - s and r2 = |x|^2 (the reference's formulas written as the
 * tensors they expand):
 *   Stokeslet   u_a += -f_a log(r2)/4 + x_a (x.f) / (2 r2)
 *               p += (x.f) / r2
 *   stresslet   u_a += -2 x_a (x.mu)(x.nu) / r2^2
 *               p += -4 (x.mu)(x.nu)/r2^2 + 2 mu.nu/r2
 * g is written out as its derivative, so the finite-difference and
 * divergence checks of the tests pin it independently of the reference.
 */
#include <math.h>
#include <stddef.h>

#include "oracle.h"

static void stokeslet(double x1, double x2, double f1, double f2, double u[2], double g[4],
                      double *p) {
    const double r2 = x1 * x1 + x2 * x2, xf = x1 * f1 + x2 * f2;
    const double lg = log(r2) / 4.0;
    u[0] += -f1 * lg + x1 * xf / (2 * r2);
    u[1] += -f2 * lg + x2 * xf / (2 * r2);
    /* g(b,a) = d/dx_b of u_a, in slot a*2 + b.  d/dx_b of -f_a log(r2)/4 = -f_a x_b / (2 r2);
       d/dx_b of x_a xf / (2 r2) = (delta_ab xf + x_a f_b) / (2 r2) - x_a xf x_b / r2^2 */
    const double x[2] = {x1, x2}, f[2] = {f1, f2};
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a)
            g[a * 2 + b] += -f[a] * x[b] / (2 * r2) + ((a == b) * xf + x[a] * f[b]) / (2 * r2) -
                            x[a] * xf * x[b] / (r2 * r2);
    *p += xf / r2;
}

static void stresslet(double x1, double x2, double m1, double m2, double n1, double n2,
                      double u[2], double g[4], double *p) {
    const double r2 = x1 * x1 + x2 * x2, r4 = r2 * r2;
    const double xm = x1 * m1 + x2 * m2, xn = x1 * n1 + x2 * n2, q = xm * xn;
    u[0] += -2 * x1 * q / r4;
    u[1] += -2 * x2 * q / r4;
    /* d/dx_b of -2 x_a q / r4 = -2 (delta_ab q + x_a (m_b xn + n_b xm)) / r4
                                 + 8 x_a q x_b / r2^3 */
    const double x[2] = {x1, x2}, m[2] = {m1, m2}, n[2] = {n1, n2};
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a)
            g[a * 2 + b] += -2 * ((a == b) * q + x[a] * (m[b] * xn + n[b] * xm)) / r4 +
                            8 * x[a] * q * x[b] / (r4 * r2);
    *p += -4 * q / r4 + 2 * (m1 * n1 + m2 * n2) / r2;
}

void oracle_st2d_direct(int nd, const double *src, int ns, int ifstok, const double *stok,
                        int istress, const double *mu, const double *nu, const double *targ,
                        int nt, double *pot, double *pre, double *grad, double thresh) {
    const double th2 = thresh * thresh;
    for (int i = 0; i < nt; ++i)
        for (int d = 0; d < nd; ++d) {
            double u[2] = {0, 0}, g[4] = {0, 0, 0, 0}, p = 0;
            for (int j = 0; j < ns; ++j) {
                const double x1 = targ[2 * i] - src[2 * j], x2 = targ[2 * i + 1] - src[2 * j + 1];
                if (x1 * x1 + x2 * x2 < th2) continue;
                const size_t k = (size_t)j * 2 * nd + d;
                if (ifstok == 1) stokeslet(x1, x2, stok[k], stok[k + nd], u, g, &p);
                if (istress == 1)
                    stresslet(x1, x2, mu[k], mu[k + nd], nu[k], nu[k + nd], u, g, &p);
            }
            const size_t o = (size_t)i * 2 * nd + d;
            pot[o] += u[0];
            pot[o + nd] += u[1];
            pre[(size_t)i * nd + d] += p;
            for (int c = 0; c < 4; ++c) grad[((size_t)i * 4 + c) * nd + d] += g[c];
        }
}
