/*
 * obh.c -- biharmonic direct sums (test infrastructure, see oracle.h).
 *
 * Restates bh2d_direct{c,d,cd}{p,g} (c_translation/src/bhkernels2d.c:30-398):
 * with z = t - s and w = 1/z, per pair (skipped when |z| <= thresh)
 *   charges (c1, c2):    vel += 2 c1 log|z| + c2 conj(w) z
 *                        g1 += c1 w,  g2 += c2 conj(w),  g3 += c1 conj(w) - c2 conj(w)^2 z
 *   dipoles (d1, d2, d3): vel += d1 w + d3 conj(w) + d2 conj(w)^2 z
 *                        g1 -= d1 w^2,  g2 += d2 conj(w)^2,
 *                        g3 -= d3 conj(w)^2 + 2 d2 conj(w)^3 z
 * so that g1 + g2 = d vel / dz and g3 = d vel / d conj(z) (the tests check
 * this by finite differences).  charges (nd,2,ns), dipoles (nd,3,ns),
 * vel (nd,nt), grad (nd,3,nt), incremented.
 */
#include <complex.h>
#include <math.h>
#include <stddef.h>

#include "oracle.h"

void oracle_bh2d_direct(int nd, const double *src, int ns, int ifc, const cplx *chg, int ifd,
                        const cplx *dip, const double *targ, int nt, int pg, cplx *vel,
                        cplx *grad, double thresh) {
    for (int i = 0; i < nt; ++i) {
        const cplx zt = targ[2 * i] + I * targ[2 * i + 1];
        for (int d = 0; d < nd; ++d) {
            cplx v = 0, g1 = 0, g2 = 0, g3 = 0;
            for (int j = 0; j < ns; ++j) {
                const cplx z = zt - (src[2 * j] + I * src[2 * j + 1]);
                if (cabs(z) <= thresh) continue;
                const cplx w = 1.0 / z, cw = conj(w);
                if (ifc == 1) {
                    const cplx c1 = chg[((size_t)j * 2 + 0) * nd + d];
                    const cplx c2 = chg[((size_t)j * 2 + 1) * nd + d];
                    v += 2.0 * c1 * log(cabs(z)) + c2 * cw * z;
                    g1 += c1 * w;
                    g2 += c2 * cw;
                    g3 += c1 * cw - c2 * cw * cw * z;
                }
                if (ifd == 1) {
                    const cplx d1 = dip[((size_t)j * 3 + 0) * nd + d];
                    const cplx d2 = dip[((size_t)j * 3 + 1) * nd + d];
                    const cplx d3 = dip[((size_t)j * 3 + 2) * nd + d];
                    v += d1 * w + d3 * cw + d2 * cw * cw * z;
                    g1 -= d1 * w * w;
                    g2 += d2 * cw * cw;
                    g3 -= d3 * cw * cw + 2.0 * d2 * cw * cw * cw * z;
                }
            }
            vel[(size_t)i * nd + d] += v;
            if (pg >= 2) {
                grad[((size_t)i * 3 + 0) * nd + d] += g1;
                grad[((size_t)i * 3 + 1) * nd + d] += g2;
                grad[((size_t)i * 3 + 2) * nd + d] += g3;
            }
        }
    }
}
