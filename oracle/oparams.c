/*
 * oparams.c -- oracle restatement of the parameter rules (test
 * infrastructure; see oracle.h).
 *   leaf size       src/laplace/lndiv2d.f:17-35 (same table as
 *                   c_translation/src/hndiv2d.c:31-47)
 *   Laplace order   c_translation/src/l2dterms.c:26-68
 *   FFT length      c_translation/src/next235.c:10-32
 */
#include "oracle.h"
#include <math.h>

int o_lndiv(double eps, int ns, int nt)
{
    static const double lim[] = {0.5, 0.5e-1, 0.5e-2, 0.5e-3, 0.5e-6, 0.5e-9, 0.5e-12, 0.5e-15};
    static const int val[] = {3, 5, 8, 10, 15, 20, 25, 45};
    for (int i = 0; i < 8; i++)
        if (eps >= lim[i]) return val[i];
    return ns + nt;
}

/* smallest j >= 2 with (sqrt(2)/2)^j / 1.5^(j+1) < eps; 1 if none up to 100 */
int o_l2dterms(double eps)
{
    for (int j = 2; j <= 100; j++) {
        double v = pow(sqrt(2.0) / 2.0, (double)j) * (1.0 / pow(1.5, (double)(j + 1)));
        if (v < eps) return j;
    }
    return 1;
}

int o_next235(double base)
{
    int n = 2 * (int)(base / 2.0 + 0.9999);
    if (n <= 0) n = 2;
    for (;; n += 2) {
        int m = n;
        while (m % 2 == 0) m /= 2;
        while (m % 3 == 0) m /= 3;
        while (m % 5 == 0) m /= 5;
        if (m == 1) return n;
    }
}

int oracle_lndiv(double eps, int ns, int nt) { return o_lndiv(eps, ns, nt); }
int oracle_l2dterms(double eps) { return o_l2dterms(eps); }
int oracle_next235(double b) { return o_next235(b); }
