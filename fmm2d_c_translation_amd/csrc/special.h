// special.h -- FP64 Hankel/Bessel functions for device code.
//
// H0^(1), H1^(1) of complex argument, our own evaluation by region:
//   |z| <= 4          ascending series (Abramowitz & Stegun 9.1.10-9.1.11);
//                     for a real positive argument the same series in real
//                     arithmetic from tabulated coefficients
//                     (hank_series_real)
//   4 < |z| < 20      Laplace-type integral
//                       H_nu(z) = sqrt(2/(pi z)) e^{i(z - nu pi/2 - pi/4)} / Gamma(nu+1/2)
//                                 * int_0^inf e^{-t} t^{nu-1/2} (1 + i t/(2z))^{nu-1/2} dt
//                     with a 20-point Gauss-Laguerre rule (gl20.h)
//   |z| >= 20         Hankel's asymptotic expansion (A&S 9.2.7)
// in the upper half plane; below it H^(1)(z) = 2 J(z) - conj(H^(1)(conj z)).
// A positive real argument above 4 takes the modulus form instead
// (hank_modulus_real: P_nu, Q_nu fitted in 4/x, hank_asym.h), in real
// arithmetic.
// Scaled sequences as in the reference (c_translation/src/h2dcommon.c:35-90,
// cdjseval2d.c:11-164): H_n(z) rscale^n by upward recurrence, J_n(z)/rscale^n
// by Miller's downward recurrence normalised through the generating
// function (Horner form, so no power of rscale is ever formed).
#pragma once

#include "common.h"
#include "gl20.h"
#include "hank_asym.h"
#include "jy_series.h"

namespace fmm {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEuler = 0.57721566490153286061;

FMM_HD cd cexp_(cd z) {
  double e = exp(z.x);
  return cd{e * cos(z.y), e * sin(z.y)};
}
FMM_HD cd clog_(cd z) { return cd{log(hypot(z.x, z.y)), atan2(z.y, z.x)}; }
FMM_HD cd csqrt_(cd z) {
  double r = hypot(z.x, z.y);
  if (r == 0.0) return cd{0, 0};
  double a = sqrt(0.5 * (r + fabs(z.x)));
  double b = 0.5 * z.y / a;
  if (z.x >= 0) return cd{a, b};
  return z.y >= 0 ? cd{fabs(b), a} : cd{fabs(b), -a};
}
FMM_HD double cabs_(cd z) { return hypot(z.x, z.y); }
FMM_HD cd ci() { return cd{0, 1}; }
FMM_HD cd mul_i(cd z) { return cd{-z.y, z.x}; }

FMM_HD void hank_series(cd z, cd& h0, cd& h1) {
  cd w = z * z * 0.25, lz = clog_(z * 0.5);
  cd t0 = mk(1, 0), t1 = mk(1, 0), j0 = mk(0, 0), j1 = mk(0, 0), s0 = mk(0, 0), s1 = mk(0, 0);
  double hk = 0.0, p1 = -kEuler, p2 = 1.0 - kEuler;
  for (int k = 0; k < 40; ++k) {
    double sg = (k & 1) ? -1.0 : 1.0;
    j0 += t0 * sg;
    j1 += t1 * sg;
    s0 += t0 * (-sg * hk);
    s1 += t1 * (sg * (p1 + p2));
    if (k > 2 && norm2(t0) < 1e-36 * (norm2(j0) + 1e-300)) break;
    t0 = t0 * w / ((double)(k + 1) * (k + 1));
    t1 = t1 * w / ((double)(k + 1) * (k + 2));
    hk += 1.0 / (k + 1);
    p1 += 1.0 / (k + 1);
    p2 += 1.0 / (k + 2);
  }
  j1 = j1 * (z * 0.5);
  cd y0 = ((lz + mk(kEuler, 0)) * j0 + s0) * (2.0 / kPi);
  cd y1 = -(inv(z) * (2.0 / kPi)) + (lz * j1) * (2.0 / kPi) - (z * 0.5) * s1 * (1.0 / kPi);
  h0 = j0 + mul_i(y0);
  h1 = j1 + mul_i(y1);
}

// Tier of a series length (jy_series.h), -1 for any other length
template <int N>
constexpr int jy_tier() {
  return N == JY_NTERMS0 ? 0 : N == JY_NTERMS1 ? 1 : N == JY_NTERMS2 ? 2 : N == JY_NTERMS3 ? 3 : -1;
}
// Polynomial c[0..n) at w and its derivative: Horner with the derivative
// chain beside it, its first step peeled (it starts from the top coefficient)
template <int n>
FMM_HD void horner_d(const double* c, double w, double& v, double& dv) {
  static_assert(n >= 2, "two terms at least");
  dv = c[n - 1];
  v = fma(c[n - 1], w, c[n - 2]);
#pragma unroll
  for (int k = n - 3; k >= 0; --k) {
    dv = fma(dv, w, v);
    v = fma(v, w, c[k]);
  }
}

// H0 = J0 + i Y0 and H1 = J1 + i Y1 of a real argument 0 < x <= JY_XMAX
// of the chosen N (jy_series.h), given xi = 1/x: the ascending series of hank_series in
// real arithmetic, J0 = P(w) and the Y0 remainder Q(w) in w = x^2/4 by
// Horner with their derivatives (for the four tier lengths, the tier's
// economised P and Q), J1 = -P'(w) x/2 and
// Y1 = -(2/pi) [J0/x - (ln(x/2) + gamma) J1 + Q'(w) x/2]
// (Abramowitz & Stegun 9.1.10, 9.1.13 and Y1 = -Y0').
// Core = true: x/2 must be a positive normal double (the table log,
// log_tab, without special cases).
// P(w), P'(w), Q(w), Q'(w) of series length N (the tier's economised forms
// for the four tier lengths)
template <int N>
FMM_HD void jy_pq(double w, double& p, double& dp, double& q, double& dq) {
  constexpr int T = jy_tier<N>();
  if constexpr (T >= 0) {
    // the tier's economised P and Q (x <= JY_XMAX<T>)
    horner_d<JY_TIER[T].n[0]>(JY_TIER[T].c[0], w, p, dp);
    horner_d<JY_TIER[T].n[1]>(JY_TIER[T].c[1], w, q, dq);
  } else if constexpr (N >= 2) {
    // the series to N terms
    horner_d<N>(JY_JA, w, p, dp);
    horner_d<N>(JY_YC, w, q, dq);
  } else {
    p = dp = q = dq = 0.0;  // no series of fewer than two terms
  }
}

template <int N, bool Core = false>
FMM_HD void hank_series_real_w(double x, double xi, double w, cd& h0, cd& h1) {
  double p, dp, q, dq;
  jy_pq<N>(w, p, dp, q, dq);
  const double hx = 0.5 * x, lg = (Core ? log_tab(hx) : log_pos(hx)) + kEuler;
  const double j1 = -dp * hx;
  const double y0 = (lg * p + q) * (2.0 / kPi);
  const double y1 = -(2.0 / kPi) * (p * xi - lg * j1 + dq * hx);
  h0 = cd{p, y0};
  h1 = cd{j1, y1};
}
// w = x^2/4 computed here
template <int N, bool Core = false>
FMM_HD void hank_series_real(double x, double xi, cd& h0, cd& h1) {
  hank_series_real_w<N, Core>(x, xi, x * x * 0.25, h0, h1);
}
template <int N>
FMM_HD void hank_series_real(double x, cd& h0, cd& h1) {
  hank_series_real<N>(x, 1.0 / x, h0, h1);
}

FMM_HD void hank_laguerre(cd z, cd& h0, cd& h1) {
  cd a = mul_i(inv(z * 2.0)), i0 = mk(0, 0), i1 = mk(0, 0);
  for (int k = 0; k < GL20_N; ++k) {
    cd s = csqrt_(mk(1, 0) + a * GL20_X[k]);
    i0 += inv(s) * GL20_W[k];
    i1 += s * (GL20_W[k] * GL20_X[k]);
  }
  cd pre = csqrt_(inv(z) * (2.0 / kPi)) * (1.0 / sqrt(kPi));
  h0 = pre * cexp_(mul_i(z - mk(kPi / 4, 0))) * i0;
  h1 = pre * cexp_(mul_i(z - mk(3 * kPi / 4, 0))) * i1 * 2.0;
}

FMM_HD void hank_asym(cd z, cd& h0, cd& h1) {
  cd pre = csqrt_(inv(z) * (2.0 / kPi));
  cd iz = mul_i(inv(z)), t0 = mk(1, 0), t1 = mk(1, 0), s0 = mk(0, 0), s1 = mk(0, 0);
  for (int k = 0; k < 40; ++k) {
    s0 += t0;
    s1 += t1;
    double m0 = (0.0 - (2.0 * k + 1) * (2.0 * k + 1)) / (8.0 * (k + 1));
    double m1 = (4.0 - (2.0 * k + 1) * (2.0 * k + 1)) / (8.0 * (k + 1));
    cd n0 = t0 * iz * m0, n1 = t1 * iz * m1;
    if (k > 4 && norm2(n0) > norm2(t0)) break;
    t0 = n0;
    t1 = n1;
    if (norm2(t0) < 1e-36 && norm2(t1) < 1e-36) {
      s0 += t0;
      s1 += t1;
      break;
    }
  }
  h0 = pre * cexp_(mul_i(z - mk(kPi / 4, 0))) * s0;
  h1 = pre * cexp_(mul_i(z - mk(3 * kPi / 4, 0))) * s1;
}

// H0, H1 of a real x > 4 in the modulus form of Hankel's expansion
// (A&S 9.2.5-9.2.6): H_nu = sqrt(2/(pi x)) (P_nu + i Q_nu) e^{i(x - (2nu+1)pi/4)},
// P_nu, Q_nu polynomials in s = 8/x - 1 (hank_asym.h).  The phase factor
// comes from sin x and cos x, e^{i(x - pi/4)} = ((cos x + sin x) + i (sin x -
// cos x)) / sqrt 2, so no rounded x - pi/4 enters it; e^{i(x - 3pi/4)} is
// -i times that.
template <int n>
FMM_HD double horner_c(const double (&c)[n], double s) {
  double v = c[n - 1];
#pragma unroll
  for (int k = n - 2; k >= 0; --k) v = fma(v, s, c[k]);
  return v;
}
FMM_HD void hank_modulus_real(double x, cd& h0, cd& h1) {
  const double u = 4.0 / x, s = fma(2.0, u, -1.0);
  const double p0 = horner_c(HA_P0, s), q0 = horner_c(HA_Q0, s);
  const double p1 = horner_c(HA_P1, s), q1 = horner_c(HA_Q1, s);
  double sn, cs;
  sincos(x, &sn, &cs);
  const double a = sqrt(u * (0.5 / kPi));  // sqrt(2 / (pi x))
  const double cr = (cs + sn) * (a * 0.70710678118654752440), ci = (sn - cs) * (a * 0.70710678118654752440);
  h0 = cd{p0 * cr - q0 * ci, p0 * ci + q0 * cr};
  h1 = cd{p1 * ci + q1 * cr, q1 * ci - p1 * cr};
}

FMM_HD void hank_upper(cd z, cd& h0, cd& h1) {
  double a = cabs_(z);
  // a positive real argument: the real series up to 4 (same function, real
  // arithmetic, a fixed number of terms), the modulus form beyond
  if (z.y == 0.0 && z.x > 0.0) {
    if (z.x <= JY_XMAX3) hank_series_real<JY_NTERMS3>(z.x, h0, h1);
    else hank_modulus_real(z.x, h0, h1);
  } else if (a <= 4.0) hank_series(z, h0, h1);
  else if (a < 20.0) hank_laguerre(z, h0, h1);
  else hank_asym(z, h0, h1);
}

// J0, J1 by Miller's recurrence with generating-function normalisation
FMM_HD void bessel_j01(cd z, cd& j0, cd& j1) {
  int top = 24 + (int)(cabs_(z) * 1.5);
  cd zi = inv(z), fp = mk(0, 0), f = mk(1e-300, 0), s = mk(0, 0);
  cd u = z.y >= 0 ? mk(0, -1) : mk(0, 1);  // -i above the axis, +i below
  // phase u^n for n = top .. 1, built downward from u^top
  cd ph = mk(1, 0);
  for (int k = 0; k < (top & 3); ++k) ph = ph * u;
  cd ui = conj(u);
  cd f1 = mk(0, 0);
  for (int n = top; n >= 1; --n) {
    s += ph * f;
    ph = ph * ui;
    cd fm = zi * f * (2.0 * n) - fp;
    fp = f;
    f = fm;
    if (n == 1) f1 = fp;
    if (norm2(f) > 1e300) {
      f = f * 1e-250; fp = fp * 1e-250; s = s * 1e-250; f1 = f1 * 1e-250;
    }
  }
  cd lam = (f + s * 2.0) * cexp_(-(u * z));
  cd li = inv(lam);
  j0 = f * li;
  j1 = f1 * li;
}

FMM_HD void hank01(cd z, cd& h0, cd& h1) {
  if (z.y >= 0) {
    hank_upper(z, h0, h1);
    return;
  }
  cd g0, g1, j0, j1;
  hank_upper(conj(z), g0, g1);
  bessel_j01(z, j0, j1);
  h0 = j0 * 2.0 - conj(g0);
  h1 = j1 * 2.0 - conj(g1);
}

// Start index of the downward recurrence (cdjseval2d.c:55-71): upward from
// nterms until |f|^2 exceeds 1e40.
FMM_HD int miller_top(int nterms, cd z) {
  cd zi = inv(z), fn = mk(1, 0), fm = mk(0, 0);
  for (int i = nterms; i <= nterms + 10000; ++i) {
    cd fp = zi * fn * (2.0 * i) - fm;
    if (norm2(fp) > 1e40) return i + 1;
    fm = fn;
    fn = fp;
  }
  return nterms + 10000;
}

// Scaled J_n(z)/rs^n into out[n*stride], n = 0..nmax, with the downward
// recurrence started where an ntop-term evaluation starts (nmax <= ntop):
// the stored values are bit for bit those of jvals(ntop, ...).
FMM_HD void jvals_part(int ntop, int nmax, cd z, double rs, cd* out, int stride) {
  if (cabs_(z) < 1e-200) {
    out[0] = mk(1, 0);
    for (int n = 1; n <= nmax; ++n) out[n * stride] = mk(0, 0);
    return;
  }
  int top = miller_top(ntop, z);
  cd zr = inv(z) * rs;       // g_{n-1} = 2n (rs/z) g_n - rs^2 g_{n+1}
  double r2 = rs * rs;
  cd u = z.y >= 0 ? mk(0, -1) : mk(0, 1);
  cd ur = u * rs;
  cd gp = mk(0, 0), g = mk(1, 0), H = mk(0, 0);  // g = g_{top-1}, gp = g_top
  // Horner: H <- u rs (g_n + H), n = top-1 .. 1
  for (int n = top - 1; n >= 1; --n) {
    if (n <= nmax) out[n * stride] = g;
    H = ur * (g + H);
    cd gm = zr * g * (2.0 * n) - gp * r2;
    gp = g;
    g = gm;
    if (norm2(g) > 1e40) {
      g = g * 1e-40; gp = gp * 1e-40; H = H * 1e-40;
      for (int k = n; k <= nmax && k <= top - 1; ++k) out[k * stride] = out[k * stride] * 1e-40;
    }
  }
  for (int n = top; n <= nmax; ++n) out[n * stride] = mk(0, 0);
  cd lam = (g + H * 2.0) * cexp_(-(u * z));   // e^{-uz} (g_0 + 2 sum u^n rs^n g_n) = const
  cd sc = inv(lam);
  out[0] = g * sc;
  for (int n = 1; n <= nmax && n <= top - 1; ++n) out[n * stride] = out[n * stride] * sc;
}

// Scaled J_n(z)/rs^n for n = 0..nmax into out[n*stride].
FMM_HD void jvals(int nmax, cd z, double rs, cd* out, int stride) {
  jvals_part(nmax, nmax, z, rs, out, stride);
}

// J_n(x)/rs^n, n = 0..nmax, of a real argument x > 0: Miller's downward
// recurrence as in jvals, in real arithmetic, normalised with the Neumann
// sum J_0 + 2 sum_k J_2k = 1 (Abramowitz & Stegun 9.1.46).
FMM_HD void jvals_real(int nmax, double x, double rs, double* out, int stride) {
  if (x < 1e-200) {
    out[0] = 1.0;
    for (int n = 1; n <= nmax; ++n) out[n * stride] = 0.0;
    return;
  }
  const double xi = 1.0 / x;
  int top = nmax + 10000;
  {
    double fn = 1.0, fm = 0.0;
    for (int i = nmax; i <= nmax + 10000; ++i) {
      const double fp = xi * fn * (2.0 * i) - fm;
      if (fp * fp > 1e40) {
        top = i + 1;
        break;
      }
      fm = fn;
      fn = fp;
    }
  }
  const double zr = xi * rs, r2 = rs * rs;
  double gp = 0.0, g = 1.0, H = 0.0;  // H = sum over even n >= 2 of rs^n g_n (Horner)
  for (int n = top - 1; n >= 1; --n) {
    if (n <= nmax) out[n * stride] = g;
    H = rs * ((n & 1) ? H : g + H);
    const double gm = zr * g * (2.0 * n) - gp * r2;
    gp = g;
    g = gm;
    if (g * g > 1e40) {
      g *= 1e-40; gp *= 1e-40; H *= 1e-40;
      for (int k = n; k <= nmax && k <= top - 1; ++k) out[k * stride] *= 1e-40;
    }
  }
  for (int n = top; n <= nmax; ++n) out[n * stride] = 0.0;
  const double sc = 1.0 / (g + 2.0 * H);
  out[0] = g * sc;
  for (int n = 1; n <= nmax && n <= top - 1; ++n) out[n * stride] *= sc;
}

// Scaled H_n(x) rs^n of a real argument 0 < x <= JY_XMAX for N: H0, H1 from
// hank_series_real<N>, then the upward recurrence with real multipliers.
template <int N>
FMM_HD void hvals_real(int nmax, double x, double rs, cd* out, int stride) {
  if (x < 1e-200) {
    for (int n = 0; n <= nmax; ++n) out[n * stride] = mk(0, 0);
    return;
  }
  cd h0, h1;
  hank_series_real<N>(x, h0, h1);
  out[0] = h0;
  if (nmax >= 1) out[stride] = h1 * rs;
  const double zi = rs / x, s2 = rs * rs;
  for (int i = 1; i <= nmax - 1; ++i)
    out[(i + 1) * stride] = out[i * stride] * (zi * (2.0 * i)) - out[(i - 1) * stride] * s2;
}

// Scaled H_n(z) rs^n, n = 0..nmax (h2dcommon.c:35-90)
FMM_HD void hvals(int nmax, cd z, double rs, cd* out, int stride) {
  if (cabs_(z) < 1e-200) {
    for (int n = 0; n <= nmax; ++n) out[n * stride] = mk(0, 0);
    return;
  }
  cd h0, h1;
  hank01(z, h0, h1);
  out[0] = h0;
  if (nmax >= 1) out[stride] = h1 * rs;
  cd zi = inv(z) * rs;
  double s2 = rs * rs;
  for (int i = 1; i <= nmax - 1; ++i)
    out[(i + 1) * stride] = zi * out[i * stride] * (2.0 * i) - out[(i - 1) * stride] * s2;
}

}  // namespace fmm
