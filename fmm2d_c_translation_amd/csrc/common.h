// common.h -- shared types and helpers for the MI355X 2D FMM engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstddef>

#include "log_table.h"

#define FMM_HD __host__ __device__ inline

namespace fmm {

// Error codes written to `ier`.  0/4/13 keep their reference meaning
// (success / workspace allocation failure / expansion order over 10000,
// c_translation/src/hfmm2d.c:1810-1815, h2dterms.c:62-64); codes >= 100
// are this engine's own and are documented in include/fmm2d_hip.h.
enum : int {
  IER_OK = 0,
  IER_ALLOC = 4,
  IER_NTERMS = 13,
  IER_HIP = 101,        // a HIP runtime call failed
  IER_ARGS = 102,       // argument outside what the engine supports
  IER_DEPTH = 103,      // tree would need more than kMaxDepth levels
  IER_NOGPU = 104,      // no usable gfx950 device
  IER_HIGHFREQ = 105,   // Helmholtz box size above 16 wavelengths (diagonal form)
  IER_P_TOO_LARGE = 106, // expansion order beyond the kernels' register/LDS budget
  IER_SCRATCH = 107      // beyond the allocation-free entry point's fixed scratch
};

// Keys hold 2 bits per level below the root: kMaxDepth levels in 58 bits.
constexpr int kMaxDepth = 29;

struct cd {  // complex double, 16-byte aligned, layout of double _Complex
  double x, y;
};

FMM_HD cd mk(double a, double b) { return cd{a, b}; }
FMM_HD cd operator+(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
FMM_HD cd operator-(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
FMM_HD cd operator-(cd a) { return cd{-a.x, -a.y}; }
FMM_HD cd operator*(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
FMM_HD cd operator*(cd a, double s) { return cd{a.x * s, a.y * s}; }
FMM_HD cd operator*(double s, cd a) { return cd{a.x * s, a.y * s}; }
FMM_HD cd operator/(cd a, double s) { return cd{a.x / s, a.y / s}; }
FMM_HD cd& operator+=(cd& a, cd b) { a.x += b.x; a.y += b.y; return a; }
FMM_HD cd& operator-=(cd& a, cd b) { a.x -= b.x; a.y -= b.y; return a; }
FMM_HD cd conj(cd a) { return cd{a.x, -a.y}; }
FMM_HD double norm2(cd a) { return a.x * a.x + a.y * a.y; }
// 1/z by Smith's scaling: no overflow for |z| near the double range ends.
FMM_HD cd inv(cd z) {
  if (fabs(z.x) >= fabs(z.y)) {
    double r = z.y / z.x, d = z.x + z.y * r;
    return cd{1.0 / d, -r / d};
  }
  double r = z.x / z.y, d = z.y + z.x * r;
  return cd{r / d, -1.0 / d};
}
FMM_HD cd cdiv(cd a, cd b) { return a * inv(b); }
// z^n, n >= 0, by repeated squaring
FMM_HD cd cpow_n(cd z, int n) {
  cd r = cd{1.0, 0.0};
  for (; n > 0; n >>= 1) {
    if (n & 1) r = r * z;
    z = z * z;
  }
  return r;
}

// Natural log of a positive finite double.  x = 2^k m with m in
// [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m - 1)/(m + 1), |s| < 0.172,
// from the series 2 (s + s^3/3 + ... + s^23/23) (the first omitted term is
// below 1e-18 of the sum); k ln 2 in two parts (Cody and Waite), the high
// part with 21 significant bits so that k ln2_hi is exact.  About thirty
// instructions and a couple of ulp, without branches: zero gives -inf,
// +inf gives +inf, negative and NaN arguments give NaN.  log_pos_core is
// the same for a positive normal finite x, without the special cases.
FMM_HD double log_pos_core(double x) {
  int k;
  double m = frexp(x, &k);  // [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    --k;
  }
  const double a = m - 1.0, b = m + 1.0;
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(b);  // refined by two Newton steps
  r = fma(r, fma(-b, r, 1.0), r);
  r = fma(r, fma(-b, r, 1.0), r);
#else
  const double r = 1.0 / b;
#endif
  const double s = a * r, s2 = s * s;
  double p = 1.0 / 23.0;
  p = fma(p, s2, 1.0 / 21.0);
  p = fma(p, s2, 1.0 / 19.0);
  p = fma(p, s2, 1.0 / 17.0);
  p = fma(p, s2, 1.0 / 15.0);
  p = fma(p, s2, 1.0 / 13.0);
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  const double lm = fma(2.0 * s * s2, p, 2.0 * s);
  const double ln2_hi = 0x1.62e42p-1, ln2_lo = 0x1.fdf473de6af28p-22;
  const double kd = (double)k;
  return fma(kd, ln2_hi, fma(kd, ln2_lo, lm));
}
FMM_HD double log_pos(double x) {
  const double v = log_pos_core(x);
  const bool normal = x > 0.0 && x <= 1.7976931348623157e308;
  return normal ? v : x == 0.0 ? -HUGE_VAL : x > 0.0 ? x : NAN;
}

// 1/sqrt(x) for a positive normal finite x: the hardware estimate refined
// by one third-order step, y + y e (1/2 + 3/8 e) with e = 1 - x y^2
// (Householder's method for x y^2 = 1), without the special-value checks of
// the library rsqrt.
FMM_HD double rsqrt_pos(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
#else
  return 1.0 / std::sqrt(x);
#endif
}

// Natural log of a positive normal finite double by table: x = 2^e m with
// m in [1, 2) rounded to the nearest c_j = 1 + j/256, whose table row holds
// r_j ~ 1/c_j and L_j = -ln r_j exactly rounded (log_table.h; from c_j >=
// sqrt 2 the row holds L_j - ln 2 and e counts one more), so ln x =
// e ln 2 + L_j + log1p(u) with u = m r_j - 1 (one FMA, |u| <= 2^-9) and
// log1p(u) = u - u^2/2 + ... + u^7/7 (the first omitted term is below
// 1e-23); e ln 2 in two parts as in log_pos_core.  Next to x = 1 the table
// term is exactly 0.  About 18 instructions and one table load, against
// about 30 for log_pos_core; no special cases.
struct LogRow {
  double r, l;
};
static constexpr LogRow kLogTab[LOG_TAB_N] = {LOG_TAB_ROWS};
FMM_HD double log_tab(double x) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  const uint64_t frac = b & 0x000fffffffffffffull;
  const int j = (int)((frac + (uint64_t(1) << 43)) >> 44);  // 0 .. 256
  const int e = (int)((b >> 52) & 0x7ff) - 1023 + (j >= LOG_TAB_SHIFT);
  const double m = __builtin_bit_cast(double, frac | 0x3ff0000000000000ull);
  const LogRow t = kLogTab[j];
  const double u = fma(m, t.r, -1.0);
  double p = 1.0 / 7.0;
  p = fma(p, u, -1.0 / 6.0);
  p = fma(p, u, 1.0 / 5.0);
  p = fma(p, u, -1.0 / 4.0);
  p = fma(p, u, 1.0 / 3.0);
  p = fma(p, u, -0.5);
  const double l1p = fma(p * u, u, u);
  const double ln2_hi = 0x1.62e42p-1, ln2_lo = 0x1.fdf473de6af28p-22;
  const double ed = (double)e;
  return fma(ed, ln2_hi, fma(ed, ln2_lo, t.l + l1p));
}

// Quadrant digit of a point relative to a box centre: bit 0 set for the
// right half (x - cx >= 0), bit 1 for the upper half (y - cy >= 0).  Child
// j = digit + 1 in the reference numbering (-x,-y), (+x,-y), (-x,+y), (+x,+y)
// (c_translation/src/tree_routs2d.c:77-92, pts_tree2d.c:700-745).
FMM_HD int quadrant(double x, double y, double cx, double cy) {
  return (x - cx < 0 ? 0 : 1) | (y - cy < 0 ? 0 : 2);
}

// Child centre from the parent centre and the child box size, the same
// floating-point operations as the reference's tree refinement
// (tree_routs2d.c:81-84): c + (+-1) * bs / 2.
FMM_HD double child_coord(double c, int positive, double bs_child) {
  return c + (positive ? 1.0 : -1.0) * bs_child / 2;
}

// Morton path key of a point: digit of level l (1-based) in bits
// 2*(kMaxDepth-l) .. +1, obtained by walking down from the root centre with
// the reference's centre recurrence, so the digit at every level is exactly
// the child the reference's stable partition would put the point in.
FMM_HD uint64_t point_key(double x, double y, double cx, double cy, double bs0) {
  uint64_t key = 0;
  double bs = bs0;
  for (int l = 1; l <= kMaxDepth; ++l) {
    int q = quadrant(x, y, cx, cy);
    bs = bs / 2;
    cx = child_coord(cx, q & 1, bs);
    cy = child_coord(cy, q & 2, bs);
    key |= (uint64_t)q << (2 * (kMaxDepth - l));
  }
  return key;
}

FMM_HD uint64_t box_key_lo(uint64_t prefix, int level) {
  return prefix << (2 * (kMaxDepth - level));
}

}  // namespace fmm

#define HIP_TRY(expr)                         \
  do {                                        \
    hipError_t _e = (expr);                   \
    if (_e != hipSuccess) return fmm::IER_HIP; \
  } while (0)
