// wideband.hip -- the Helmholtz FMM's diagonal (high-frequency) translations
// for boxes of more than 16 wavelengths (c_translation/src/wideband2d.c:33-290,
// used by hfmm2d.c:946-960, 993-1051, 1057-1160, 1165-1254).
//
// An expansion a_n, |n| <= p, is turned into its signature, the length-nsig
// DFT  sig_k = sum_n a_n exp(-2 pi i n k / nsig)  (h2d_mptosig,
// wideband2d.c:52-80); a translation is then a pointwise product with the
// DFT of the translation kernel B(m), |m| <= P, divided by nsig
// (h2d_mkmpshift / h2d_mkm2ltrans, :120-226), and the inverse DFT truncated
// to the output order gives the translated expansion (h2d_sig2exp, :85-115):
//   out_j = sum_n a_n B(j - n),
// exactly the dense translation of the low-frequency path, because
// nsig >= 2 P + 1 keeps the circular convolution free of aliasing.  The
// kernel is B(m) = v_|m| e^{-i m theta} (-1)^{m, m < 0}, with theta the angle
// of c2 - c1 minus pi, v = J_m(k r) for the shifts (M2M, L2L) and
// H_m(k r) for M2L, at scale 1 (the scale is 1 on every such level).
//
// GPU layout: one row of nsig complex values per (box, channel); every
// transform is a row of a batched FFT (k_fft_rows, a Stockham mixed-radix
// 4/2/3/5 FFT of the library's own, nsig = next235(...) has no other
// factors); the diagonal step is one block per (output box, channel)
// accumulating over the 49 list-2 offsets or the 4 child quadrants.
#include <algorithm>
#include <vector>

#include "engine.h"
#include "special.h"

namespace fmm {
namespace {

constexpr int kFftThreads = 256;
constexpr int kMaxFactors = 48;

struct FftPlan {
  int n, nf;
  int f[kMaxFactors];  // radices in pass order: 4s, then 2, 3s, 5s
};

bool fft_plan(int n, FftPlan& p) {
  p.n = n;
  p.nf = 0;
  int m = n;
  while (m % 4 == 0 && p.nf < kMaxFactors) { p.f[p.nf++] = 4; m /= 4; }
  while (m % 2 == 0 && p.nf < kMaxFactors) { p.f[p.nf++] = 2; m /= 2; }
  while (m % 3 == 0 && p.nf < kMaxFactors) { p.f[p.nf++] = 3; m /= 3; }
  while (m % 5 == 0 && p.nf < kMaxFactors) { p.f[p.nf++] = 5; m /= 5; }
  return m == 1;
}

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// exp(sign * 2 pi i q / d), 0 <= q < d, with the argument reduced exactly
__device__ inline double2 root(int q, int d, int sign) {
  double s, c;
  sincospi(2.0 * (double)q / (double)d, &s, &c);
  return make_double2(c, sign * s);
}

// Unnormalised DFT of every row (rows of n = nsig values, one block per
// row): sign -1 is FFTPACK's zfftf, +1 its zfftb.  Stockham autosort: the
// pass of radix R over sub-transforms of length Ns combines R of them,
//   y[(j / Ns) Ns R + j % Ns + r Ns] = sum_q w^(q r) x[j + q n / R] t^q,
// t = exp(sign 2 pi i (j % Ns) / (Ns R)) the twiddle, w = exp(sign 2 pi i / R);
// passes alternate between the row and its scratch copy.
__global__ void __launch_bounds__(kFftThreads) k_fft_rows(double2* rows, double2* tmp, FftPlan P,
                                                          int sign) {
  const int n = P.n;
  double2* a = rows + (size_t)blockIdx.x * n;
  double2* b = tmp + (size_t)blockIdx.x * n;
  int Ns = 1;
  for (int f = 0; f < P.nf; ++f) {
    const int R = P.f[f], m = n / R;
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
      const int k = j % Ns;
      double2 v[5];
      for (int q = 0; q < R; ++q) {
        v[q] = a[j + q * m];
        if (q && k) v[q] = cmul(v[q], root(k * q, Ns * R, sign));
      }
      double2 o[5];
      if (R == 2) {
        o[0] = cadd(v[0], v[1]);
        o[1] = csub(v[0], v[1]);
      } else if (R == 4) {
        const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const double2 t2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        const double2 t3 = make_double2(-sign * d.y, sign * d.x);  // (sign i) d
        o[0] = cadd(t0, t2);
        o[1] = cadd(t1, t3);
        o[2] = csub(t0, t2);
        o[3] = csub(t1, t3);
      } else {
        for (int r = 0; r < R; ++r) {
          double2 s = v[0];
          for (int q = 1; q < R; ++q) s = cadd(s, cmul(v[q], root((q * r) % R, R, sign)));
          o[r] = s;
        }
      }
      const int d0 = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; ++r) b[d0 + r * Ns] = o[r];
    }
    __syncthreads();
    double2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  if (P.nf & 1) {  // the result is in the scratch row
    double2* dst = rows + (size_t)blockIdx.x * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) dst[j] = a[j];
  }
}

hipError_t fft_rows(hipStream_t st, Scratch& S, double2* rows, int nrows, int n, int sign) {
  if (nrows <= 0) return hipSuccess;
  FftPlan P;
  if (!fft_plan(n, P)) return hipErrorInvalidValue;
  double2* tmp = S.get<double2>((size_t)nrows * n);
  if (!tmp) return hipErrorOutOfMemory;
  k_fft_rows<<<nrows, kFftThreads, 0, st>>>(rows, tmp, P, sign);
  return hipGetLastError();
}

// c2 - c1 of slot s of a translation kind (0 M2M, 1 M2L, 2 L2L) at box size
// bs (the parent's for M2M and L2L); false for the near M2L offsets, which
// no list-2 pair uses
__device__ inline bool slot_offset(int kind, int s, double bs, double& ox, double& oy) {
  if (kind == 1) {
    const int dx = s % 7 - 3, dy = s / 7 - 3;  // source at (dx, dy) bs from the target
    if (abs(dx) <= 1 && abs(dy) <= 1) return false;
    ox = -dx * bs;
    oy = -dy * bs;
    return true;
  }
  const double h = bs / 4;  // child centre = parent centre +- h
  const double sx = (s & 1) ? h : -h, sy = (s & 2) ? h : -h;
  ox = kind == 0 ? -sx : sx;
  oy = kind == 0 ? -sy : sy;
  return true;
}

// Translation vectors, one row of nsig per slot: B(m) / nsig at index
// m mod nsig, |m| <= P (h2d_mkmpshift, h2d_mkm2ltrans before their FFT).
// Thread 0 evaluates the Bessel or Hankel values into vbuf (P + 4 per slot).
__global__ void __launch_bounds__(256) k_hf_kernel(int kind, int P, int nsig, double bs, cd zk,
                                                   cd* vbuf, double2* T) {
  const int s = blockIdx.x;
  double2* row = T + (size_t)s * nsig;
  double ox = 0, oy = 0;
  if (!slot_offset(kind, s, bs, ox, oy)) {
    for (int i = threadIdx.x; i < nsig; i += blockDim.x) row[i] = make_double2(0, 0);
    return;
  }
  cd* v = vbuf + (size_t)s * (P + 4);
  const double r = sqrt(ox * ox + oy * oy), th = atan2(oy, ox) - kPi;
  if (threadIdx.x == 0) {
    if (kind == 1) hvals(P + 1, zk * r, 1.0, v, 1);
    else jvals(P + 3, zk * r, 1.0, v, 1);
  }
  __syncthreads();
  const double inv = 1.0 / nsig;
  for (int i = threadIdx.x; i < nsig; i += blockDim.x) {
    int m;
    if (i <= P) m = i;
    else if (i >= nsig - P) m = i - nsig;
    else {
      row[i] = make_double2(0, 0);
      continue;
    }
    const int am = abs(m);
    double sn, cs;
    sincos(-m * th, &sn, &cs);
    double2 b = cmul(make_double2(v[am].x, v[am].y), make_double2(cs, sn));
    const double sg = (m < 0 && (am & 1)) ? -inv : inv;
    row[i] = make_double2(b.x * sg, b.y * sg);
  }
}

// rows[(r nd + d) nsig + k]: expansion of box first + r, channel d, laid out
// for its signature (order n at index n mod nsig, zeros elsewhere)
__global__ void k_hf_rows(const double2* __restrict__ X, size_t xs, int first, int p, int nd,
                          int nsig, double2* __restrict__ rows) {
  const int rd = blockIdx.x, r = rd / nd, d = rd - r * nd;
  const double2* x = X + (size_t)(first + r) * xs + (size_t)d * (2 * p + 1) + p;
  double2* row = rows + (size_t)rd * nsig;
  for (int i = threadIdx.x; i < nsig; i += blockDim.x) {
    double2 v = make_double2(0, 0);
    if (i <= p) v = x[i];
    else if (i >= nsig - p) v = x[i - nsig];
    row[i] = v;
  }
}

// out[(w nd + d)] = sum over the slots s of work item w with an input box j:
// T[s] .* in[(j - first) nd + d]  (h2d_diagtrans, wideband2d.c:33-47)
__global__ void k_hf_diag(const int* __restrict__ map, int nslot, const double2* __restrict__ T,
                          const double2* __restrict__ in, int first, int nd, int nsig,
                          double2* __restrict__ out) {
  const int wd = blockIdx.x, w = wd / nd, d = wd - w * nd;
  const int* mw = map + (size_t)w * nslot;
  double2* o = out + (size_t)wd * nsig;
  for (int k = threadIdx.x; k < nsig; k += blockDim.x) {
    double2 acc = make_double2(0, 0);
    for (int s = 0; s < nslot; ++s) {
      const int j = mw[s];
      if (j < 0) continue;
      const double2 t = T[(size_t)s * nsig + k];
      const double2 x = in[((size_t)(j - first) * nd + d) * nsig + k];
      acc.x += t.x * x.x - t.y * x.y;
      acc.y += t.x * x.y + t.y * x.x;
    }
    o[k] = acc;
  }
}

// Y[work[w]] += the inverse transform of row (w, d) truncated to order p
// (h2d_sig2exp, wideband2d.c:85-115)
__global__ void k_hf_add(const int* __restrict__ work, const double2* __restrict__ rows, int p,
                         int nd, int nsig, double2* __restrict__ Y, size_t ys) {
  const int wd = blockIdx.x, w = wd / nd, d = wd - w * nd;
  const double2* row = rows + (size_t)wd * nsig;
  double2* y = Y + (size_t)work[w] * ys + (size_t)d * (2 * p + 1) + p;
  for (int n = (int)threadIdx.x - p; n <= p; n += blockDim.x) {
    const double2 v = row[n >= 0 ? n : n + nsig];
    y[n] = make_double2(y[n].x + v.x, y[n].y + v.y);
  }
}

}  // namespace

int next235(double base) {
  // next235.c:10-32: the smallest even product of powers of 2, 3 and 5 at
  // least base
  int r = 2 * (int)(base / 2.0 + 0.9999);
  if (r <= 0) r = 2;
  for (;; r += 2) {
    int m = r;
    while (m % 2 == 0) m /= 2;
    while (m % 3 == 0) m /= 3;
    while (m % 5 == 0) m /= 5;
    if (m == 1) return r;
  }
}

int hf_translate(hipStream_t st, Scratch& S, std::complex<double> zk, const HfJob& j,
                 const double2* X, size_t xs, double2* Y, size_t ys, int nd) {
  if (j.nwork <= 0) return IER_OK;
  const int P = j.pin + j.pout, nsig = j.nsig;
  if (nsig < 2 * P + 1 || j.nslot < 1) return IER_ARGS;
  const int nslot = j.nslot;
  cd* vbuf = S.get<cd>((size_t)nslot * (P + 4));
  double2* T = S.get<double2>((size_t)nslot * nsig);
  double2* in = S.get<double2>((size_t)j.in_count * nd * nsig);
  double2* out = S.get<double2>((size_t)j.nwork * nd * nsig);
  if (S.err) return S.err;
  const cd k = mk(zk.real(), zk.imag());
  k_hf_kernel<<<nslot, 256, 0, st>>>(j.kind, P, nsig, j.bs, k, vbuf, T);
  HIP_TRY(fft_rows(st, S, T, nslot, nsig, -1));
  k_hf_rows<<<j.in_count * nd, 256, 0, st>>>(X, xs, j.in_first, j.pin, nd, nsig, in);
  HIP_TRY(fft_rows(st, S, in, j.in_count * nd, nsig, -1));
  k_hf_diag<<<j.nwork * nd, 256, 0, st>>>(j.map, nslot, T, in, j.in_first, nd, nsig, out);
  HIP_TRY(fft_rows(st, S, out, j.nwork * nd, nsig, +1));
  k_hf_add<<<j.nwork * nd, 256, 0, st>>>(j.work, out, j.pout, nd, nsig, Y, ys);
  HIP_TRY(hipGetLastError());
  return S.err;
}

int fft_rows_device(hipStream_t st, Scratch& S, double2* rows, int nrows, int n, int sign) {
  return fft_rows(st, S, rows, nrows, n, sign) == hipSuccess ? IER_OK : IER_ARGS;
}

}  // namespace fmm
