// direct.hip -- O(N*M) direct sums on the GPU behind the reference's
// c2d_direct* entry points (c_translation/src/cauchykernels2d.c:31-442).
// One thread per target; sources staged through LDS in tiles of 256.
// Outputs are incremented, as in the reference.
#include "../../include/fmm2d_hip.h"
#include "engine.h"

namespace fmm {
namespace {

constexpr int TILE = 256;

// skip test per reference routine: 0 rr<=th^2, 1 rr<th^2 (charges),
// 2 |z|<=th, 3 |z|<th (dipole-only routines compare cabs(z))
template <bool C, bool Q, int PG>
__global__ void __launch_bounds__(TILE) k_c2d(const double2* src, int ns, const double2* chg,
                                              const double2* dip, const double2* trg, int nt,
                                              int nd, int mode, double th, double2* pot,
                                              double2* grad, double2* hess) {
  __shared__ double2 sx[TILE];
  __shared__ double2 sc[TILE];
  __shared__ double2 sq[TILE];
  const int j = blockIdx.x * TILE + threadIdx.x;
  const bool act = j < nt;
  const double2 x = act ? trg[j] : make_double2(0, 0);
  for (int d = 0; d < nd; ++d) {
    cd P = mk(0, 0), G = mk(0, 0), H = mk(0, 0);
    for (int s0 = 0; s0 < ns; s0 += TILE) {
      int cnt = min(TILE, ns - s0);
      if ((int)threadIdx.x < cnt) {
        sx[threadIdx.x] = src[s0 + threadIdx.x];
        if (C) sc[threadIdx.x] = chg[(size_t)(s0 + threadIdx.x) * nd + d];
        if (Q) sq[threadIdx.x] = dip[(size_t)(s0 + threadIdx.x) * nd + d];
      }
      __syncthreads();
      if (act)
        for (int m = 0; m < cnt; ++m) {
          double dx = x.x - sx[m].x, dy = x.y - sx[m].y;
          double rr = dx * dx + dy * dy;
          bool skip;
          if (mode == 0) skip = rr <= th * th;
          else if (mode == 1) skip = rr < th * th;
          else if (mode == 2) skip = hypot(dx, dy) <= th;
          else skip = hypot(dx, dy) < th;
          if (skip) continue;
          double ri = 1.0 / rr;
          cd zi = mk(dx * ri, -dy * ri), z2 = -(zi * zi);
          if (C) {
            cd c = mk(sc[m].x, sc[m].y);
            P += c * (0.5 * log(rr));
            if (PG >= 2) G += c * zi;
            if (PG >= 3) H += c * z2;
          }
          if (Q) {
            cd q = mk(sq[m].x, sq[m].y);
            P += q * zi;
            if (PG >= 2) G += q * z2;
            if (PG >= 3) H += q * ((z2 * zi) * -2.0);
          }
        }
      __syncthreads();
    }
    if (act) {
      size_t o = (size_t)j * nd + d;
      pot[o] = make_double2(pot[o].x + P.x, pot[o].y + P.y);
      if (PG >= 2) grad[o] = make_double2(grad[o].x + G.x, grad[o].y + G.y);
      if (PG >= 3) hess[o] = make_double2(hess[o].x + H.x, hess[o].y + H.y);
    }
  }
}

template <bool C, bool Q, int PG>
int c2d(int nd, const double* sources, int ns, const void* charge, const void* dipstr,
        const double* targ, int nt, void* pot, void* grad, void* hess, double th, int mode) {
  if (nd < 1 || ns < 0 || nt < 0) return IER_ARGS;
  if (ns == 0 || nt == 0) return IER_OK;
  Ctx c;
  int e = get_ctx(c);
  if (e) return e;
  Scratch S(c.st);
  double2* s = S.get<double2>(ns);
  double2* t = S.get<double2>(nt);
  double2* cc = C ? S.get<double2>((size_t)nd * ns) : nullptr;
  double2* qq = Q ? S.get<double2>((size_t)nd * ns) : nullptr;
  size_t no = (size_t)nd * nt;
  double2* po = S.get<double2>(no);
  double2* go = PG >= 2 ? S.get<double2>(no) : nullptr;
  double2* ho = PG >= 3 ? S.get<double2>(no) : nullptr;
  if (S.err) return S.err;
  HIP_TRY(hipMemcpyAsync(s, sources, 16 * (size_t)ns, hipMemcpyHostToDevice, c.st));
  HIP_TRY(hipMemcpyAsync(t, targ, 16 * (size_t)nt, hipMemcpyHostToDevice, c.st));
  if (C) HIP_TRY(hipMemcpyAsync(cc, charge, 16 * (size_t)nd * ns, hipMemcpyHostToDevice, c.st));
  if (Q) HIP_TRY(hipMemcpyAsync(qq, dipstr, 16 * (size_t)nd * ns, hipMemcpyHostToDevice, c.st));
  HIP_TRY(hipMemcpyAsync(po, pot, 16 * no, hipMemcpyHostToDevice, c.st));
  if (go) HIP_TRY(hipMemcpyAsync(go, grad, 16 * no, hipMemcpyHostToDevice, c.st));
  if (ho) HIP_TRY(hipMemcpyAsync(ho, hess, 16 * no, hipMemcpyHostToDevice, c.st));
  k_c2d<C, Q, PG><<<(nt + TILE - 1) / TILE, TILE, 0, c.st>>>(s, ns, cc, qq, t, nt, nd, mode, th,
                                                             po, go, ho);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(pot, po, 16 * no, hipMemcpyDeviceToHost, c.st));
  if (go) HIP_TRY(hipMemcpyAsync(grad, go, 16 * no, hipMemcpyDeviceToHost, c.st));
  if (ho) HIP_TRY(hipMemcpyAsync(hess, ho, 16 * no, hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipStreamSynchronize(c.st));
  return IER_OK;
}

}  // namespace
}  // namespace fmm

extern "C" {
void c2d_directcp_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c,
                   const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, const double* th) {
  fmm::c2d<true, false, 1>(*nd, s, *ns, c, nullptr, t, *nt, pot, nullptr, nullptr, *th, 0);
}
void c2d_directcg_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c,
                   const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad,
                   const double* th) {
  fmm::c2d<true, false, 2>(*nd, s, *ns, c, nullptr, t, *nt, pot, grad, nullptr, *th, 1);
}
void c2d_directch_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c,
                   const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad,
                   fmm2d_cplx* hess, const double* th) {
  fmm::c2d<true, false, 3>(*nd, s, *ns, c, nullptr, t, *nt, pot, grad, hess, *th, 1);
}
void c2d_directdp_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* q,
                   const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, const double* th) {
  fmm::c2d<false, true, 1>(*nd, s, *ns, nullptr, q, t, *nt, pot, nullptr, nullptr, *th, 2);
}
void c2d_directdg_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* q,
                   const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad,
                   const double* th) {
  fmm::c2d<false, true, 2>(*nd, s, *ns, nullptr, q, t, *nt, pot, grad, nullptr, *th, 3);
}
void c2d_directdh_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* q,
                   const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad,
                   fmm2d_cplx* hess, const double* th) {
  fmm::c2d<false, true, 3>(*nd, s, *ns, nullptr, q, t, *nt, pot, grad, hess, *th, 3);
}
void c2d_directcdp_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns,
                    const fmm2d_cplx* c, const fmm2d_cplx* q, const double* t,
                    const fmm2d_int* nt, fmm2d_cplx* pot, const double* th) {
  fmm::c2d<true, true, 1>(*nd, s, *ns, c, q, t, *nt, pot, nullptr, nullptr, *th, 0);
}
void c2d_directcdg_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns,
                    const fmm2d_cplx* c, const fmm2d_cplx* q, const double* t,
                    const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, const double* th) {
  fmm::c2d<true, true, 2>(*nd, s, *ns, c, q, t, *nt, pot, grad, nullptr, *th, 1);
}
void c2d_directcdh_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns,
                    const fmm2d_cplx* c, const fmm2d_cplx* q, const double* t,
                    const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess,
                    const double* th) {
  fmm::c2d<true, true, 3>(*nd, s, *ns, c, q, t, *nt, pot, grad, hess, *th, 1);
}
}
