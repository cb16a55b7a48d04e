// direct.hip -- O(N*M) direct sums on the GPU behind the reference's
// c2d_direct* (c_translation/src/cauchykernels2d.c:31-442), l2d_direct*
// (lapkernels2d.c:19-494), r2d_direct* (rlapkernels2d.c:16-491) and
// h2d_direct* (helmkernels2d.c:21-520) entry points.  One thread per target,
// sources staged through LDS in tiles of 256; when there are few targets
// the sources are also split into chunks across blocks (launch_chunked).  Outputs are incremented, as in the reference; each routine keeps
// its own skip test (rr vs thresh^2 or |z| vs thresh, <= or <).
#include "../../include/fmm2d_hip.h"
#include "engine.h"
#include "special.h"

namespace fmm {
namespace {

constexpr int TILE = 256;

// skip test per reference routine: 0 rr<=th^2, 1 rr<th^2 (charges),
// 2 |z|<=th, 3 |z|<th (dipole-only routines compare cabs(z))
template <bool C, bool Q, int PG>
__global__ void __launch_bounds__(TILE) k_c2d(const double2* src, int ns, const double2* chg,
                                              const double2* dip, const double2* trg, int nt,
                                              int nd, int mode, double th, double2* pot,
                                              double2* grad, double2* hess, int len, size_t sp, size_t sg, size_t sh) {
  // sources of this block's chunk; outputs of chunk y go to slice y
  const int sb = blockIdx.y * len, se = min(ns, sb + len);
  pot += blockIdx.y * sp;
  if (PG >= 2) grad += blockIdx.y * sg;
  if (PG >= 3) hess += blockIdx.y * sh;
  __shared__ double2 sx[TILE];
  __shared__ double2 sc[TILE];
  __shared__ double2 sq[TILE];
  const int j = blockIdx.x * TILE + threadIdx.x;
  const bool act = j < nt;
  const double2 x = act ? trg[j] : make_double2(0, 0);
  for (int d = 0; d < nd; ++d) {
    cd P = mk(0, 0), G = mk(0, 0), H = mk(0, 0);
    for (int s0 = sb; s0 < se; s0 += TILE) {
      int cnt = min(TILE, se - s0);
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

// log kernel with complex densities (lapkernels2d.c): charge c log r,
// dipole d (v . grad_y log|x - y|); gradient (nd,2,nt), Hessian (nd,3,nt)
template <bool C, bool Q, int PG>
__global__ void __launch_bounds__(TILE) k_l2d(const double2* src, int ns, const double2* chg,
                                              const double2* dip, const double* dv,
                                              const double2* trg, int nt, int nd, bool le,
                                              double th2, double2* pot, double2* grad,
                                              double2* hess, int len, size_t sp, size_t sg, size_t sh) {
  // sources of this block's chunk; outputs of chunk y go to slice y
  const int sb = blockIdx.y * len, se = min(ns, sb + len);
  pot += blockIdx.y * sp;
  if (PG >= 2) grad += blockIdx.y * sg;
  if (PG >= 3) hess += blockIdx.y * sh;
  __shared__ double2 sx[TILE];
  __shared__ double2 sc[TILE];
  __shared__ double2 sq[TILE];
  __shared__ double2 sv[TILE];
  const int j = blockIdx.x * TILE + threadIdx.x;
  const bool act = j < nt;
  const double2 x = act ? trg[j] : make_double2(0, 0);
  for (int d = 0; d < nd; ++d) {
    cd P = mk(0, 0), Gx = mk(0, 0), Gy = mk(0, 0), Hxx = mk(0, 0), Hxy = mk(0, 0), Hyy = mk(0, 0);
    for (int s0 = sb; s0 < se; s0 += TILE) {
      int cnt = min(TILE, se - s0);
      if ((int)threadIdx.x < cnt) {
        const size_t i = s0 + threadIdx.x;
        sx[threadIdx.x] = src[i];
        if (C) sc[threadIdx.x] = chg[i * nd + d];
        if (Q) {
          sq[threadIdx.x] = dip[i * nd + d];
          sv[threadIdx.x] = make_double2(dv[(i * 2 + 0) * nd + d], dv[(i * 2 + 1) * nd + d]);
        }
      }
      __syncthreads();
      if (act)
        for (int m = 0; m < cnt; ++m) {
          const double dx = x.x - sx[m].x, dy = x.y - sx[m].y;
          const double dx2 = dx * dx, dy2 = dy * dy, rr = dx2 + dy2;
          if (le ? rr <= th2 : rr < th2) continue;
          const double rr2 = rr * rr;
          if (C) {
            const cd c = mk(sc[m].x, sc[m].y);
            P += c * (log(rr) / 2);
            if (PG >= 2) { Gx += c * (dx / rr); Gy += c * (dy / rr); }
            if (PG >= 3) {
              Hxx += c * ((rr - 2 * dx2) / rr2);
              Hxy += c * (-2 * dx * dy / rr2);
              Hyy += c * ((rr - 2 * dy2) / rr2);
            }
          }
          if (Q) {
            const cd q = mk(sq[m].x, sq[m].y), d1 = q * sv[m].x, d2 = q * sv[m].y;
            P += d1 * (-dx / rr) + d2 * (-dy / rr);
            if (PG >= 2) {
              const double a1 = -(rr - 2 * dx2) / rr2, a2 = 2 * dx * dy / rr2;
              const double b2 = -(rr - 2 * dy2) / rr2;
              Gx += d1 * a1 + d2 * a2;
              Gy += d1 * a2 + d2 * b2;
            }
            if (PG >= 3) {
              const double rr3 = rr2 * rr;
              const double e1 = -2 * dx * (dx2 - 3 * dy2) / rr3, e2 = 2 * dy * (dy2 - 3 * dx2) / rr3;
              Hxx += d1 * e1 + d2 * e2;
              Hxy += d1 * e2 + d2 * (-e1);
              Hyy += d1 * (-e1) + d2 * (-e2);
            }
          }
        }
      __syncthreads();
    }
    if (act) {
      auto add = [](double2* o, cd v) { *o = make_double2(o->x + v.x, o->y + v.y); };
      add(pot + (size_t)j * nd + d, P);
      if (PG >= 2) {
        add(grad + ((size_t)j * 2 + 0) * nd + d, Gx);
        add(grad + ((size_t)j * 2 + 1) * nd + d, Gy);
      }
      if (PG >= 3) {
        add(hess + ((size_t)j * 3 + 0) * nd + d, Hxx);
        add(hess + ((size_t)j * 3 + 1) * nd + d, Hxy);
        add(hess + ((size_t)j * 3 + 2) * nd + d, Hyy);
      }
    }
  }
}

// Helmholtz, (i/4) H0(k r) kernel (helmkernels2d.c): charge (i/4) c H0,
// dipole (i/4) d k H1(kr) (x - y).v / r; the potential-only charge routine
// skips on |z| <= thresh, every other routine on <
template <bool C, bool Q, int PG>
__global__ void __launch_bounds__(TILE) k_h2d(cd zk, const double2* src, int ns,
                                              const double2* chg, const double2* dip,
                                              const double* dv, const double2* trg, int nt,
                                              int nd, bool le, double th, double2* pot,
                                              double2* grad, double2* hess, int len, size_t sp, size_t sg, size_t sh) {
  // sources of this block's chunk; outputs of chunk y go to slice y
  const int sb = blockIdx.y * len, se = min(ns, sb + len);
  pot += blockIdx.y * sp;
  if (PG >= 2) grad += blockIdx.y * sg;
  if (PG >= 3) hess += blockIdx.y * sh;
  __shared__ double2 sx[TILE];
  __shared__ double2 sc[TILE];
  __shared__ double2 sq[TILE];
  __shared__ double2 sv[TILE];
  const int j = blockIdx.x * TILE + threadIdx.x;
  const bool act = j < nt;
  const double2 x = act ? trg[j] : make_double2(0, 0);
  const cd i4 = mk(0, 0.25), ik4 = zk * mk(0, 0.25);
  for (int d = 0; d < nd; ++d) {
    cd Pc = mk(0, 0), Pd = mk(0, 0), Gx = mk(0, 0), Gy = mk(0, 0);
    cd Hxx = mk(0, 0), Hxy = mk(0, 0), Hyy = mk(0, 0);
    for (int s0 = sb; s0 < se; s0 += TILE) {
      int cnt = min(TILE, se - s0);
      if ((int)threadIdx.x < cnt) {
        const size_t i = s0 + threadIdx.x;
        sx[threadIdx.x] = src[i];
        if (C) sc[threadIdx.x] = chg[i * nd + d];
        if (Q) {
          sq[threadIdx.x] = dip[i * nd + d];
          sv[threadIdx.x] = make_double2(dv[(i * 2 + 0) * nd + d], dv[(i * 2 + 1) * nd + d]);
        }
      }
      __syncthreads();
      if (act)
        for (int m = 0; m < cnt; ++m) {
          const double dx = x.x - sx[m].x, dy = x.y - sx[m].y;
          const double rr = dx * dx + dy * dy, r = sqrt(rr);
          const cd z = zk * r;
          const double az = cabs_(z);
          if (le ? az <= th : az < th) continue;
          cd h0, h1;
          hank01(z, h0, h1);
          if (C) {
            const cd c = mk(sc[m].x, sc[m].y);
            Pc += h0 * c;
            if (PG >= 2) {
              const cd g = -(h1 * c * (1.0 / r));
              Gx += g * dx;
              Gy += g * dy;
            }
            if (PG >= 3) {
              const cd f = c * (1.0 / (r * rr)), h2z = -(z * h0) + h1 * 2.0;
              Hxx += f * (h2z * (dx * dx) - h1 * rr);
              Hxy += f * (h2z * (dx * dy));
              Hyy += f * (h2z * (dy * dy) - h1 * rr);
            }
          }
          if (Q) {
            const cd q = mk(sq[m].x, sq[m].y);
            const double v1 = sv[m].x, v2 = sv[m].y, ct = dx / r, st = dy / r;
            Pd += h1 * (1.0 / r) * q * (dx * v1 + dy * v2);
            if (PG >= 2) {
              const cd h2 = h1 * inv(z) * 2.0 - h0;
              const cd f1 = h2 * (ct * ct - 0.5) - h0 * 0.5, f2 = h2 * (ct * st);
              const cd f3 = h2 * (st * st - 0.5) - h0 * 0.5, qk = -(q * zk);
              Gx += qk * (f1 * v1 + f2 * v2);
              Gy += qk * (f2 * v1 + f3 * v2);
              if (PG >= 3) {
                const cd h3 = h2 * inv(z) * 4.0 - h1;
                const cd a = (h1 * 0.75 - h3 * (0.5 * (ct * ct - 0.5 - st * st))) * ct;
                const cd bb = (h1 * 0.25 - h3 * (0.5 * (1.5 * ct * ct - 0.5 * st * st))) * st;
                const cd c2 = (h1 * 0.25 - h3 * (0.5 * (1.5 * st * st - 0.5 * ct * ct))) * ct;
                const cd e = (h1 * 0.75 - h3 * (0.5 * (st * st - 0.5 - ct * ct))) * st;
                const cd qk2 = -(q * zk * zk);
                Hxx += qk2 * (a * v1 + bb * v2);
                Hxy += qk2 * (bb * v1 + c2 * v2);
                Hyy += qk2 * (c2 * v1 + e * v2);
              }
            }
          }
        }
      __syncthreads();
    }
    if (act) {
      auto add = [](double2* o, cd v) { *o = make_double2(o->x + v.x, o->y + v.y); };
      add(pot + (size_t)j * nd + d, Pc * i4 + Pd * ik4);
      if (PG >= 2) {
        add(grad + ((size_t)j * 2 + 0) * nd + d, Gx * ik4);
        add(grad + ((size_t)j * 2 + 1) * nd + d, Gy * ik4);
      }
      if (PG >= 3) {
        add(hess + ((size_t)j * 3 + 0) * nd + d, Hxx * ik4);
        add(hess + ((size_t)j * 3 + 1) * nd + d, Hxy * ik4);
        add(hess + ((size_t)j * 3 + 2) * nd + d, Hyy * ik4);
      }
    }
  }
}

// Laplace log kernel with real densities and outputs (rlapkernels2d.c:16-491):
// charge c: pot += c log(rr)/2, grad += c (dx, dy)/rr, hess += c ((rr - 2dx^2),
// -2dxdy, (rr - 2dy^2))/rr^2; dipole d1 = q v1, d2 = q v2: pot += d1 p1 + d2 p2
// with p = -(dx, dy)/rr, and its derivatives.  le: skip on rr <= th^2, else <.
template <bool C, bool Q, int PG>
__global__ void __launch_bounds__(TILE) k_r2d(const double2* src, int ns, const double* chg,
                                              const double* dip, const double* dv,
                                              const double2* trg, int nt, int nd, bool le,
                                              double th2, double* pot, double* grad,
                                              double* hess, int len, size_t sp, size_t sg, size_t sh) {
  // sources of this block's chunk; outputs of chunk y go to slice y
  const int sb = blockIdx.y * len, se = min(ns, sb + len);
  pot += blockIdx.y * sp;
  if (PG >= 2) grad += blockIdx.y * sg;
  if (PG >= 3) hess += blockIdx.y * sh;
  __shared__ double2 sx[TILE];
  __shared__ double sc[TILE];
  __shared__ double2 sd[TILE];  // (q v1, q v2)
  const int j = blockIdx.x * TILE + threadIdx.x;
  const bool act = j < nt;
  const double2 x = act ? trg[j] : make_double2(0, 0);
  for (int d = 0; d < nd; ++d) {
    double P = 0, Gx = 0, Gy = 0, Hxx = 0, Hxy = 0, Hyy = 0;
    for (int s0 = sb; s0 < se; s0 += TILE) {
      int cnt = min(TILE, se - s0);
      if ((int)threadIdx.x < cnt) {
        const size_t i = s0 + threadIdx.x;
        sx[threadIdx.x] = src[i];
        if (C) sc[threadIdx.x] = chg[i * nd + d];
        if (Q) {
          const double q = dip[i * nd + d];
          sd[threadIdx.x] = make_double2(q * dv[(i * 2 + 0) * nd + d], q * dv[(i * 2 + 1) * nd + d]);
        }
      }
      __syncthreads();
      if (act)
        for (int m = 0; m < cnt; ++m) {
          const double dx = x.x - sx[m].x, dy = x.y - sx[m].y;
          const double dx2 = dx * dx, dy2 = dy * dy, rr = dx2 + dy2;
          if (le ? rr <= th2 : rr < th2) continue;
          const double rr2 = rr * rr;
          if (C) {
            const double c = sc[m];
            P += c * (log(rr) / 2);
            if (PG >= 2) { Gx += c * (dx / rr); Gy += c * (dy / rr); }
            if (PG >= 3) {
              Hxx += c * ((rr - 2 * dx2) / rr2);
              Hxy += c * (-2 * dx * dy / rr2);
              Hyy += c * ((rr - 2 * dy2) / rr2);
            }
          }
          if (Q) {
            const double d1 = sd[m].x, d2 = sd[m].y;
            P += d1 * (-dx / rr) + d2 * (-dy / rr);
            if (PG >= 2) {
              const double a1 = -(rr - 2 * dx2) / rr2, a2 = 2 * dx * dy / rr2;
              const double b2 = -(rr - 2 * dy2) / rr2;
              Gx += d1 * a1 + d2 * a2;
              Gy += d1 * a2 + d2 * b2;
            }
            if (PG >= 3) {
              const double rr3 = rr2 * rr;
              const double e1 = -2 * dx * (dx2 - 3 * dy2) / rr3, e2 = 2 * dy * (dy2 - 3 * dx2) / rr3;
              Hxx += d1 * e1 + d2 * e2;
              Hxy += d1 * e2 - d2 * e1;
              Hyy += -(d1 * e1) - d2 * e2;
            }
          }
        }
      __syncthreads();
    }
    if (act) {
      pot[(size_t)j * nd + d] += P;
      if (PG >= 2) {
        grad[((size_t)j * 2 + 0) * nd + d] += Gx;
        grad[((size_t)j * 2 + 1) * nd + d] += Gy;
      }
      if (PG >= 3) {
        hess[((size_t)j * 3 + 0) * nd + d] += Hxx;
        hess[((size_t)j * 3 + 1) * nd + d] += Hxy;
        hess[((size_t)j * 3 + 2) * nd + d] += Hyy;
      }
    }
  }
}

// Sources are split into chunks over the grid's y dimension when the
// targets alone give too few blocks (a few targets against many sources):
// chunk y writes its partial sums to slice y of zeroed buffers, which
// k_dsum then adds to the outputs in chunk order (deterministic).
struct Chunks {
  int n = 1, len = 0;
};
inline Chunks chunks_for(int ns, int nt) {
  const int tb = (nt + TILE - 1) / TILE;
  const int n = std::min((ns + 1023) / 1024, (2048 + tb - 1) / tb);
  if (n <= 1) return {1, ns};
  const int len = ((ns + n - 1) / n + TILE - 1) / TILE * TILE;
  return {(ns + len - 1) / len, len};
}
__global__ void k_dsum(const double* part, int nchunk, size_t n, double* out) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double a = out[e];
  for (int y = 0; y < nchunk; ++y) a += part[y * n + e];
  out[e] = a;
}
// Launches kern(grid, pot, grad, hess, len, slice strides) over the chunks
// of ns sources; the outputs are no (pot), no*ncg (grad) and no*nch (hess)
// elements of T (double2 or double), incremented.
template <class T, class K>
hipError_t launch_chunked(hipStream_t st, Scratch& S, int ns, int nt, size_t no, int ncg,
                          int nch, int PG, T* po, T* go, T* ho, K&& kern) {
  const Chunks ch = chunks_for(ns, nt);
  const unsigned tb = (unsigned)((nt + TILE - 1) / TILE);
  if (ch.n == 1) {
    kern(dim3(tb, 1), po, go, ho, ns, size_t(0), size_t(0), size_t(0));
    return hipGetLastError();
  }
  const size_t ng = PG >= 2 ? no * ncg : 0, nh = PG >= 3 ? no * nch : 0;
  T* pp = S.get<T>(ch.n * no);
  T* pg = ng ? S.get<T>(ch.n * ng) : nullptr;
  T* ph = nh ? S.get<T>(ch.n * nh) : nullptr;
  if (S.err) return hipErrorOutOfMemory;
  hipError_t e = hipMemsetAsync(pp, 0, sizeof(T) * ch.n * no, st);
  if (e == hipSuccess && pg) e = hipMemsetAsync(pg, 0, sizeof(T) * ch.n * ng, st);
  if (e == hipSuccess && ph) e = hipMemsetAsync(ph, 0, sizeof(T) * ch.n * nh, st);
  if (e != hipSuccess) return e;
  kern(dim3(tb, (unsigned)ch.n), pp, pg, ph, ch.len, no, ng, nh);
  constexpr int w = sizeof(T) / sizeof(double);
  auto sum = [&](const T* part, size_t n, T* out) {
    const size_t m = n * w;
    k_dsum<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(reinterpret_cast<const double*>(part),
                                                       ch.n, m, reinterpret_cast<double*>(out));
  };
  sum(pp, no, po);
  if (pg) sum(pg, ng, go);
  if (ph) sum(ph, nh, ho);
  return hipGetLastError();
}

// Stokes (stokkernels2d.c:11-173): Stokeslets f and stresslets (mu, nu) at
// the sources, real outputs velocity pot (nd,2,nt), pressure pre (nd,nt)
// and velocity gradient grad (nd,2,2,nt), each incremented; pairs with
// r^2 < thresh^2 skipped.  With d = t - s, r2 = |d|^2, rt = d1^2 - d2^2:
//   Stokeslet: pot -= f log(r2)/4 - d (d.f)/(2 r2); pre += (d.f)/r2;
//     grad11 -= rt (d.f)/(2 r2^2), grad22 += the same,
//     grad21 += (pl2 rt - 4 f1 d2 d1^2)/(2 r2^2),
//     grad12 -= (pl2 rt + 4 f2 d1 d2^2)/(2 r2^2), pl2 = d1 f2 + d2 f1;
//   stresslet, pl = d.mu, pl2 = d.nu, pv = mu.nu:
//     pot -= 2 d pl pl2 / r2^2; pre -= 4 pl pl2 / r2^2 - 2 pv / r2;
//     grad(a,b) -= 2 d_b (mu_a pl2 + nu_a pl)/r2^2 - 8 d_a d_b pl pl2/r2^3
//       (plus 2 pl pl2 / r2^2 on the diagonal).
// grad(a, b) = d u_b / d x_a is grad[((t*2 + b)*2 + a)*nd + d] (column-major
// (nd,2,2,nt)).
// Chunk slices as the other direct sums: pre is the first output, pot the
// second, grad the third (launch_chunked's pot / grad / hess).
template <bool STOK, bool STRS>
__global__ void __launch_bounds__(TILE) k_st2d(const double2* src, int ns, const double* stok,
                                               const double* mu, const double* nu,
                                               const double2* trg, int nt, int nd, double th2,
                                               double* pre, double* pot, double* grad, int len,
                                               size_t sp, size_t sg, size_t sh) {
  const int sb = blockIdx.y * len, se = min(ns, sb + len);
  pre += blockIdx.y * sp;
  pot += blockIdx.y * sg;
  grad += blockIdx.y * sh;
  __shared__ double2 sx[TILE];
  __shared__ double2 sf[TILE];
  __shared__ double2 smu[TILE];
  __shared__ double2 snu[TILE];
  const int j = blockIdx.x * TILE + threadIdx.x;
  const bool act = j < nt;
  const double2 x = act ? trg[j] : make_double2(0, 0);
  for (int d = 0; d < nd; ++d) {
    double u1 = 0, u2 = 0, pr = 0, g11 = 0, g21 = 0, g12 = 0, g22 = 0;
    for (int s0 = sb; s0 < se; s0 += TILE) {
      const int cnt = min(TILE, se - s0);
      if ((int)threadIdx.x < cnt) {
        const size_t i = s0 + threadIdx.x;
        sx[threadIdx.x] = src[i];
        if (STOK) sf[threadIdx.x] = make_double2(stok[(i * 2 + 0) * nd + d], stok[(i * 2 + 1) * nd + d]);
        if (STRS) {
          smu[threadIdx.x] = make_double2(mu[(i * 2 + 0) * nd + d], mu[(i * 2 + 1) * nd + d]);
          snu[threadIdx.x] = make_double2(nu[(i * 2 + 0) * nd + d], nu[(i * 2 + 1) * nd + d]);
        }
      }
      __syncthreads();
      if (act)
        for (int m = 0; m < cnt; ++m) {
          const double d1 = x.x - sx[m].x, d2 = x.y - sx[m].y;
          const double r2 = d1 * d1 + d2 * d2;
          if (r2 < th2) continue;
          const double rt = d1 * d1 - d2 * d2, r4 = r2 * r2;
          if (STOK) {
            const double f1 = sf[m].x, f2 = sf[m].y;
            const double lg = log_pos(r2) / 4.0;
            const double pl = d1 * f1 + d2 * f2, pl2 = d1 * f2 + d2 * f1;
            u1 += d1 * pl / r2 / 2.0 - f1 * lg;
            u2 += d2 * pl / r2 / 2.0 - f2 * lg;
            g11 -= rt * pl / r4 / 2.0;
            g22 += rt * pl / r4 / 2.0;
            g21 += (pl2 * rt - f1 * 4.0 * d2 * (d1 * d1)) / 2.0 / r4;
            g12 -= (pl2 * rt + f2 * 4.0 * d1 * (d2 * d2)) / 2.0 / r4;
            pr += pl / r2;
          }
          if (STRS) {
            const double m1 = smu[m].x, m2 = smu[m].y, n1 = snu[m].x, n2 = snu[m].y;
            const double pl = d1 * m1 + d2 * m2, pl2 = d1 * n1 + d2 * n2;
            const double pv = m1 * n1 + m2 * n2, r6 = r4 * r2;
            const double q = pl * pl2;
            u1 -= 2 * d1 * q / r4;
            u2 -= 2 * d2 * q / r4;
            g11 -= 2 * q / r4 + 2 * d1 * m1 * pl2 / r4 + 2 * d1 * n1 * pl / r4 - 8 * d1 * d1 * q / r6;
            g21 -= 2 * d1 * m2 * pl2 / r4 + 2 * d1 * n2 * pl / r4 - 8 * d1 * d2 * q / r6;
            g12 -= 2 * d2 * m1 * pl2 / r4 + 2 * d2 * n1 * pl / r4 - 8 * d1 * d2 * q / r6;
            g22 -= 2 * q / r4 + 2 * d2 * m2 * pl2 / r4 + 2 * d2 * n2 * pl / r4 - 8 * d2 * d2 * q / r6;
            pr -= 4 * q / r4 - 2 * pv / r2;
          }
        }
      __syncthreads();
    }
    if (act) {
      const size_t t = (size_t)j;
      pot[(t * 2 + 0) * nd + d] += u1;
      pot[(t * 2 + 1) * nd + d] += u2;
      pre[t * nd + d] += pr;
      grad[((t * 2 + 0) * 2 + 0) * nd + d] += g11;
      grad[((t * 2 + 0) * 2 + 1) * nd + d] += g21;
      grad[((t * 2 + 1) * 2 + 0) * nd + d] += g12;
      grad[((t * 2 + 1) * 2 + 1) * nd + d] += g22;
    }
  }
}

// Stokes direct sums through the host staging of the other routines
template <bool STOK, bool STRS>
int st2d(int nd, const double* sources, int ns, const double* stok, const double* mu,
         const double* nu, const double* targ, int nt, double* pot, double* pre, double* grad,
         double th) {
  if (nd < 1 || ns < 0 || nt < 0) return IER_ARGS;
  if (ns == 0 || nt == 0) return IER_OK;
  Ctx c;
  int e = get_ctx(c);
  if (e) return e;
  Scratch S(c.st);
  const size_t nq = (size_t)2 * nd * ns, no = (size_t)nd * nt;
  double2* s = S.get<double2>(ns);
  double2* t = S.get<double2>(nt);
  double* fs = STOK ? S.get<double>(nq) : nullptr;
  double* ms = STRS ? S.get<double>(nq) : nullptr;
  double* vs = STRS ? S.get<double>(nq) : nullptr;
  double* pr = S.get<double>(no);
  double* po = S.get<double>(2 * no);
  double* go = S.get<double>(4 * no);
  if (S.err) return S.err;
  HIP_TRY(hipMemcpyAsync(s, sources, 16 * (size_t)ns, hipMemcpyHostToDevice, c.st));
  HIP_TRY(hipMemcpyAsync(t, targ, 16 * (size_t)nt, hipMemcpyHostToDevice, c.st));
  if (STOK) HIP_TRY(hipMemcpyAsync(fs, stok, 8 * nq, hipMemcpyHostToDevice, c.st));
  if (STRS) {
    HIP_TRY(hipMemcpyAsync(ms, mu, 8 * nq, hipMemcpyHostToDevice, c.st));
    HIP_TRY(hipMemcpyAsync(vs, nu, 8 * nq, hipMemcpyHostToDevice, c.st));
  }
  HIP_TRY(hipMemcpyAsync(pr, pre, 8 * no, hipMemcpyHostToDevice, c.st));
  HIP_TRY(hipMemcpyAsync(po, pot, 16 * no, hipMemcpyHostToDevice, c.st));
  HIP_TRY(hipMemcpyAsync(go, grad, 32 * no, hipMemcpyHostToDevice, c.st));
  HIP_TRY(launch_chunked(c.st, S, ns, nt, no, 2, 4, 3, pr, po, go,
                         [&](dim3 grid, double* a, double* b, double* g, int len, size_t x,
                             size_t y, size_t z) {
                           k_st2d<STOK, STRS><<<grid, TILE, 0, c.st>>>(
                               s, ns, fs, ms, vs, t, nt, nd, th * th, a, b, g, len, x, y, z);
                         }));
  HIP_TRY(hipMemcpyAsync(pre, pr, 8 * no, hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipMemcpyAsync(pot, po, 16 * no, hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipMemcpyAsync(grad, go, 32 * no, hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipStreamSynchronize(c.st));
  return IER_OK;
}

// host staging for the direct sums: copy in, run, copy the outputs back
struct DirectIO {
  int nd, ns, nt, ncg, nch;  // gradient / Hessian components per output
  const double *sources, *targ, *dipvec;
  const void *charge, *dipstr;
  void *pot, *grad, *hess;
};

template <class Launch>
int run_direct(const DirectIO& io, bool C, bool Q, int PG, Launch&& launch) {
  if (io.nd < 1 || io.ns < 0 || io.nt < 0) return IER_ARGS;
  if (io.ns == 0 || io.nt == 0) return IER_OK;
  Ctx c;
  int e = get_ctx(c);
  if (e) return e;
  Scratch S(c.st);
  const int nd = io.nd, ns = io.ns, nt = io.nt;
  double2* s = S.get<double2>(ns);
  double2* t = S.get<double2>(nt);
  double2* cc = C ? S.get<double2>((size_t)nd * ns) : nullptr;
  double2* qq = Q ? S.get<double2>((size_t)nd * ns) : nullptr;
  double* vv = Q ? S.get<double>((size_t)2 * nd * ns) : nullptr;
  const size_t no = (size_t)nd * nt;
  double2* po = S.get<double2>(no);
  double2* go = PG >= 2 ? S.get<double2>(no * io.ncg) : nullptr;
  double2* ho = PG >= 3 ? S.get<double2>(no * io.nch) : nullptr;
  if (S.err) return S.err;
  HIP_TRY(hipMemcpyAsync(s, io.sources, 16 * (size_t)ns, hipMemcpyHostToDevice, c.st));
  HIP_TRY(hipMemcpyAsync(t, io.targ, 16 * (size_t)nt, hipMemcpyHostToDevice, c.st));
  if (C) HIP_TRY(hipMemcpyAsync(cc, io.charge, 16 * (size_t)nd * ns, hipMemcpyHostToDevice, c.st));
  if (Q) {
    HIP_TRY(hipMemcpyAsync(qq, io.dipstr, 16 * (size_t)nd * ns, hipMemcpyHostToDevice, c.st));
    HIP_TRY(hipMemcpyAsync(vv, io.dipvec, 16 * (size_t)nd * ns, hipMemcpyHostToDevice, c.st));
  }
  HIP_TRY(hipMemcpyAsync(po, io.pot, 16 * no, hipMemcpyHostToDevice, c.st));
  if (go) HIP_TRY(hipMemcpyAsync(go, io.grad, 16 * no * io.ncg, hipMemcpyHostToDevice, c.st));
  if (ho) HIP_TRY(hipMemcpyAsync(ho, io.hess, 16 * no * io.nch, hipMemcpyHostToDevice, c.st));
  HIP_TRY(launch_chunked(c.st, S, ns, nt, no, io.ncg, io.nch, PG, po, go, ho,
                         [&](dim3 grid, double2* p, double2* g, double2* h, int len, size_t a,
                             size_t b, size_t d) { launch(c.st, grid, s, cc, qq, vv, t, p, g, h,
                                                          len, a, b, d); }));
  HIP_TRY(hipMemcpyAsync(io.pot, po, 16 * no, hipMemcpyDeviceToHost, c.st));
  if (go) HIP_TRY(hipMemcpyAsync(io.grad, go, 16 * no * io.ncg, hipMemcpyDeviceToHost, c.st));
  if (ho) HIP_TRY(hipMemcpyAsync(io.hess, ho, 16 * no * io.nch, hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipStreamSynchronize(c.st));
  return IER_OK;
}

template <bool C, bool Q, int PG>
int l2d(int nd, const double* sources, int ns, const void* charge, const void* dipstr,
        const double* dipvec, const double* targ, int nt, void* pot, void* grad, void* hess,
        double th, bool le) {
  DirectIO io{nd, ns, nt, 2, 3, sources, targ, dipvec, charge, dipstr, pot, grad, hess};
  return run_direct(io, C, Q, PG, [&](hipStream_t st, dim3 grid, const double2* s,
                                      const double2* cc, const double2* qq, const double* vv,
                                      const double2* t, double2* po, double2* go, double2* ho,
                                      int len, size_t a, size_t b, size_t d) {
    k_l2d<C, Q, PG><<<grid, TILE, 0, st>>>(s, ns, cc, qq, vv, t, nt, nd, le, th * th, po, go, ho,
                                           len, a, b, d);
  });
}

template <bool C, bool Q, int PG>
int h2d(int nd, const fmm2d_cplx* zk, const double* sources, int ns, const void* charge,
        const void* dipstr, const double* dipvec, const double* targ, int nt, void* pot,
        void* grad, void* hess, double th, bool le) {
  const double* z = reinterpret_cast<const double*>(zk);
  const cd k = mk(z[0], z[1]);
  DirectIO io{nd, ns, nt, 2, 3, sources, targ, dipvec, charge, dipstr, pot, grad, hess};
  return run_direct(io, C, Q, PG, [&](hipStream_t st, dim3 grid, const double2* s,
                                      const double2* cc, const double2* qq, const double* vv,
                                      const double2* t, double2* po, double2* go, double2* ho,
                                      int len, size_t a, size_t b, size_t d) {
    k_h2d<C, Q, PG><<<grid, TILE, 0, st>>>(k, s, ns, cc, qq, vv, t, nt, nd, le, th, po, go, ho,
                                           len, a, b, d);
  });
}

// real-valued direct sums: real strengths, dipole vectors and outputs
template <bool C, bool Q, int PG>
int r2d(int nd, const double* sources, int ns, const double* charge, const double* dipstr,
        const double* dipvec, const double* targ, int nt, double* pot, double* grad,
        double* hess, double th, bool le) {
  if (nd < 1 || ns < 0 || nt < 0) return IER_ARGS;
  if (ns == 0 || nt == 0) return IER_OK;
  Ctx c;
  int e = get_ctx(c);
  if (e) return e;
  Scratch S(c.st);
  const size_t nq = (size_t)nd * ns, no = (size_t)nd * nt;
  double2* s = S.get<double2>(ns);
  double2* t = S.get<double2>(nt);
  double* cc = C ? S.get<double>(nq) : nullptr;
  double* qq = Q ? S.get<double>(nq) : nullptr;
  double* vv = Q ? S.get<double>(2 * nq) : nullptr;
  double* po = S.get<double>(no);
  double* go = PG >= 2 ? S.get<double>(2 * no) : nullptr;
  double* ho = PG >= 3 ? S.get<double>(3 * no) : nullptr;
  if (S.err) return S.err;
  HIP_TRY(hipMemcpyAsync(s, sources, 16 * (size_t)ns, hipMemcpyHostToDevice, c.st));
  HIP_TRY(hipMemcpyAsync(t, targ, 16 * (size_t)nt, hipMemcpyHostToDevice, c.st));
  if (C) HIP_TRY(hipMemcpyAsync(cc, charge, 8 * nq, hipMemcpyHostToDevice, c.st));
  if (Q) {
    HIP_TRY(hipMemcpyAsync(qq, dipstr, 8 * nq, hipMemcpyHostToDevice, c.st));
    HIP_TRY(hipMemcpyAsync(vv, dipvec, 16 * nq, hipMemcpyHostToDevice, c.st));
  }
  HIP_TRY(hipMemcpyAsync(po, pot, 8 * no, hipMemcpyHostToDevice, c.st));
  if (go) HIP_TRY(hipMemcpyAsync(go, grad, 16 * no, hipMemcpyHostToDevice, c.st));
  if (ho) HIP_TRY(hipMemcpyAsync(ho, hess, 24 * no, hipMemcpyHostToDevice, c.st));
  HIP_TRY(launch_chunked(c.st, S, ns, nt, no, 2, 3, PG, po, go, ho,
                         [&](dim3 grid, double* p, double* g, double* h, int len, size_t a,
                             size_t b, size_t d) {
                           k_r2d<C, Q, PG><<<grid, TILE, 0, c.st>>>(s, ns, cc, qq, vv, t, nt, nd,
                                                                    le, th * th, p, g, h, len, a,
                                                                    b, d);
                         }));
  HIP_TRY(hipMemcpyAsync(pot, po, 8 * no, hipMemcpyDeviceToHost, c.st));
  if (go) HIP_TRY(hipMemcpyAsync(grad, go, 16 * no, hipMemcpyDeviceToHost, c.st));
  if (ho) HIP_TRY(hipMemcpyAsync(hess, ho, 24 * no, hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipStreamSynchronize(c.st));
  return IER_OK;
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
  HIP_TRY(launch_chunked(c.st, S, ns, nt, no, 1, 1, PG, po, go, ho,
                         [&](dim3 grid, double2* p, double2* g, double2* h, int len, size_t a,
                             size_t b, size_t d) {
                           k_c2d<C, Q, PG><<<grid, TILE, 0, c.st>>>(s, ns, cc, qq, t, nt, nd, mode,
                                                                    th, p, g, h, len, a, b, d);
                         }));
  HIP_TRY(hipMemcpyAsync(pot, po, 16 * no, hipMemcpyDeviceToHost, c.st));
  if (go) HIP_TRY(hipMemcpyAsync(grad, go, 16 * no, hipMemcpyDeviceToHost, c.st));
  if (ho) HIP_TRY(hipMemcpyAsync(hess, ho, 16 * no, hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipStreamSynchronize(c.st));
  return IER_OK;
}

}  // namespace

int h2d_direct_cdg(int nd, const fmm2d_cplx* zk, const double* s, int ns, const fmm2d_cplx* c,
                   const fmm2d_cplx* q, const double* v, const double* t, int nt,
                   fmm2d_cplx* pot, fmm2d_cplx* grad, double th) {
  return h2d<true, true, 2>(nd, zk, s, ns, c, q, v, t, nt, pot, grad, nullptr, th, false);
}

}  // namespace fmm

extern "C" {
void st2ddirectstokg_(const fmm2d_int* nd, const double* sources, const double* stoklet,
                      const fmm2d_int* ns, const double* targ, const fmm2d_int* nt, double* pot,
                      double* pre, double* grad, const double* thresh) {
  fmm::st2d<true, false>(*nd, sources, *ns, stoklet, nullptr, nullptr, targ, *nt, pot, pre, grad,
                         *thresh);
}
void st2ddirectstokstrsg_(const fmm2d_int* nd, const double* sources, const fmm2d_int* ifstoklet,
                          const double* stoklet, const fmm2d_int* istress, const double* strslet,
                          const double* strsvec, const fmm2d_int* ns, const double* targ,
                          const fmm2d_int* nt, double* pot, double* pre, double* grad,
                          const double* thresh) {
  const bool f = *ifstoklet == 1, m = *istress == 1;
  if (f && m)
    fmm::st2d<true, true>(*nd, sources, *ns, stoklet, strslet, strsvec, targ, *nt, pot, pre, grad,
                          *thresh);
  else if (f)
    fmm::st2d<true, false>(*nd, sources, *ns, stoklet, nullptr, nullptr, targ, *nt, pot, pre,
                           grad, *thresh);
  else if (m)
    fmm::st2d<false, true>(*nd, sources, *ns, nullptr, strslet, strsvec, targ, *nt, pot, pre,
                           grad, *thresh);
}
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
void l2d_directcp_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, const double* th) {
  fmm::l2d<true, false, 1>(*nd, s, *ns, c, nullptr, nullptr, t, *nt, pot, nullptr, nullptr, *th, true);
}
void h2d_directcp_(const fmm2d_int* nd, const fmm2d_cplx* zk, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, const double* th) {
  fmm::h2d<true, false, 1>(*nd, zk, s, *ns, c, nullptr, nullptr, t, *nt, pot, nullptr, nullptr, *th, true);
}
void l2d_directcg_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, const double* th) {
  fmm::l2d<true, false, 2>(*nd, s, *ns, c, nullptr, nullptr, t, *nt, pot, grad, nullptr, *th, false);
}
void h2d_directcg_(const fmm2d_int* nd, const fmm2d_cplx* zk, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, const double* th) {
  fmm::h2d<true, false, 2>(*nd, zk, s, *ns, c, nullptr, nullptr, t, *nt, pot, grad, nullptr, *th, false);
}
void l2d_directch_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, const double* th) {
  fmm::l2d<true, false, 3>(*nd, s, *ns, c, nullptr, nullptr, t, *nt, pot, grad, hess, *th, false);
}
void h2d_directch_(const fmm2d_int* nd, const fmm2d_cplx* zk, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, const double* th) {
  fmm::h2d<true, false, 3>(*nd, zk, s, *ns, c, nullptr, nullptr, t, *nt, pot, grad, hess, *th, false);
}
void l2d_directdp_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, const double* th) {
  fmm::l2d<false, true, 1>(*nd, s, *ns, nullptr, q, v, t, *nt, pot, nullptr, nullptr, *th, true);
}
void h2d_directdp_(const fmm2d_int* nd, const fmm2d_cplx* zk, const double* s, const fmm2d_int* ns, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, const double* th) {
  fmm::h2d<false, true, 1>(*nd, zk, s, *ns, nullptr, q, v, t, *nt, pot, nullptr, nullptr, *th, false);
}
void l2d_directdg_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, const double* th) {
  fmm::l2d<false, true, 2>(*nd, s, *ns, nullptr, q, v, t, *nt, pot, grad, nullptr, *th, true);
}
void h2d_directdg_(const fmm2d_int* nd, const fmm2d_cplx* zk, const double* s, const fmm2d_int* ns, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, const double* th) {
  fmm::h2d<false, true, 2>(*nd, zk, s, *ns, nullptr, q, v, t, *nt, pot, grad, nullptr, *th, false);
}
void l2d_directdh_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, const double* th) {
  fmm::l2d<false, true, 3>(*nd, s, *ns, nullptr, q, v, t, *nt, pot, grad, hess, *th, true);
}
void h2d_directdh_(const fmm2d_int* nd, const fmm2d_cplx* zk, const double* s, const fmm2d_int* ns, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, const double* th) {
  fmm::h2d<false, true, 3>(*nd, zk, s, *ns, nullptr, q, v, t, *nt, pot, grad, hess, *th, false);
}
void l2d_directcdp_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, const double* th) {
  fmm::l2d<true, true, 1>(*nd, s, *ns, c, q, v, t, *nt, pot, nullptr, nullptr, *th, true);
}
void h2d_directcdp_(const fmm2d_int* nd, const fmm2d_cplx* zk, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, const double* th) {
  fmm::h2d<true, true, 1>(*nd, zk, s, *ns, c, q, v, t, *nt, pot, nullptr, nullptr, *th, false);
}
void l2d_directcdg_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, const double* th) {
  fmm::l2d<true, true, 2>(*nd, s, *ns, c, q, v, t, *nt, pot, grad, nullptr, *th, false);
}
void h2d_directcdg_(const fmm2d_int* nd, const fmm2d_cplx* zk, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, const double* th) {
  fmm::h2d<true, true, 2>(*nd, zk, s, *ns, c, q, v, t, *nt, pot, grad, nullptr, *th, false);
}
void l2d_directcdh_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, const double* th) {
  fmm::l2d<true, true, 3>(*nd, s, *ns, c, q, v, t, *nt, pot, grad, hess, *th, false);
}
void h2d_directcdh_(const fmm2d_int* nd, const fmm2d_cplx* zk, const double* s, const fmm2d_int* ns, const fmm2d_cplx* c, const fmm2d_cplx* q, const double* v, const double* t, const fmm2d_int* nt, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, const double* th) {
  fmm::h2d<true, true, 3>(*nd, zk, s, *ns, c, q, v, t, *nt, pot, grad, hess, *th, false);
}
// rlapkernels2d.c: skip operators per routine -- cp, dp, dg, dh, cdp: rr <= th^2
// (:33,149,191,251,318); cg, ch, cdg, cdh: rr < th^2 (:65,106,364,434)
void r2d_directcp_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const double* c, const double* t, const fmm2d_int* nt, double* pot, const double* th) {
  fmm::r2d<true, false, 1>(*nd, s, *ns, c, nullptr, nullptr, t, *nt, pot, nullptr, nullptr, *th, true);
}
void r2d_directcg_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const double* c, const double* t, const fmm2d_int* nt, double* pot, double* grad, const double* th) {
  fmm::r2d<true, false, 2>(*nd, s, *ns, c, nullptr, nullptr, t, *nt, pot, grad, nullptr, *th, false);
}
void r2d_directch_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const double* c, const double* t, const fmm2d_int* nt, double* pot, double* grad, double* hess, const double* th) {
  fmm::r2d<true, false, 3>(*nd, s, *ns, c, nullptr, nullptr, t, *nt, pot, grad, hess, *th, false);
}
void r2d_directdp_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const double* q, const double* v, const double* t, const fmm2d_int* nt, double* pot, const double* th) {
  fmm::r2d<false, true, 1>(*nd, s, *ns, nullptr, q, v, t, *nt, pot, nullptr, nullptr, *th, true);
}
void r2d_directdg_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const double* q, const double* v, const double* t, const fmm2d_int* nt, double* pot, double* grad, const double* th) {
  fmm::r2d<false, true, 2>(*nd, s, *ns, nullptr, q, v, t, *nt, pot, grad, nullptr, *th, true);
}
void r2d_directdh_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const double* q, const double* v, const double* t, const fmm2d_int* nt, double* pot, double* grad, double* hess, const double* th) {
  fmm::r2d<false, true, 3>(*nd, s, *ns, nullptr, q, v, t, *nt, pot, grad, hess, *th, true);
}
void r2d_directcdp_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const double* c, const double* q, const double* v, const double* t, const fmm2d_int* nt, double* pot, const double* th) {
  fmm::r2d<true, true, 1>(*nd, s, *ns, c, q, v, t, *nt, pot, nullptr, nullptr, *th, true);
}
void r2d_directcdg_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const double* c, const double* q, const double* v, const double* t, const fmm2d_int* nt, double* pot, double* grad, const double* th) {
  fmm::r2d<true, true, 2>(*nd, s, *ns, c, q, v, t, *nt, pot, grad, nullptr, *th, false);
}
void r2d_directcdh_(const fmm2d_int* nd, const double* s, const fmm2d_int* ns, const double* c, const double* q, const double* v, const double* t, const fmm2d_int* nt, double* pot, double* grad, double* hess, const double* th) {
  fmm::r2d<true, true, 3>(*nd, s, *ns, c, q, v, t, *nt, pot, grad, hess, *th, false);
}
}
