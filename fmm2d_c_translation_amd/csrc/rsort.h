// rsort.h -- stable LSD radix sort of (key, value) pairs on the device, 8 key
// bits per pass, issuing no memsets: one kernel zeroes the state of every
// pass, one counts the digits of every pass, then one kernel per pass.
//
// A pass deals the keys to tiles of kTile (4 waves x kItems x 64 lanes; wave
// w holds keys base + w * 64 kItems + 64 i + lane).  Each wave ranks its keys
// stably with no barrier -- item by item it finds the lanes that share a
// digit from eight ballots and bumps its own digit counters in LDS -- then
// the tile publishes its digit counts and takes the counts of the tiles
// before it by decoupled look-back (tile ids come from an atomic counter, so
// a tile only waits for tiles that already run).  The tile is then laid out
// in sorted order in LDS (tile digit start + waves before + rank in the
// wave) and written out in runs of equal digits, each key at digit start +
// tiles before + its place in the tile's run.
// The look-back wait is bounded: past kSpinLimit polls it sets *err and
// gives up, so a fault in the chain yields a flagged wrong order rather than
// a hang.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "engine.h"

namespace fmm {
namespace rsort {
namespace {

constexpr int kThreads = 256;  // one thread per digit in the tile epilogue
constexpr int kWaves = kThreads / 64;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
constexpr uint32_t kCount = (1u << 30) - 1;   // look-back word: count
constexpr uint32_t kAggregate = 1u << 30;     // this tile's counts only
constexpr uint32_t kInclusive = 2u << 30;     // counts of this tile and all before
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ inline uint32_t ld_relaxed(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_relaxed(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_rs_init(uint32_t* state, size_t nwords) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
       i += (size_t)gridDim.x * blockDim.x)
    state[i] = 0;
}

// digit counts of every pass: hist[p * 256 + d]
// TH threads per block: 1024 for large sorts (more waves per LDS histogram),
// 256 for small ones, whose blocks then also fit beside other streams' work
template <class K, int TH>
__global__ void __launch_bounds__(TH)
k_rs_hist(const K* __restrict__ key, int n, const int* n_dev, int begin, int end, int npass,
          uint32_t* hist) {
  __shared__ uint32_t h[8 * 256];
  if (n_dev) n = min(n, *n_dev);
  for (int i = threadIdx.x; i < npass * 256; i += TH) h[i] = 0;
  __syncthreads();
  for (long i = (long)blockIdx.x * TH + threadIdx.x; i < n; i += (long)gridDim.x * TH) {
    const K k = key[i];
    for (int p = 0; p < npass; ++p) {
      const int sh = begin + 8 * p, wd = min(8, end - sh);
      atomicAdd(&h[p * 256 + (int)((k >> sh) & (K)((1u << wd) - 1))], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < npass * 256; i += TH)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// one pass on key bits [sh, sh + wd): kin/vin -> kout/vout (V: with values)
template <class K, bool V>
__global__ void __launch_bounds__(kThreads)
k_rs_pass(const K* __restrict__ kin, K* __restrict__ kout, const int* __restrict__ vin,
          int* __restrict__ vout, int n, const int* n_dev, int sh, int wd, const uint32_t* hist,
          uint32_t* look, uint32_t* counter, uint32_t* err) {
  __shared__ uint32_t s_cnt[kWaves][256];  // per wave and digit: keys ranked so far
  __shared__ uint32_t s_gofs[256];         // per digit: output position - tile slot
  __shared__ uint32_t s_wsum[2][kWaves];
  __shared__ K s_key[kTile];
  __shared__ int s_val[V ? kTile : 1];
  __shared__ int s_tile;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) s_tile = (int)atomicAdd(counter, 1u);
#pragma unroll
  for (int w = 0; w < kWaves; ++w) s_cnt[w][tid] = 0;
  __syncthreads();
  const int tile = s_tile;
  // a count on the device: tiles past it leave at once (no tile before them
  // waits on them)
  if (n_dev) n = min(n, *n_dev);
  if ((long)tile * kTile >= n) return;
  const long base = (long)tile * kTile + (long)wv * (kItems * 64) + lane;
  const uint32_t mask = (1u << wd) - 1;
  K k[kItems];
  int v[kItems];
  uint32_t rk[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const long idx = base + i * 64;
    const bool ok = idx < n;
    k[i] = ok ? kin[idx] : (K)0;
    if (V) v[i] = ok ? vin[idx] : 0;
  }
  // rank inside the wave: its keys in index order are item-major, lane-minor
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;  // lanes below this one
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const bool ok = base + i * 64 < n;
    const int d = (int)((k[i] >> sh) & mask);
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t below = (uint32_t)__popcll(peers & lt);
    uint32_t old = 0;
    if (ok) old = s_cnt[wv][d];
    // the wave's LDS accesses complete in order, so every lane has read the
    // count before the digit's first lane stores the new one; the wave
    // barriers keep the compiler from moving the accesses across each other
    __builtin_amdgcn_wave_barrier();
    if (ok && below == 0) s_cnt[wv][d] = old + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    rk[i] = old + below;
  }
  __syncthreads();
  // digit tid: this tile's count, published at once for the tiles after it
  uint32_t cw[kWaves], cnt = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    cw[w] = s_cnt[w][tid];
    cnt += cw[w];
  }
  uint32_t* mine = look + (size_t)tile * 256 + tid;
  st_relaxed(mine, (tile == 0 ? kInclusive : kAggregate) | cnt);
  // exclusive scans over the digits of the pass's histogram (global digit
  // starts) and of this tile's counts (tile-local digit starts): wave scans,
  // then the wave totals
  const uint32_t h = hist[tid];
  uint32_t ih = h, ic = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t yh = __shfl_up(ih, o), yc = __shfl_up(ic, o);
    if (lane >= o) {
      ih += yh;
      ic += yc;
    }
  }
  if (lane == 63) {
    s_wsum[0][wv] = ih;
    s_wsum[1][wv] = ic;
  }
  // the tiles before this one, by decoupled look-back
  uint32_t before = 0;
  if (tile > 0) {
    for (int t = tile - 1; t >= 0; --t) {
      uint32_t x = ld_relaxed(look + (size_t)t * 256 + tid);
      uint32_t spins = 0;
      while (!(x & ~kCount) && spins < kSpinLimit) {
        __builtin_amdgcn_s_sleep(1);
        x = ld_relaxed(look + (size_t)t * 256 + tid);
        ++spins;
      }
      if (!(x & ~kCount)) {
        st_relaxed(err, 1u);
        break;
      }
      before += x & kCount;
      if (x & kInclusive) break;
    }
    st_relaxed(mine, kInclusive | (before + cnt));
  }
  __syncthreads();
  for (int w = 0; w < wv; ++w) {
    ih += s_wsum[0][w];
    ic += s_wsum[1][w];
  }
  const uint32_t tex = ic - cnt;  // first slot of digit tid in the tile
  uint32_t run = tex;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    s_cnt[w][tid] = run;
    run += cw[w];
  }
  // output position of tile slot j holding digit d: s_gofs[d] + j (mod 2^32)
  s_gofs[tid] = (ih - h) + before - tex;
  __syncthreads();
  // the tile in sorted order in LDS, then written out in runs per digit
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if (base + i * 64 >= n) continue;
    const int d = (int)((k[i] >> sh) & mask);
    const uint32_t j = s_cnt[wv][d] + rk[i];
    s_key[j] = k[i];
    if (V) s_val[j] = v[i];
  }
  __syncthreads();
  const int tn = (int)min((long)kTile, (long)n - (long)tile * kTile);
#pragma unroll 4
  for (int j = tid; j < tn; j += kThreads) {
    const K key = s_key[j];
    const uint32_t pos = s_gofs[(int)((key >> sh) & mask)] + (uint32_t)j;
    kout[pos] = key;
    if (V) vout[pos] = s_val[j];
  }
}

// Stable sort of n < 2^30 keys (and values, unless vin is null) by key bits
// [begin, end) of K: kin/vin -> kout/vout, which must not overlap them.
// n_dev (device memory, optional): the count actually sorted, min(n,
// *n_dev), read by the kernels; n then only sizes the launch and the
// buffers, and outputs past the count are left unwritten.
// *err (device-visible: device or pinned host memory) is set to 1 if a
// look-back wait gave up; the caller clears it and reads it after the
// stream synchronises.
template <class K>
int radix_sort(Ctx& c, Scratch& S, const K* kin, K* kout, const int* vin, int* vout, int n,
               int begin, int end, uint32_t* err, const int* n_dev = nullptr) {
  if (n <= 0) return IER_OK;
  if (begin < 0 || end <= begin || end > (int)(8 * sizeof(K)) || n >= (1 << 30) || !err ||
      (vin && !vout))
    return IER_ARGS;
  const int npass = (end - begin + 7) / 8;
  const int ntiles = (n + kTile - 1) / kTile;
  K* kb = npass > 1 ? S.get<K>(n) : nullptr;
  int* vb = (npass > 1 && vin) ? S.get<int>(n) : nullptr;
  const size_t nlook = (size_t)npass * ntiles * 256;
  const size_t nwords = (size_t)npass * 256 + nlook + npass;
  uint32_t* state = S.get<uint32_t>(nwords);
  if (S.err) return S.err;
  uint32_t* hist = state;
  uint32_t* look = state + (size_t)npass * 256;
  uint32_t* counter = look + nlook;
  uint32_t* e = err;
  k_rs_init<<<(unsigned)std::min<size_t>(1024, (nwords + 255) / 256), 256, 0, c.st>>>(state, nwords);
  if (n >= (1 << 20))
    k_rs_hist<K, 1024><<<std::min(256, (n + 4095) / 4096), 1024, 0, c.st>>>(kin, n, n_dev, begin,
                                                                          end, npass, hist);
  else
    k_rs_hist<K, 256><<<std::min(256, (n + 1023) / 1024), 256, 0, c.st>>>(kin, n, n_dev, begin,
                                                                        end, npass, hist);
  const K* ks = kin;
  const int* vs = vin;
  for (int p = 0; p < npass; ++p) {
    const bool last = ((npass - 1 - p) & 1) == 0;  // passes alternate, the last one into kout
    K* kd = last ? kout : kb;
    int* vd = last ? vout : vb;
    const int sh = begin + 8 * p, wd = std::min(8, end - sh);
    if (vin)
      k_rs_pass<K, true><<<ntiles, kThreads, 0, c.st>>>(ks, kd, vs, vd, n, n_dev, sh, wd, hist + 256 * p,
                                                        look + (size_t)p * ntiles * 256,
                                                        counter + p, e);
    else
      k_rs_pass<K, false><<<ntiles, kThreads, 0, c.st>>>(ks, kd, nullptr, nullptr, n, n_dev, sh, wd,
                                                         hist + 256 * p,
                                                         look + (size_t)p * ntiles * 256,
                                                         counter + p, e);
    ks = kd;
    vs = vd;
  }
  return hipGetLastError() == hipSuccess ? IER_OK : IER_HIP;
}

// After a sort whose values are used as indices before the host can read
// *err: when a wait gave up, the unsorted values (vin is left intact) replace
// the output, so consumers still see valid indices in a valid (unsorted)
// order and the failure is reported when the host reads *err.
// One block: the error word (pinned host memory) is read once, by one
// thread, and the copy only runs when it is set.
__global__ void __launch_bounds__(256) k_rs_fallback(const uint32_t* err, const int* vin,
                                                     int* vout, int n) {
  __shared__ uint32_t flag;
  if (threadIdx.x == 0) flag = ld_relaxed(err);
  __syncthreads();
  if (!flag) return;
  for (int i = threadIdx.x; i < n; i += blockDim.x) vout[i] = vin[i];
}
inline void fallback_unsorted(hipStream_t st, const uint32_t* err, const int* vin, int* vout,
                              int n) {
  if (n > 0) k_rs_fallback<<<1, 256, 0, st>>>(err, vin, vout, n);
}

}  // namespace
}  // namespace rsort
}  // namespace fmm
