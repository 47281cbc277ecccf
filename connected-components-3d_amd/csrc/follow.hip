// Follow-ons of the labelling path that stay in HBM: per-label statistics
// (cc3d/fastcc3d.pyx:681-938) and dust removal (cc3d/__init__.py:71-155).
//
// statistics: voxel counts (uint32, wrapping like the reference's uint32
// array), bounding boxes along the numpy axes (min/max per axis) and centroids
// (sum of coordinates / count in float64; the sums are exact integers, so the
// result is independent of summation order).  A workgroup sums a brick of
// the volume per label in an LDS table and then flushes one set of global
// atomics per label present in the brick.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/cc3d_amd.h"
#include "common.hpp"

namespace cc3d_amd {

template <class L>
__global__ __launch_bounds__(WG) void k_label_range(const L* lab, int64_t n, int is_signed,
                                                    unsigned long long* umax, long long* smin,
                                                    long long* smax) {
  unsigned long long mx = 0;
  long long lo = 0x7FFFFFFFFFFFFFFFll, hi = -0x7FFFFFFFFFFFFFFFll - 1;
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) {
    const L v = lab[i];
    if (is_signed) {
      long long s;
      if (sizeof(L) == 1) s = (int8_t)v;
      else if (sizeof(L) == 2) s = (int16_t)v;
      else if (sizeof(L) == 4) s = (int32_t)v;
      else s = (long long)v;
      lo = s < lo ? s : lo;
      hi = s > hi ? s : hi;
    } else {
      mx = (unsigned long long)v > mx ? (unsigned long long)v : mx;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long a = __shfl_xor(mx, o);
    const long long b = __shfl_xor(lo, o), c = __shfl_xor(hi, o);
    mx = a > mx ? a : mx;
    lo = b < lo ? b : lo;
    hi = c > hi ? c : hi;
  }
  // one set of atomics per block: each lands on a single address
  __shared__ unsigned long long wmx[WG / 64];
  __shared__ long long wlo[WG / 64], whi[WG / 64];
  if ((threadIdx.x & 63) == 0) {
    wmx[threadIdx.x >> 6] = mx;
    wlo[threadIdx.x >> 6] = lo;
    whi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < WG / 64; w++) {
      mx = wmx[w] > mx ? wmx[w] : mx;
      lo = wlo[w] < lo ? wlo[w] : lo;
      hi = whi[w] > hi ? whi[w] : hi;
    }
    atomicMax(umax, mx);
    atomicMin(smin, lo);
    atomicMax(smax, hi);
  }
}

__global__ void k_range_init(unsigned long long* umax, long long* smin, long long* smax) {
  *umax = 0;
  *smin = 0x7FFFFFFFFFFFFFFFll;
  *smax = -0x7FFFFFFFFFFFFFFFll - 1;
}

__global__ void k_stats_init(uint32_t* counts, uint32_t* bbox, unsigned long long* sums, int64_t n1, int nd) {
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n1; i += (int64_t)gridDim.x * WG) {
    counts[i] = 0;
    for (int a = 0; a < nd; a++) {
      if (bbox) {
        bbox[2 * nd * i + 2 * a] = 0xFFFFFFFFu;
        bbox[2 * nd * i + 2 * a + 1] = 0u;
      }
      if (sums) sums[nd * i + a] = 0ull;
    }
  }
}

// Bricks of the label volume in memory order: m0 is the memory-fastest
// dimension; p[k] is the numpy axis of memory dimension k (>= nd: a unit axis
// that is not reported).  A brick is (16 << sl) x (1 << yl) x (1 << zl)
// voxels, sl + yl + zl = 8: each of the 256 threads walks one 16-voxel
// segment of one row.
struct StatGeo {
  int64_t m0, m1, m2;
  int64_t nbx, nby, nbricks;
  int p[3];
  int nd, sl, yl, zl;
};

constexpr int ST_SEG = 16;
constexpr int ST_SLOTS = 512;  // per-brick table; slot 0 is label 0's
constexpr int ST_PROBES = 16;  // then the run goes straight to global atomics

template <class L>
__device__ inline void load_seg(const L* p, L (&v)[ST_SEG]) {
  if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
    constexpr int NV = ST_SEG * sizeof(L) / 16;
    uint4 w[NV];
    const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int j = 0; j < NV; j++) w[j] = q[j];
    __builtin_memcpy(v, w, sizeof(v));
  } else {
#pragma unroll
    for (int i = 0; i < ST_SEG; i++) v[i] = p[i];
  }
}

// Per-label statistics, W: bit 0 bounding boxes, bit 1 coordinate sums,
// bit 2 counts are final already (a counting pass ran before): no count
// atomics, and a label whose count in the brick equals its total lies in
// this brick only, so its boxes and sums are written with plain stores.
// Each thread splits its segment into runs of one label; each run is added
// to the brick's LDS table (count, min/max and coordinate sums relative to
// the brick origin; along a row only the fastest coordinate varies), and
// after the brick each occupied slot costs one set of global atomics.
// Label 0, present in most bricks of most volumes, is summed in registers
// over every brick of the workgroup and flushed once.  A run whose label
// finds no slot within ST_PROBES probes is flushed to global memory
// directly.  Counts wrap modulo 2^32 like the reference's uint32 array.
template <class L, int W>
__global__ __launch_bounds__(WG) void k_stats(const L* lab, StatGeo g, uint32_t* counts, uint32_t* bbox,
                                              unsigned long long* sums) {
  constexpr bool BB = (W & 1) != 0, SM = (W & 2) != 0, CF = (W & 4) != 0;
  constexpr unsigned long long EMPTY = ~0ull;  // never a label: labels are at most N < 2^64 - 1
  __shared__ unsigned long long tkey[ST_SLOTS];
  __shared__ uint32_t tcnt[ST_SLOTS];
  __shared__ uint32_t tmn[BB ? 3 * ST_SLOTS : 1], tmx[BB ? 3 * ST_SLOTS : 1];
  __shared__ uint32_t tsm[SM ? 3 * ST_SLOTS : 1];
  const int t = threadIdx.x, nd = g.nd;
  auto reset = [&](int s) {
    tkey[s] = EMPTY;
    tcnt[s] = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if constexpr (BB) {
        tmn[3 * s + k] = 0xFFFFFFFFu;
        tmx[3 * s + k] = 0u;
      }
      if constexpr (SM) tsm[3 * s + k] = 0u;
    }
  };
  for (int s = t; s < ST_SLOTS; s += WG) reset(s);
  const int seg = t & ((1 << g.sl) - 1), row = t >> g.sl;
  const int lx = seg * ST_SEG, ly = row & ((1 << g.yl) - 1), lz = row >> g.yl;
  uint32_t c0 = 0;
  bool seen0 = false;
  int64_t mn0[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, mx0[3] = {-1, -1, -1};
  unsigned long long sm0[3] = {0, 0, 0};
  __syncthreads();
  for (int64_t b = blockIdx.x; b < g.nbricks; b += gridDim.x) {
    const int64_t bx = b % g.nbx, r = b / g.nbx, by = r % g.nby, bz = r / g.nby;
    const int64_t o[3] = {bx << (g.sl + 4), by << g.yl, bz << g.zl};
    const int64_t x = o[0] + lx, y = o[1] + ly, z = o[2] + lz;
    if (x < g.m0 && y < g.m1 && z < g.m2) {
      const L* p = lab + (x + g.m0 * (y + g.m1 * z));
      const int len = (int)(g.m0 - x < ST_SEG ? g.m0 - x : ST_SEG);
      // bit i (1 <= i < len): a run starts at i; bit len ends the last run
      uint32_t m = 1u << len;
      if (len == ST_SEG) {
        L v[ST_SEG];
        load_seg(p, v);
#pragma unroll
        for (int i = 1; i < ST_SEG; i++) m |= (uint32_t)(v[i] != v[i - 1]) << i;
      } else {
        L prev = p[0];
        for (int i = 1; i < len; i++) {
          const L v = p[i];
          m |= (uint32_t)(v != prev) << i;
          prev = v;
        }
      }
      for (int xs = 0; xs < len;) {
        const int xe = __builtin_ctz(m & ~((2u << xs) - 1u));
        const unsigned long long key = (unsigned long long)p[xs];
        const uint32_t c = (uint32_t)(xe - xs);
        const int xa = lx + xs, xb = lx + xe - 1;
        int s = 0;
        if (key != 0) {
          s = -1;
          uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 55);
          for (int q = 0; q < ST_PROBES; q++) {
            h = h ? h : 1u;
            unsigned long long k = *reinterpret_cast<volatile unsigned long long*>(tkey + h);
            if (k == EMPTY) k = atomicCAS(tkey + h, EMPTY, key);
            if (k == EMPTY || k == key) {
              s = (int)h;
              break;
            }
            h = (h + 1) & (ST_SLOTS - 1);
          }
        }
        const uint32_t lo[3] = {(uint32_t)xa, (uint32_t)ly, (uint32_t)lz};
        const uint32_t hi[3] = {(uint32_t)xb, (uint32_t)ly, (uint32_t)lz};
        const uint32_t ls[3] = {(uint32_t)(xa + xb) * c / 2u, c * (uint32_t)ly, c * (uint32_t)lz};
        if (s >= 0) {
          atomicAdd(tcnt + s, c);
#pragma unroll
          for (int k = 0; k < 3; k++) {
            if constexpr (BB) {
              atomicMin(tmn + 3 * s + k, lo[k]);
              atomicMax(tmx + 3 * s + k, hi[k]);
            }
            if constexpr (SM) atomicAdd(tsm + 3 * s + k, ls[k]);
          }
        } else {
          if constexpr (!CF) atomicAdd(counts + key, c);
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const int a = g.p[k];
            if (a >= nd) continue;
            if constexpr (BB) {
              atomicMin(bbox + 2 * nd * key + 2 * a, (uint32_t)(o[k] + lo[k]));
              atomicMax(bbox + 2 * nd * key + 2 * a + 1, (uint32_t)(o[k] + hi[k]));
            }
            if constexpr (SM) atomicAdd(sums + nd * key + a, (unsigned long long)c * o[k] + ls[k]);
          }
        }
        xs = xe;
      }
    }
    __syncthreads();
    for (int s = t; s < ST_SLOTS; s += WG) {
      const uint32_t c = tcnt[s];
      if (c == 0) continue;  // never claimed: a claimed slot always has a run added
      if (s == 0) {
        c0 += c;
        seen0 = true;
#pragma unroll
        for (int k = 0; k < 3; k++) {
          if constexpr (BB) {
            const int64_t a = o[k] + tmn[k], z2 = o[k] + tmx[k];
            mn0[k] = a < mn0[k] ? a : mn0[k];
            mx0[k] = z2 > mx0[k] ? z2 : mx0[k];
          }
          if constexpr (SM) sm0[k] += (unsigned long long)c * o[k] + tsm[k];
        }
      } else {
        const unsigned long long key = tkey[s];
        if constexpr (!CF) atomicAdd(counts + key, c);
        const bool only = CF && counts[key] == c;  // the label's every voxel is in this slot
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const int a = g.p[k];
          if (a >= nd) continue;
          const uint32_t lo = (uint32_t)(o[k] + (BB ? tmn[3 * s + k] : 0u));
          const uint32_t hi = (uint32_t)(o[k] + (BB ? tmx[3 * s + k] : 0u));
          const unsigned long long sm = (unsigned long long)c * o[k] + (SM ? tsm[3 * s + k] : 0u);
          if (only) {
            if constexpr (BB) {
              bbox[2 * nd * key + 2 * a] = lo;
              bbox[2 * nd * key + 2 * a + 1] = hi;
            }
            if constexpr (SM) sums[nd * key + a] = sm;
          } else {
            if constexpr (BB) {
              atomicMin(bbox + 2 * nd * key + 2 * a, lo);
              atomicMax(bbox + 2 * nd * key + 2 * a + 1, hi);
            }
            if constexpr (SM) atomicAdd(sums + nd * key + a, sm);
          }
        }
      }
      reset(s);
    }
    __syncthreads();
  }
  if (t == 0 && seen0) {
    if constexpr (!CF) atomicAdd(counts, c0);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int a = g.p[k];
      if (a >= nd) continue;
      if constexpr (BB) {
        atomicMin(bbox + 2 * a, (uint32_t)mn0[k]);
        atomicMax(bbox + 2 * a + 1, (uint32_t)mx0[k]);
      }
      if constexpr (SM) atomicAdd(sums + a, sm0[k]);
    }
  }
}

__global__ void k_centroids(const uint32_t* counts, const unsigned long long* sums, double* cen, int64_t n1,
                            int nd) {
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n1; i += (int64_t)gridDim.x * WG) {
    const uint32_t c = counts[i];
    for (int a = 0; a < nd; a++)
      cen[nd * i + a] = c == 0 ? __builtin_nan("") : (double)sums[nd * i + a] / (double)c;
  }
}

// dust: erase[l] = 1 when voxels of label l are to be zeroed.
__global__ void k_dust_table(const uint32_t* counts, int64_t N, double lo, double hi, int ranged, int invert,
                             uint8_t* erase, unsigned long long* nmask) {
  unsigned long long m = 0;
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i <= N; i += (int64_t)gridDim.x * WG) {
    bool mask = false;
    if (i >= 1) {
      const double sz = (double)counts[i];
      mask = ranged ? !(lo <= sz && sz < hi) : (sz < lo);
    }
    m += mask;
    erase[i] = (uint8_t)(mask != (invert != 0));
  }
  // one atomic per block: the count lands on a single address
  __shared__ unsigned long long wsum[WG / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m += __shfl_xor(m, o);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = 0;
    for (int w = 0; w < WG / 64; w++) b += wsum[w];
    if (b) atomicAdd(nmask, b);
  }
}

template <class T, class L>
__global__ __launch_bounds__(WG) void k_dust_apply(T* img, const L* lab, const uint8_t* erase, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG)
    if (erase[(uint64_t)lab[i]]) img[i] = T(0);
}

// largest_k renumbering: lut[l] = output label of label l (0: dropped).
__global__ void k_lut_clear(uint32_t* lut, int64_t n1) {
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n1; i += (int64_t)gridDim.x * WG) lut[i] = 0u;
}

__global__ void k_lut_set(uint32_t* lut, const uint64_t* keep, int64_t k, uint64_t N) {
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < k; i += (int64_t)gridDim.x * WG) {
    const uint64_t l = keep[i];
    if (l <= N) lut[l] = (uint32_t)(i + 1);
  }
}

template <class L, class O>
__global__ __launch_bounds__(WG) void k_keep_apply(O* out, const L* lab, const uint32_t* lut, uint64_t N, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) {
    const uint64_t l = (uint64_t)lab[i];
    out[i] = O(l <= N ? lut[l] : 0u);
  }
}

template <class T, class L>
__global__ __launch_bounds__(WG) void k_dust_copy(T* __restrict__ dst, const T* __restrict__ src, const L* lab,
                                                 const uint8_t* erase, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG)
    dst[i] = erase[(uint64_t)lab[i]] ? T(0) : src[i];
}

static thread_local std::string f_err;

static int ffail(int code, const char* msg) {
  f_err = msg;
  return code;
}

static inline unsigned fgrid(int64_t n, int64_t per, int64_t cap) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + per - 1) / per, cap));
}

// One workgroup per brick, at most as many as are resident at once: each
// workgroup then walks its bricks in a grid-stride loop.
template <class L, int W>
static void launch_stats_w(const void* labels, const StatGeo& g, uint32_t* counts, uint32_t* bbox,
                           unsigned long long* sums, hipStream_t s) {
  int dev = 0, cus = 0, occ = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(k_stats<L, W>), WG, 0);
  const int64_t cap = (int64_t)std::max(cus, 1) * std::max(occ, 1);
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(g.nbricks, cap));
  hipLaunchKernelGGL((k_stats<L, W>), dim3(grid), dim3(WG), 0, s, static_cast<const L*>(labels), g, counts, bbox,
                     sums);
}

// With boxes or sums and fewer than 2^32 voxels (no count can wrap), counts
// come first in a pass of their own, and the second pass writes the labels
// that lie in a single brick (most labels of a noisy volume) without atomics.
template <class L>
static void launch_stats(int w, int64_t n, const void* labels, const StatGeo& g, uint32_t* counts, uint32_t* bbox,
                         unsigned long long* sums, hipStream_t s) {
  if (w != 0 && n < (1ll << 32)) {
    launch_stats_w<L, 0>(labels, g, counts, bbox, sums, s);
    w |= 4;
  }
  switch (w) {
    case 0: launch_stats_w<L, 0>(labels, g, counts, bbox, sums, s); break;
    case 1: launch_stats_w<L, 1>(labels, g, counts, bbox, sums, s); break;
    case 2: launch_stats_w<L, 2>(labels, g, counts, bbox, sums, s); break;
    case 3: launch_stats_w<L, 3>(labels, g, counts, bbox, sums, s); break;
    case 5: launch_stats_w<L, 5>(labels, g, counts, bbox, sums, s); break;
    case 6: launch_stats_w<L, 6>(labels, g, counts, bbox, sums, s); break;
    default: launch_stats_w<L, 7>(labels, g, counts, bbox, sums, s); break;
  }
}

}  // namespace cc3d_amd

using namespace cc3d_amd;

extern "C" {

const char* cc3d_amd_follow_last_error(void) { return f_err.c_str(); }

int cc3d_amd_label_range(const void* labels, int dtype, int64_t voxels, int is_signed, uint64_t* umax,
                         int64_t* smin, int64_t* smax, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned long long* d = nullptr;
  if (hipMallocAsync(reinterpret_cast<void**>(&d), 64, s) != hipSuccess) return ffail(CC3D_AMD_ERR_NOMEM, "alloc");
  long long* smn = reinterpret_cast<long long*>(d + 1);
  long long* smx = reinterpret_cast<long long*>(d + 2);
  hipLaunchKernelGGL(k_range_init, dim3(1), dim3(1), 0, s, d, smn, smx);
  const dim3 g(fgrid(voxels, WG, 8192));
  switch (dtype) {
    case DT_U8: hipLaunchKernelGGL((k_label_range<uint8_t>), g, dim3(WG), 0, s, (const uint8_t*)labels, voxels, is_signed, d, smn, smx); break;
    case DT_U16: hipLaunchKernelGGL((k_label_range<uint16_t>), g, dim3(WG), 0, s, (const uint16_t*)labels, voxels, is_signed, d, smn, smx); break;
    case DT_U32: hipLaunchKernelGGL((k_label_range<uint32_t>), g, dim3(WG), 0, s, (const uint32_t*)labels, voxels, is_signed, d, smn, smx); break;
    case DT_U64: hipLaunchKernelGGL((k_label_range<uint64_t>), g, dim3(WG), 0, s, (const uint64_t*)labels, voxels, is_signed, d, smn, smx); break;
    default: (void)hipFreeAsync(d, s); return ffail(CC3D_AMD_ERR_INVALID, "labels must be an integer type");
  }
  unsigned long long h[3];
  if (hipMemcpyAsync(h, d, 24, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return ffail(CC3D_AMD_ERR_HIP, "label range readback failed");
  (void)hipFreeAsync(d, s);
  *umax = h[0];
  *smin = (int64_t)h[1];
  *smax = (int64_t)h[2];
  return CC3D_AMD_OK;
}

// counts: uint32[N+1]; bbox: uint32[(N+1)*2*nd] as min0,max0,min1,max1[,min2,max2]
// (min = 0xFFFFFFFF, max = 0 for absent labels); centroids: float64[(N+1)*nd];
// sums_scratch: uint64[(N+1)*nd].  bbox and centroids may be null (not
// computed); centroids need sums_scratch.
int cc3d_amd_statistics(const void* labels, int dtype, int64_t s0, int64_t s1, int64_t s2, int ndim, int f_order,
                        uint64_t N, uint32_t* counts, uint32_t* bbox, double* centroids,
                        unsigned long long* sums_scratch, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!counts || (centroids && !sums_scratch)) return ffail(CC3D_AMD_ERR_INVALID, "statistics needs counts (and sums_scratch for centroids)");
  if (dtype != DT_U8 && dtype != DT_U16 && dtype != DT_U32 && dtype != DT_U64)
    return ffail(CC3D_AMD_ERR_INVALID, "labels must be an integer type");
  const int64_t n1 = (int64_t)N + 1, n = s0 * s1 * s2;
  const int nd = ndim == 2 ? 2 : 3;
  unsigned long long* sums = centroids ? sums_scratch : nullptr;
  hipLaunchKernelGGL(k_stats_init, dim3(fgrid(n1, WG, 8192)), dim3(WG), 0, s, counts, bbox, sums, n1, nd);
  if (n > 0) {
    StatGeo g;
    int64_t m[3];
    if (f_order) {
      m[0] = s0, m[1] = s1, m[2] = s2;
      g.p[0] = 0, g.p[1] = 1, g.p[2] = 2;
    } else {
      m[0] = s2, m[1] = s1, m[2] = s0;
      g.p[0] = 2, g.p[1] = 1, g.p[2] = 0;
    }
    for (int r = 0; r < 2 && m[0] == 1; r++) {  // a unit fastest dimension: the next one is contiguous
      const int64_t m0 = m[0];
      const int p0 = g.p[0];
      m[0] = m[1], m[1] = m[2], m[2] = m0;
      g.p[0] = g.p[1], g.p[1] = g.p[2], g.p[2] = p0;
    }
    g.m0 = m[0], g.m1 = m[1], g.m2 = m[2];
    g.nd = nd;
    g.sl = g.m0 >= 48 ? 2 : g.m0 >= 24 ? 1 : 0;
    while (g.sl < 8 && g.m1 * g.m2 < (1ll << (8 - g.sl)) && (16ll << g.sl) < g.m0) g.sl++;  // few long rows
    const int rl = 8 - g.sl;
    if (g.m2 == 1) {
      g.yl = rl;
    } else {
      g.yl = (rl + 1) / 2;
      while (g.yl > 0 && (1ll << (g.yl - 1)) >= g.m1) g.yl--;
    }
    g.zl = rl - g.yl;
    g.nbx = (g.m0 + (16ll << g.sl) - 1) >> (g.sl + 4);
    g.nby = (g.m1 + (1ll << g.yl) - 1) >> g.yl;
    g.nbricks = g.nbx * g.nby * ((g.m2 + (1ll << g.zl) - 1) >> g.zl);
    const int w = (bbox ? 1 : 0) | (sums ? 2 : 0);
    switch (dtype) {
      case DT_U8: launch_stats<uint8_t>(w, n, labels, g, counts, bbox, sums, s); break;
      case DT_U16: launch_stats<uint16_t>(w, n, labels, g, counts, bbox, sums, s); break;
      case DT_U32: launch_stats<uint32_t>(w, n, labels, g, counts, bbox, sums, s); break;
      default: launch_stats<uint64_t>(w, n, labels, g, counts, bbox, sums, s); break;
    }
  }
  if (centroids)
    hipLaunchKernelGGL(k_centroids, dim3(fgrid(n1, WG, 8192)), dim3(WG), 0, s, counts, sums, centroids, n1, nd);
  if (hipGetLastError() != hipSuccess) return ffail(CC3D_AMD_ERR_HIP, "statistics launch failed");
  return CC3D_AMD_OK;
}

// Zeroes img where the component label's size fails the threshold test of
// cc3d/__init__.py:121-150.  erase_scratch: uint8[N+1].  *n_masked (device
// pointer, uint64) receives len(to_mask).
int cc3d_amd_dust_apply(void* img, int img_dtype, const void* labels, int label_dtype, int64_t voxels,
                        const uint32_t* counts, uint64_t N, double lo, double hi, int ranged, int invert,
                        uint8_t* erase_scratch, unsigned long long* n_masked, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_dust_table, dim3(fgrid((int64_t)N + 1, WG, 8192)), dim3(WG), 0, s, counts, (int64_t)N, lo,
                     hi, ranged, invert, erase_scratch, n_masked);
  const dim3 g(fgrid(voxels, WG, 65536));
#define CC3D_DA(T, L) hipLaunchKernelGGL((k_dust_apply<T, L>), g, dim3(WG), 0, s, (T*)img, (const L*)labels, erase_scratch, voxels)
#define CC3D_DL(T)                                                  \
  switch (label_dtype) {                                            \
    case DT_U8: CC3D_DA(T, uint8_t); break;                         \
    case DT_U16: CC3D_DA(T, uint16_t); break;                       \
    case DT_U32: CC3D_DA(T, uint32_t); break;                       \
    case DT_U64: CC3D_DA(T, uint64_t); break;                       \
    default: return ffail(CC3D_AMD_ERR_INVALID, "labels must be unsigned integers"); \
  }
  switch (img_dtype) {
    case DT_U8: CC3D_DL(uint8_t); break;
    case DT_U16: CC3D_DL(uint16_t); break;
    case DT_U32: CC3D_DL(uint32_t); break;
    case DT_U64: CC3D_DL(uint64_t); break;
    case DT_F32: CC3D_DL(uint32_t); break;  // zeroing is bitwise: +0.0 has all bits clear
    case DT_F64: CC3D_DL(uint64_t); break;
    default: return ffail(CC3D_AMD_ERR_INVALID, "bad image dtype");
  }
#undef CC3D_DL
#undef CC3D_DA
  if (hipGetLastError() != hipSuccess) return ffail(CC3D_AMD_ERR_HIP, "dust launch failed");
  return CC3D_AMD_OK;
}

// largest_k's renumbering without fastremap (cc3d/__init__.py:266-275):
// out[v] = i + 1 where labels[v] == keep[i], else 0.  keep: device uint64[k];
// lut_scratch: device uint32[N+1]; labels are <= N.  out_dtype DT_U8 with
// k == 1 gives the boolean mask of the k == 1 branch (:250-254).
int cc3d_amd_keep_labels(const void* labels, int label_dtype, int64_t voxels, const uint64_t* keep, int64_t k,
                         uint64_t N, uint32_t* lut_scratch, void* out, int out_dtype, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!labels || !out || !lut_scratch || (k > 0 && !keep) || k < 0 || voxels < 0)
    return ffail(CC3D_AMD_ERR_INVALID, "keep_labels: bad argument");
  if (k > 0xFFFFFFFFll) return ffail(CC3D_AMD_ERR_INVALID, "keep_labels: k must fit 32 bits");
  const auto unsigned_code = [](int d) { return d == DT_U8 || d == DT_U16 || d == DT_U32 || d == DT_U64; };
  if (!unsigned_code(label_dtype)) return ffail(CC3D_AMD_ERR_INVALID, "keep_labels: labels must be unsigned integers");
  if (!unsigned_code(out_dtype)) return ffail(CC3D_AMD_ERR_INVALID, "keep_labels: output must be u8/u16/u32/u64");
  const int64_t n1 = (int64_t)N + 1;
  hipLaunchKernelGGL(k_lut_clear, dim3(fgrid(n1, WG, 8192)), dim3(WG), 0, s, lut_scratch, n1);
  if (k > 0) hipLaunchKernelGGL(k_lut_set, dim3(fgrid(k, WG, 8192)), dim3(WG), 0, s, lut_scratch, keep, k, N);
  const dim3 g(fgrid(voxels, WG, 65536));
#define CC3D_KA(L, O) hipLaunchKernelGGL((k_keep_apply<L, O>), g, dim3(WG), 0, s, (O*)out, (const L*)labels, lut_scratch, N, voxels)
#define CC3D_KO(L)                                                   \
  switch (out_dtype) {                                               \
    case DT_U8: CC3D_KA(L, uint8_t); break;                          \
    case DT_U16: CC3D_KA(L, uint16_t); break;                        \
    case DT_U32: CC3D_KA(L, uint32_t); break;                        \
    case DT_U64: CC3D_KA(L, uint64_t); break;                        \
    default: return ffail(CC3D_AMD_ERR_INVALID, "keep_labels: output must be u8/u16/u32/u64"); \
  }
  if (voxels > 0) {
    switch (label_dtype) {
      case DT_U8: CC3D_KO(uint8_t); break;
      case DT_U16: CC3D_KO(uint16_t); break;
      case DT_U32: CC3D_KO(uint32_t); break;
      case DT_U64: CC3D_KO(uint64_t); break;
      default: return ffail(CC3D_AMD_ERR_INVALID, "keep_labels: labels must be unsigned integers");
    }
  }
#undef CC3D_KO
#undef CC3D_KA
  if (hipGetLastError() != hipSuccess) return ffail(CC3D_AMD_ERR_HIP, "keep_labels launch failed");
  return CC3D_AMD_OK;
}

// dust into a separate image: dst[v] = 0 where the component of v fails the
// test of cc3d_amd_dust_apply, else src[v] (the copy and the masking of
// cc3d/__init__.py:121-150 in one pass).  Same arguments otherwise.
int cc3d_amd_dust_copy(const void* src, void* dst, int img_dtype, const void* labels, int label_dtype,
                       int64_t voxels, const uint32_t* counts, uint64_t N, double lo, double hi, int ranged,
                       int invert, uint8_t* erase_scratch, unsigned long long* n_masked, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!src || !dst || !labels || !counts || !erase_scratch || !n_masked || voxels < 0)
    return ffail(CC3D_AMD_ERR_INVALID, "dust_copy: bad argument");
  const auto unsigned_code = [](int d) { return d == DT_U8 || d == DT_U16 || d == DT_U32 || d == DT_U64; };
  if (!unsigned_code(label_dtype)) return ffail(CC3D_AMD_ERR_INVALID, "labels must be unsigned integers");
  if (!unsigned_code(img_dtype) && img_dtype != DT_F32 && img_dtype != DT_F64)
    return ffail(CC3D_AMD_ERR_INVALID, "bad image dtype");
  hipLaunchKernelGGL(k_dust_table, dim3(fgrid((int64_t)N + 1, WG, 8192)), dim3(WG), 0, s, counts, (int64_t)N, lo,
                     hi, ranged, invert, erase_scratch, n_masked);
  const dim3 g(fgrid(voxels, WG, 65536));
#define CC3D_DC(T, L) \
  hipLaunchKernelGGL((k_dust_copy<T, L>), g, dim3(WG), 0, s, (T*)dst, (const T*)src, (const L*)labels, erase_scratch, voxels)
#define CC3D_DCL(T)                                 \
  switch (label_dtype) {                            \
    case DT_U8: CC3D_DC(T, uint8_t); break;         \
    case DT_U16: CC3D_DC(T, uint16_t); break;       \
    case DT_U32: CC3D_DC(T, uint32_t); break;       \
    default: CC3D_DC(T, uint64_t); break;           \
  }
  switch (img_dtype) {  // zeroing is bitwise: +0.0 has all bits clear
    case DT_U8: CC3D_DCL(uint8_t); break;
    case DT_U16: CC3D_DCL(uint16_t); break;
    case DT_U32: case DT_F32: CC3D_DCL(uint32_t); break;
    default: CC3D_DCL(uint64_t); break;
  }
#undef CC3D_DCL
#undef CC3D_DC
  if (hipGetLastError() != hipSuccess) return ffail(CC3D_AMD_ERR_HIP, "dust launch failed");
  return CC3D_AMD_OK;
}

}  // extern "C"
