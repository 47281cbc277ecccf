// Follow-ons of the labelling path that stay in HBM: per-label statistics
// (cc3d/fastcc3d.pyx:681-938) and dust removal (cc3d/__init__.py:71-155).
//
// statistics: voxel counts (uint32, wrapping like the reference's uint32
// array), bounding boxes along the numpy axes (min/max per axis) and centroids
// (sum of coordinates / count in float64; the sums are exact integers, so the
// result is independent of summation order).  Each thread walks a short
// contiguous run of voxels and flushes its atomics only when the label
// changes, so long runs of one label cost one set of atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/cc3d_amd.h"
#include "common.hpp"

namespace cc3d_amd {

constexpr int STAT_RUN = 16;

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
  if ((threadIdx.x & 63) == 0) {
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
      bbox[2 * nd * i + 2 * a] = 0xFFFFFFFFu;
      bbox[2 * nd * i + 2 * a + 1] = 0u;
      sums[nd * i + a] = 0ull;
    }
  }
}

// Coordinates along numpy axes (s0, s1, s2) of linear index i: F order has
// axis 0 fastest, C order axis 2 (or 1 in 2D) fastest.
template <class L>
__global__ __launch_bounds__(WG) void k_stats(const L* lab, int64_t s0, int64_t s1, int64_t s2, int nd,
                                              int f_order, uint32_t* counts, uint32_t* bbox,
                                              unsigned long long* sums) {
  const int64_t n = s0 * s1 * s2;
  const int64_t start = ((int64_t)blockIdx.x * WG + threadIdx.x) * STAT_RUN;
  auto coords = [&](int64_t i, int64_t& c0, int64_t& c1, int64_t& c2) {
    if (f_order) {
      c0 = i % s0;
      const int64_t r = i / s0;
      c1 = r % s1;
      c2 = r / s1;
    } else {
      c2 = i % s2;
      const int64_t r = i / s2;
      c1 = r % s1;
      c0 = r / s1;
    }
  };
  uint64_t cur = start < n ? (uint64_t)lab[start] : 0;
  uint32_t cnt = 0;
  int64_t mn[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, mx[3] = {-1, -1, -1};
  unsigned long long sm[3] = {0, 0, 0};
  // Label 0 (background) is in most runs of most volumes: it is accumulated
  // per thread and reduced per block, so it costs one set of atomics per
  // block instead of one per run on a single hot address.
  uint32_t cnt0 = 0;
  int64_t mn0[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, mx0[3] = {-1, -1, -1};
  unsigned long long sm0[3] = {0, 0, 0};
  auto flush = [&]() {
    if (cur == 0) {
      cnt0 += cnt;
      for (int a = 0; a < 3; a++) {
        mn0[a] = mn[a] < mn0[a] ? mn[a] : mn0[a];
        mx0[a] = mx[a] > mx0[a] ? mx[a] : mx0[a];
        sm0[a] += sm[a];
      }
      return;
    }
    atomicAdd(counts + cur, cnt);
    for (int a = 0; a < nd; a++) {
      atomicMin(bbox + 2 * nd * cur + 2 * a, (uint32_t)mn[a]);
      atomicMax(bbox + 2 * nd * cur + 2 * a + 1, (uint32_t)mx[a]);
      atomicAdd(sums + nd * cur + a, sm[a]);
    }
  };
  if (start < n) {
    const int64_t end = start + STAT_RUN < n ? start + STAT_RUN : n;
    for (int64_t i = start; i < end; i++) {
      const uint64_t l = (uint64_t)lab[i];
      if (l != cur) {
        flush();
        cur = l;
        cnt = 0;
        for (int a = 0; a < 3; a++) {
          mn[a] = INT64_MAX;
          mx[a] = -1;
          sm[a] = 0;
        }
      }
      int64_t c[3];
      coords(i, c[0], c[1], c[2]);
      cnt++;
      for (int a = 0; a < nd; a++) {
        mn[a] = c[a] < mn[a] ? c[a] : mn[a];
        mx[a] = c[a] > mx[a] ? c[a] : mx[a];
        sm[a] += (unsigned long long)c[a];
      }
    }
    flush();
  }
  // block reduction of the background accumulators
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    cnt0 += __shfl_xor(cnt0, o);
    for (int a = 0; a < 3; a++) {
      const int64_t m1 = __shfl_xor(mn0[a], o), m2 = __shfl_xor(mx0[a], o);
      mn0[a] = m1 < mn0[a] ? m1 : mn0[a];
      mx0[a] = m2 > mx0[a] ? m2 : mx0[a];
      sm0[a] += __shfl_xor(sm0[a], o);
    }
  }
  __shared__ uint32_t bc[WG / 64];
  __shared__ long long bmn[WG / 64][3], bmx[WG / 64][3];
  __shared__ unsigned long long bsm[WG / 64][3];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    bc[wv] = cnt0;
    for (int a = 0; a < 3; a++) {
      bmn[wv][a] = mn0[a];
      bmx[wv][a] = mx0[a];
      bsm[wv][a] = sm0[a];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < WG / 64; w++) {
      cnt0 += bc[w];
      for (int a = 0; a < 3; a++) {
        mn0[a] = bmn[w][a] < mn0[a] ? bmn[w][a] : mn0[a];
        mx0[a] = bmx[w][a] > mx0[a] ? bmx[w][a] : mx0[a];
        sm0[a] += bsm[w][a];
      }
    }
    if (cnt0) {
      atomicAdd(counts, cnt0);
      for (int a = 0; a < nd; a++) {
        atomicMin(bbox + 2 * a, (uint32_t)mn0[a]);
        atomicMax(bbox + 2 * a + 1, (uint32_t)mx0[a]);
        atomicAdd(sums + a, sm0[a]);
      }
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
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m += __shfl_xor(m, o);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(nmask, m);
}

template <class T, class L>
__global__ __launch_bounds__(WG) void k_dust_apply(T* img, const L* lab, const uint8_t* erase, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG)
    if (erase[(uint64_t)lab[i]]) img[i] = T(0);
}

static thread_local std::string f_err;

static int ffail(int code, const char* msg) {
  f_err = msg;
  return code;
}

static inline unsigned fgrid(int64_t n, int64_t per, int64_t cap) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + per - 1) / per, cap));
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
// (min = 0xFFFFFFFF, max = 0 for absent labels); centroids: float64[(N+1)*nd].
// sums_scratch: uint64[(N+1)*nd].
int cc3d_amd_statistics(const void* labels, int dtype, int64_t s0, int64_t s1, int64_t s2, int ndim, int f_order,
                        uint64_t N, uint32_t* counts, uint32_t* bbox, double* centroids,
                        unsigned long long* sums_scratch, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n1 = (int64_t)N + 1, n = s0 * s1 * s2;
  const int nd = ndim == 2 ? 2 : 3;
  hipLaunchKernelGGL(k_stats_init, dim3(fgrid(n1, WG, 8192)), dim3(WG), 0, s, counts, bbox, sums_scratch, n1, nd);
  const dim3 g(fgrid(n, (int64_t)WG * STAT_RUN, 1 << 30));
  switch (dtype) {
    case DT_U8: hipLaunchKernelGGL((k_stats<uint8_t>), g, dim3(WG), 0, s, (const uint8_t*)labels, s0, s1, s2, nd, f_order, counts, bbox, sums_scratch); break;
    case DT_U16: hipLaunchKernelGGL((k_stats<uint16_t>), g, dim3(WG), 0, s, (const uint16_t*)labels, s0, s1, s2, nd, f_order, counts, bbox, sums_scratch); break;
    case DT_U32: hipLaunchKernelGGL((k_stats<uint32_t>), g, dim3(WG), 0, s, (const uint32_t*)labels, s0, s1, s2, nd, f_order, counts, bbox, sums_scratch); break;
    case DT_U64: hipLaunchKernelGGL((k_stats<uint64_t>), g, dim3(WG), 0, s, (const uint64_t*)labels, s0, s1, s2, nd, f_order, counts, bbox, sums_scratch); break;
    default: return ffail(CC3D_AMD_ERR_INVALID, "labels must be an integer type");
  }
  hipLaunchKernelGGL(k_centroids, dim3(fgrid(n1, WG, 8192)), dim3(WG), 0, s, counts, sums_scratch, centroids, n1, nd);
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

}  // extern "C"
