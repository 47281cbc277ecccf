// The binding's single-foreground-row shortcut (cc3d/fastcc3d.pyx:469-470,
// :644-679) on the GPU.
//
// When delta == 0 and only one row (y, z) of the volume holds foreground, the
// reference labels that row in Python instead of calling its C++ core:
//
//   last = 0 (a C int64)
//   for each x: v == 0          -> background, last = 0
//               v == last       -> continue the current label
//               else            -> N += 1, out = N, last = int64(v)
//
// `last` is a C int64, so a float row is compared with the truncated value of
// the run's first voxel: [2.5, 2.5] gives two labels, [2.5, 2.0] one.
// Because a voxel that continues a run equals `last`, which is an integer, its
// own truncation equals `last` too.  So voxel x starts a new label exactly
// when v[x] != 0 and (x == 0 or v[x] != trunc(v[x-1])): a local rule, and the
// label of a foreground voxel is the number of starts at or before it.
//
// Conversion failures raise in the reference at the start voxel that causes
// them (non-convertible values never continue a run):
//   NaN                      -> ValueError "cannot convert float NaN to integer"
//   +-infinity               -> OverflowError "cannot convert float infinity to integer"
//   float outside [-2^63, 2^63), uint64 >= 2^63
//                            -> OverflowError "Python int too large to convert to C long"
// The reference stores each label into the output array before converting
// the value, so a label beyond the output dtype's range raises first
// (numpy's "Python integer M out of bounds for uint16").  The kernels report
// the first failing start's rank; the caller orders the two.
#include <hip/hip_runtime.h>

#include <math.h>

#include <algorithm>
#include <string>

#include "../../include/cc3d_amd.h"
#include "common.hpp"

namespace cc3d_amd {

constexpr int SR_PER_THREAD = 16;
constexpr int SR_CHUNK = WG * SR_PER_THREAD;

struct SrResult {
  unsigned long long N;         // starts in the row
  unsigned long long bad;       // (position << 2) | kind of the first failing start; ~0 if none
  unsigned long long bad_rank;  // 1-based rank of that start among the starts
};

// Failure kind of a start voxel's conversion to int64 (0 = converts).
template <class T>
__device__ inline int sr_bad(T v, int check_u64) {
  if constexpr (sizeof(T) == 8 && T(-1) > T(0)) {
    return (check_u64 && (v >> 63)) ? 3 : 0;
  } else if constexpr (T(-1) > T(0)) {
    return 0;
  } else {
    if (v != v) return 1;
    if (v == T(INFINITY) || v == T(-INFINITY)) return 2;
    if (v >= T(9223372036854775808.0) || v < T(-9223372036854775808.0)) return 3;
    return 0;
  }
}

// The value a run started at v continues with: v itself for integers, its
// truncation for floats (NaN when the start does not convert, so nothing
// continues it).
template <class T>
__device__ inline T sr_carry(T v) {
  if constexpr (T(-1) > T(0)) {
    return v;
  } else {
    if (sr_bad<T>(v, 0)) return T(NAN);
    return sizeof(T) == 4 ? T(truncf((float)v)) : T(trunc((double)v));
  }
}

template <class T>
__device__ inline bool sr_start(const T* row, int64_t x) {
  const T v = row[x];
  if (v == T(0)) return false;  // -0.0 is background too
  if (x == 0) return true;
  return !(v == sr_carry<T>(row[x - 1]));
}

__device__ inline uint32_t sr_block_scan(uint32_t v, uint32_t* tmp, uint32_t& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) tmp[w] = x;
  __syncthreads();
  uint32_t off = 0;
  total = 0;
#pragma unroll
  for (int i = 0; i < WG / 64; i++) {
    const uint32_t t = tmp[i];
    if (i < w) off += t;
    total += t;
  }
  __syncthreads();
  return off + x - v;
}

// Per chunk: number of starts; the first failing start of the row.
template <class T>
__global__ __launch_bounds__(WG) void k_sr_count(const T* row, int64_t sx, int check_u64, uint32_t* bcnt,
                                                 SrResult* res) {
  __shared__ uint32_t tmp[WG / 64];
  const int64_t x0 = (int64_t)blockIdx.x * SR_CHUNK + (int64_t)threadIdx.x * SR_PER_THREAD;
  uint32_t c = 0;
  unsigned long long bad = ~0ull;
  for (int k = 0; k < SR_PER_THREAD; k++) {
    const int64_t x = x0 + k;
    if (x >= sx) break;
    if (sr_start<T>(row, x)) {
      c++;
      const int kind = sr_bad<T>(row[x], check_u64);
      if (kind && bad == ~0ull) bad = ((unsigned long long)x << 2) | (unsigned)kind;
    }
  }
  if (bad != ~0ull) atomicMin(&res->bad, bad);
  uint32_t total;
  sr_block_scan(c, tmp, total);
  if (threadIdx.x == 0) bcnt[blockIdx.x] = total;
}

// Exclusive scan of the chunk counts (one workgroup); N.
__global__ __launch_bounds__(WG) void k_sr_scan(uint32_t* bcnt, int64_t nblk, SrResult* res) {
  __shared__ uint32_t tmp[WG / 64];
  unsigned long long carry = 0;
  for (int64_t off = 0; off < nblk; off += WG) {
    const int64_t i = off + threadIdx.x;
    const uint32_t v = i < nblk ? bcnt[i] : 0u;
    uint32_t total;
    const uint32_t ex = sr_block_scan(v, tmp, total);
    if (i < nblk) bcnt[i] = (uint32_t)(carry + ex);
    carry += total;
  }
  if (threadIdx.x == 0) res->N = carry;
}

// Labels of the row: the number of starts at or before each foreground voxel.
template <class T, class OUT>
__global__ __launch_bounds__(WG) void k_sr_write(const T* row, int64_t sx, const uint32_t* bcnt, OUT* out,
                                                 SrResult* res) {
  __shared__ uint32_t tmp[WG / 64];
  const int64_t x0 = (int64_t)blockIdx.x * SR_CHUNK + (int64_t)threadIdx.x * SR_PER_THREAD;
  uint32_t flags = 0, c = 0;
  for (int k = 0; k < SR_PER_THREAD; k++) {
    const int64_t x = x0 + k;
    if (x >= sx) break;
    if (sr_start<T>(row, x)) {
      flags |= 1u << k;
      c++;
    }
  }
  uint32_t total;
  uint32_t run = bcnt[blockIdx.x] + sr_block_scan(c, tmp, total);
  const unsigned long long bad = res->bad;
  const int64_t badx = bad == ~0ull ? -1 : (int64_t)(bad >> 2);
  for (int k = 0; k < SR_PER_THREAD; k++) {
    const int64_t x = x0 + k;
    if (x >= sx) break;
    if ((flags >> k) & 1u) run++;
    if (x == badx) res->bad_rank = run;
    out[x] = row[x] == T(0) ? OUT(0) : OUT(run);
  }
}

__global__ void k_sr_init(SrResult* res) {
  res->N = 0;
  res->bad = ~0ull;
  res->bad_rank = 0;
}

template <class OUT>
__global__ void k_sr_zero(OUT* out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = OUT(0);
}

static thread_local std::string sr_err;

template <class T, class OUT>
static hipError_t sr_launch(const void* in, int64_t sx, int64_t voxels, int64_t row, int check_u64, void* out,
                            uint32_t* bcnt, SrResult* res, int64_t nblk, hipStream_t s) {
  const T* r = static_cast<const T*>(in) + row * sx;
  OUT* o = static_cast<OUT*>(out);
  const unsigned zgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((voxels + WG - 1) / WG, 4096));
  hipLaunchKernelGGL(k_sr_init, dim3(1), dim3(1), 0, s, res);
  hipLaunchKernelGGL(k_sr_zero<OUT>, dim3(zgrid), dim3(WG), 0, s, o, voxels);
  hipLaunchKernelGGL(k_sr_count<T>, dim3((unsigned)nblk), dim3(WG), 0, s, r, sx, check_u64, bcnt, res);
  hipLaunchKernelGGL(k_sr_scan, dim3(1), dim3(WG), 0, s, bcnt, nblk, res);
  hipLaunchKernelGGL((k_sr_write<T, OUT>), dim3((unsigned)nblk), dim3(WG), 0, s, r, sx, bcnt, o + row * sx, res);
  return hipGetLastError();
}

template <class T>
static hipError_t sr_dispatch_out(int out_dtype, const void* in, int64_t sx, int64_t voxels, int64_t row,
                                  int check_u64, void* out, uint32_t* bcnt, SrResult* res, int64_t nblk,
                                  hipStream_t s) {
  switch (out_dtype) {
    case DT_U16: return sr_launch<T, uint16_t>(in, sx, voxels, row, check_u64, out, bcnt, res, nblk, s);
    case DT_U32: return sr_launch<T, uint32_t>(in, sx, voxels, row, check_u64, out, bcnt, res, nblk, s);
    default: return sr_launch<T, uint64_t>(in, sx, voxels, row, check_u64, out, bcnt, res, nblk, s);
  }
}

}  // namespace cc3d_amd

using namespace cc3d_amd;

extern "C" {

int cc3d_amd_special_row(const void* in, int in_dtype, int check_u64, int64_t sx, int64_t voxels, int64_t row,
                         void* out, int out_dtype, uint64_t* N, int* fail_kind, uint64_t* fail_rank,
                         void* stream) {
  if (!in || !out || !N || !fail_kind || !fail_rank || sx <= 0 || voxels <= 0 || voxels % sx != 0 || row < 0 ||
      (row + 1) * sx > voxels || in_dtype < DT_U8 || in_dtype > DT_F64 || out_dtype < DT_U16 ||
      out_dtype > DT_U64) {
    sr_err = "cc3d_amd_special_row: invalid arguments";
    return CC3D_AMD_ERR_INVALID;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t nblk = (sx + SR_CHUNK - 1) / SR_CHUNK;
  void* ws = nullptr;
  const size_t ws_bytes = 256 + (size_t)nblk * sizeof(uint32_t);
  hipError_t e = hipMallocAsync(&ws, ws_bytes, s);
  if (e != hipSuccess) {
    sr_err = std::string("cc3d_amd_special_row: ") + hipGetErrorString(e);
    return CC3D_AMD_ERR_NOMEM;
  }
  SrResult* res = static_cast<SrResult*>(ws);
  uint32_t* bcnt = reinterpret_cast<uint32_t*>(static_cast<char*>(ws) + 256);
  switch (in_dtype) {
    case DT_U8: e = sr_dispatch_out<uint8_t>(out_dtype, in, sx, voxels, row, 0, out, bcnt, res, nblk, s); break;
    case DT_U16: e = sr_dispatch_out<uint16_t>(out_dtype, in, sx, voxels, row, 0, out, bcnt, res, nblk, s); break;
    case DT_U32: e = sr_dispatch_out<uint32_t>(out_dtype, in, sx, voxels, row, 0, out, bcnt, res, nblk, s); break;
    case DT_U64:
      e = sr_dispatch_out<uint64_t>(out_dtype, in, sx, voxels, row, check_u64, out, bcnt, res, nblk, s);
      break;
    case DT_F32: e = sr_dispatch_out<float>(out_dtype, in, sx, voxels, row, 0, out, bcnt, res, nblk, s); break;
    default: e = sr_dispatch_out<double>(out_dtype, in, sx, voxels, row, 0, out, bcnt, res, nblk, s); break;
  }
  SrResult h{};
  if (e == hipSuccess) e = hipMemcpyAsync(&h, res, sizeof(h), hipMemcpyDeviceToHost, s);
  const hipError_t ef = hipFreeAsync(ws, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = ef;
  if (e != hipSuccess) {
    sr_err = std::string("cc3d_amd_special_row: ") + hipGetErrorString(e);
    return CC3D_AMD_ERR_HIP;
  }
  *N = h.N;
  *fail_kind = h.bad == ~0ull ? 0 : (int)(h.bad & 3ull);
  *fail_rank = h.bad == ~0ull ? 0 : h.bad_rank;
  return CC3D_AMD_OK;
}

const char* cc3d_amd_special_row_last_error(void) { return sr_err.c_str(); }

}  // extern "C"
