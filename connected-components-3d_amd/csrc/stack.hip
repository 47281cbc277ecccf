// Streaming a volume through the GPU in z-slabs (cc3d/__init__.py:353-501,
// connected_components_stack, and volumes too large for one labelling call).
//
// Each slab is labelled on its own (labels 1..N_k, numbered by first
// appearance inside the slab).  Slab k's labels become provisional ids
// offset_k + label, offset_k = N_0 + ... + N_{k-1}; since slabs follow each
// other in z and labels follow first appearance inside a slab, provisional
// ids increase with first appearance over the whole volume.  The seam
// between slab k-1's last plane and slab k's first plane joins provisional
// ids in a 64-bit forest (union by the smaller id, so every tree is rooted at
// its component's first appearance).  At the end each provisional id maps to
// 1 + the rank of its root among the roots -- the whole volume's labels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/cc3d_amd.h"
#include "common.hpp"

namespace cc3d_amd {

hipError_t scan_root_bits(const uint32_t* bits, int64_t nwords, uint32_t* bsum, uint32_t* prefix, Counters* ctr,
                          hipStream_t s);

namespace {

using u64 = unsigned long long;

__device__ inline u64 ffind(u64* F, u64 a) {
  for (;;) {
    const u64 p = __hip_atomic_load(F + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (p == a) return a;
    const u64 g = __hip_atomic_load(F + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (g != p) atomicMin(F + a, g);  // halving; parents only decrease
    a = p;
  }
}

// Union by the smaller id: link the larger root under the smaller one with
// atomicMin, retrying from what the atomic saw.
__device__ inline void funion(u64* F, u64 a, u64 b) {
  for (;;) {
    a = ffind(F, a);
    b = ffind(F, b);
    if (a == b) return;
    if (a < b) {
      const u64 t = a;
      a = b;
      b = t;
    }
    const u64 old = atomicMin(F + a, b);
    if (old == a) return;
    a = old;
  }
}

template <class T>
__device__ inline bool linked(T a, T b, bool binary) {
  return binary ? true : a == b;  // both labels nonzero: both voxels are foreground
}

// Seam unions: voxel (x, y) of slab k-1's last plane with its neighbours in
// slab k's first plane -- (x, y) for 6-connectivity, the 3 x 3 around it for
// 26 (18: the 3 x 3 without the corners).
template <class T>
__global__ __launch_bounds__(WG) void k_seam(const T* __restrict__ pv, const uint32_t* __restrict__ pl,
                                             const T* __restrict__ cv, const uint32_t* __restrict__ cl, int64_t sx,
                                             int64_t sy, u64 poff, u64 coff, int nb, bool binary, u64* F) {
  const int64_t n = sx * sy;
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) {
    const uint32_t a = pl[i];
    if (!a) continue;
    const T va = pv[i];
    const int64_t x = i % sx, y = i / sx;
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int m = (dx != 0) + (dy != 0);
        if (nb == 6 && m > 0) continue;
        if (nb == 18 && m > 1) continue;
        const int64_t qx = x + dx, qy = y + dy;
        if (qx < 0 || qx >= sx || qy < 0 || qy >= sy) continue;
        const int64_t j = qx + sx * qy;
        const uint32_t b = cl[j];
        if (!b || !linked<T>(va, cv[j], binary)) continue;
        funion(F, poff + a, coff + b);
      }
  }
}

// Foreground voxels of slab k's first plane that took a new provisional
// label in the slab's own pass (no linked backward neighbour in the plane)
// but have a linked neighbour in slab k-1's last plane: the whole volume's
// pass gives them none.
template <class T>
__global__ __launch_bounds__(WG) void k_seam_prov(const T* __restrict__ pv, const T* __restrict__ cv, int64_t sx,
                                                  int64_t sy, int nb, bool binary, u64* count) {
  const int64_t n = sx * sy;
  u64 c = 0;
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) {
    const T v = cv[i];
    if (v == T(0)) continue;
    const int64_t x = i % sx, y = i / sx;
    auto lk = [&](T q) { return q != T(0) && (binary || q == v); };
    bool back = x > 0 && lk(cv[i - 1]);
    if (nb != 6 && y > 0) {
      back = back || lk(cv[i - sx]) || (x > 0 && lk(cv[i - sx - 1])) || (x < sx - 1 && lk(cv[i - sx + 1]));
    } else if (y > 0) {
      back = back || lk(cv[i - sx]);
    }
    if (back) continue;
    bool below = false;
    for (int dy = -1; dy <= 1 && !below; dy++)
      for (int dx = -1; dx <= 1 && !below; dx++) {
        const int m = (dx != 0) + (dy != 0);
        if (nb == 6 && m > 0) continue;
        if (nb == 18 && m > 1) continue;
        const int64_t qx = x + dx, qy = y + dy;
        if (qx < 0 || qx >= sx || qy < 0 || qy >= sy) continue;
        below = lk(pv[qx + sx * qy]);
      }
    if (below) c++;
  }
  if (c) atomicAdd(count, c);
}

__global__ __launch_bounds__(WG) void k_forest_init(u64* F, u64 first, u64 last) {
  for (u64 p = first + (u64)blockIdx.x * WG + threadIdx.x; p < last; p += (u64)gridDim.x * WG) F[p] = p;
}

// Roots of ids 1..n: compress every id onto its root; roots set their bit.
__global__ __launch_bounds__(WG) void k_forest_roots(u64* F, u64 n, uint32_t* bits) {
  for (u64 p = 1 + (u64)blockIdx.x * WG + threadIdx.x; p <= n; p += (u64)gridDim.x * WG) {
    const u64 r = ffind(F, p);
    if (r == p) atomicOr(bits + (p >> 5), 1u << (p & 31));
  }
}

// Final label of every id (run after k_forest_roots finished): 1 + the rank
// of its root.  Written to a separate table so no lane reads a parent that
// another lane has already replaced by a label.
__global__ __launch_bounds__(WG) void k_forest_labels(const u64* F, u64 n, const uint32_t* bits,
                                                      const uint32_t* prefix, u64* lut) {
  for (u64 p = (u64)blockIdx.x * WG + threadIdx.x; p <= n; p += (u64)gridDim.x * WG) {
    if (p == 0) {
      lut[0] = 0;
      continue;
    }
    u64 r = p;
    for (;;) {
      const u64 q = F[r];
      if (q == r) break;
      r = q;
    }
    const uint32_t w = (uint32_t)(r >> 5);
    lut[p] = (u64)prefix[w] + (u64)__popc(bits[w] & ((1u << (r & 31)) - 1u)) + 1ull;
  }
}

__global__ __launch_bounds__(WG) void k_offset_labels(const uint32_t* __restrict__ lab, u64 n, u64 off,
                                                      u64* __restrict__ out) {
  for (u64 i = (u64)blockIdx.x * WG + threadIdx.x; i < n; i += (u64)gridDim.x * WG) {
    const uint32_t l = lab[i];
    out[i] = l ? (u64)l + off : 0ull;
  }
}

template <class O>
__global__ __launch_bounds__(WG) void k_map_labels(const u64* __restrict__ in, u64 n, const u64* __restrict__ lut,
                                                   O* __restrict__ out) {
  for (u64 i = (u64)blockIdx.x * WG + threadIdx.x; i < n; i += (u64)gridDim.x * WG) out[i] = (O)lut[in[i]];
}

static thread_local std::string g_serr;
static int sfail(int code, const std::string& msg) {
  g_serr = msg;
  return code;
}

unsigned sgrid(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + WG - 1) / WG, 1u << 16)); }

template <class T>
hipError_t launch_seam(const void* pv, const uint32_t* pl, const void* cv, const uint32_t* cl, int64_t sx, int64_t sy,
                       u64 poff, u64 coff, int nb, bool binary, u64* F, hipStream_t s) {
  hipLaunchKernelGGL((k_seam<T>), dim3(sgrid((uint64_t)(sx * sy))), dim3(WG), 0, s, static_cast<const T*>(pv), pl,
                     static_cast<const T*>(cv), cl, sx, sy, poff, coff, nb, binary, F);
  return hipGetLastError();
}

template <class T>
hipError_t launch_seam_prov(const void* pv, const void* cv, int64_t sx, int64_t sy, int nb, bool binary, u64* count,
                            hipStream_t s) {
  hipLaunchKernelGGL((k_seam_prov<T>), dim3(sgrid((uint64_t)(sx * sy))), dim3(WG), 0, s, static_cast<const T*>(pv),
                     static_cast<const T*>(cv), sx, sy, nb, binary, count);
  return hipGetLastError();
}

}  // namespace
}  // namespace cc3d_amd

using namespace cc3d_amd;

extern "C" {

const char* cc3d_amd_stack_last_error(void) { return g_serr.c_str(); }

int cc3d_amd_stack_seam(const void* prev_values, const uint32_t* prev_labels, const void* values,
                        const uint32_t* labels, int dtype, int64_t sx, int64_t sy, uint64_t prev_offset,
                        uint64_t offset, int connectivity, int binary_image, uint64_t* forest, void* stream) {
  if (!prev_values || !prev_labels || !values || !labels || !forest || sx <= 0 || sy <= 0)
    return sfail(CC3D_AMD_ERR_INVALID, "bad argument");
  if (connectivity != 6 && connectivity != 18 && connectivity != 26)
    return sfail(CC3D_AMD_ERR_INVALID, "seams join 6-, 18- or 26-connected slabs");
  hipStream_t s = static_cast<hipStream_t>(stream);
  u64* F = reinterpret_cast<u64*>(forest);
  const bool bin = binary_image != 0;
  hipError_t e;
  switch (dtype) {
    case DT_U8: e = launch_seam<uint8_t>(prev_values, prev_labels, values, labels, sx, sy, prev_offset, offset, connectivity, bin, F, s); break;
    case DT_U16: e = launch_seam<uint16_t>(prev_values, prev_labels, values, labels, sx, sy, prev_offset, offset, connectivity, bin, F, s); break;
    case DT_U32: e = launch_seam<uint32_t>(prev_values, prev_labels, values, labels, sx, sy, prev_offset, offset, connectivity, bin, F, s); break;
    case DT_U64: e = launch_seam<uint64_t>(prev_values, prev_labels, values, labels, sx, sy, prev_offset, offset, connectivity, bin, F, s); break;
    case DT_F32: e = launch_seam<float>(prev_values, prev_labels, values, labels, sx, sy, prev_offset, offset, connectivity, bin, F, s); break;
    case DT_F64: e = launch_seam<double>(prev_values, prev_labels, values, labels, sx, sy, prev_offset, offset, connectivity, bin, F, s); break;
    default: return sfail(CC3D_AMD_ERR_INVALID, "unsupported dtype");
  }
  if (e != hipSuccess) return sfail(CC3D_AMD_ERR_HIP, std::string("seam: ") + hipGetErrorString(e));
  return CC3D_AMD_OK;
}

int cc3d_amd_stack_seam_provisional(const void* prev_values, const void* values, int dtype, int64_t sx, int64_t sy,
                                    int connectivity, int binary_image, uint64_t* count, void* stream) {
  if (!prev_values || !values || !count || sx <= 0 || sy <= 0) return sfail(CC3D_AMD_ERR_INVALID, "bad argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  u64* c = reinterpret_cast<u64*>(count);
  const bool bin = binary_image != 0;
  hipError_t e;
  switch (dtype) {
    case DT_U8: e = launch_seam_prov<uint8_t>(prev_values, values, sx, sy, connectivity, bin, c, s); break;
    case DT_U16: e = launch_seam_prov<uint16_t>(prev_values, values, sx, sy, connectivity, bin, c, s); break;
    case DT_U32: e = launch_seam_prov<uint32_t>(prev_values, values, sx, sy, connectivity, bin, c, s); break;
    case DT_U64: e = launch_seam_prov<uint64_t>(prev_values, values, sx, sy, connectivity, bin, c, s); break;
    case DT_F32: e = launch_seam_prov<float>(prev_values, values, sx, sy, connectivity, bin, c, s); break;
    case DT_F64: e = launch_seam_prov<double>(prev_values, values, sx, sy, connectivity, bin, c, s); break;
    default: return sfail(CC3D_AMD_ERR_INVALID, "unsupported dtype");
  }
  if (e != hipSuccess) return sfail(CC3D_AMD_ERR_HIP, std::string("seam count: ") + hipGetErrorString(e));
  return CC3D_AMD_OK;
}

int cc3d_amd_forest_init(uint64_t* forest, uint64_t first, uint64_t last, void* stream) {
  if (!forest || last < first) return sfail(CC3D_AMD_ERR_INVALID, "bad argument");
  if (last == first) return CC3D_AMD_OK;
  hipLaunchKernelGGL(k_forest_init, dim3(sgrid(last - first)), dim3(WG), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<u64*>(forest), (u64)first, (u64)last);
  if (hipGetLastError() != hipSuccess) return sfail(CC3D_AMD_ERR_HIP, "forest init launch failed");
  return CC3D_AMD_OK;
}

int cc3d_amd_forest_labels(const uint64_t* forest_in, uint64_t n_ids, uint64_t* lut, uint64_t* N, void* stream) {
  if (!forest_in || !lut || !N) return sfail(CC3D_AMD_ERR_INVALID, "bad argument");
  if (n_ids >= 0xFFFFFFFFull) return sfail(CC3D_AMD_ERR_INVALID, "more than 2^32-2 provisional labels");
  hipStream_t s = static_cast<hipStream_t>(stream);
  u64* F = const_cast<u64*>(reinterpret_cast<const u64*>(forest_in));
  const int64_t nwords = (int64_t)((n_ids + 1 + 31) / 32);
  const int64_t nblk = (nwords + 4095) / 4096;
  char* ws = nullptr;
  const size_t bytes = (size_t)nwords * 8 + (size_t)(nblk + 1) * 4 + sizeof(Counters) + 512;
  if (hipMallocAsync(reinterpret_cast<void**>(&ws), bytes, s) != hipSuccess)
    return sfail(CC3D_AMD_ERR_NOMEM, "forest labels: could not allocate");
  uint32_t* bits = reinterpret_cast<uint32_t*>(ws);
  uint32_t* prefix = bits + nwords;
  uint32_t* bsum = prefix + nwords;
  Counters* ctr = reinterpret_cast<Counters*>(ws + (((size_t)nwords * 8 + (size_t)(nblk + 1) * 4 + 255) & ~(size_t)255));
  hipError_t e = hipMemsetAsync(bits, 0, (size_t)nwords * 4, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_forest_roots, dim3(sgrid(n_ids)), dim3(WG), 0, s, F, (u64)n_ids, bits);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = scan_root_bits(bits, nwords, bsum, prefix, ctr, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_forest_labels, dim3(sgrid(n_ids + 1)), dim3(WG), 0, s, F, (u64)n_ids, bits, prefix,
                       reinterpret_cast<u64*>(lut));
    e = hipGetLastError();
  }
  Counters h{};
  if (e == hipSuccess) e = hipMemcpyAsync(&h, ctr, sizeof(Counters), hipMemcpyDeviceToHost, s);
  const hipError_t ef = hipFreeAsync(ws, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = ef;
  if (e != hipSuccess) return sfail(CC3D_AMD_ERR_HIP, std::string("forest labels: ") + hipGetErrorString(e));
  *N = h.N;
  return CC3D_AMD_OK;
}

int cc3d_amd_offset_labels(const uint32_t* labels, uint64_t n, uint64_t offset, uint64_t* out, void* stream) {
  if ((n && (!labels || !out))) return sfail(CC3D_AMD_ERR_INVALID, "bad argument");
  if (!n) return CC3D_AMD_OK;
  hipLaunchKernelGGL(k_offset_labels, dim3(sgrid(n)), dim3(WG), 0, static_cast<hipStream_t>(stream), labels, (u64)n,
                     (u64)offset, reinterpret_cast<u64*>(out));
  if (hipGetLastError() != hipSuccess) return sfail(CC3D_AMD_ERR_HIP, "offset launch failed");
  return CC3D_AMD_OK;
}

int cc3d_amd_map_labels(const uint64_t* ids, uint64_t n, const uint64_t* lut, void* out, int out_dtype,
                        void* stream) {
  if ((n && (!ids || !lut || !out))) return sfail(CC3D_AMD_ERR_INVALID, "bad argument");
  if (!n) return CC3D_AMD_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const u64* in = reinterpret_cast<const u64*>(ids);
  const u64* lt = reinterpret_cast<const u64*>(lut);
  switch (out_dtype) {
    case DT_U16: hipLaunchKernelGGL((k_map_labels<uint16_t>), dim3(sgrid(n)), dim3(WG), 0, s, in, (u64)n, lt, (uint16_t*)out); break;
    case DT_U32: hipLaunchKernelGGL((k_map_labels<uint32_t>), dim3(sgrid(n)), dim3(WG), 0, s, in, (u64)n, lt, (uint32_t*)out); break;
    case DT_U64: hipLaunchKernelGGL((k_map_labels<u64>), dim3(sgrid(n)), dim3(WG), 0, s, in, (u64)n, lt, (u64*)out); break;
    default: return sfail(CC3D_AMD_ERR_INVALID, "out_dtype must be uint16, uint32 or uint64");
  }
  if (hipGetLastError() != hipSuccess) return sfail(CC3D_AMD_ERR_HIP, "map launch failed");
  return CC3D_AMD_OK;
}

}  // extern "C"
