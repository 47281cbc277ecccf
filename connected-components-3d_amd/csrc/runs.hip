// Runs of a dense labelling and painting them back (cc3d/cc3d_graphs.hpp:
// 470-523, bound at cc3d/fastcc3d.pyx:1254-1308 as runs / draw, and used by
// each at :1310-1372).
//
// runs: the maximal runs of equal values of the flat array, grouped by label
// in ascending label order, each label's runs in position order; label 0 is
// left out, except that a one-voxel array yields its single run whatever its
// value (the reference's special case).  On the GPU: a bitmask of run starts
// (one thread per 32 positions), its popcount prefix (the scan of the
// labelling path), a scatter of (label, start) per run, and a stable radix
// sort of the runs by label.
//
// draw: every position of the given runs is set to one value.  The
// reference checks each run as it reaches it ("Invalid run." when the end
// lies past the image or start >= end) after painting the runs before it;
// here the first invalid run is found first, the runs before it are painted,
// and the same error is returned.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <string>

#include "../../include/cc3d_amd.h"
#include "common.hpp"

namespace cc3d_amd {

hipError_t scan_root_bits(const uint32_t* bits, int64_t nwords, uint32_t* bsum, uint32_t* prefix, Counters* ctr,
                          hipStream_t s);

namespace {

template <class T>
__global__ __launch_bounds__(WG) void k_runs_mark(const T* __restrict__ lab, uint64_t n, uint32_t* bits,
                                                  uint64_t nwords) {
  for (uint64_t w = (uint64_t)blockIdx.x * WG + threadIdx.x; w < nwords; w += (uint64_t)gridDim.x * WG) {
    const uint64_t i0 = 32 * w;
    uint32_t m = 0;
    T prev = i0 > 0 ? lab[i0 - 1] : T(0);
    for (int j = 0; j < 32; j++) {
      const uint64_t i = i0 + j;
      if (i >= n) break;
      const T v = lab[i];
      if (i == 0 || v != prev) m |= 1u << j;
      prev = v;
    }
    bits[w] = m;
  }
}

template <class T>
__global__ __launch_bounds__(WG) void k_runs_scatter(const T* __restrict__ lab, const uint32_t* bits,
                                                     const uint32_t* prefix, uint64_t nwords, unsigned long long* key,
                                                     uint32_t* idx, unsigned long long* start) {
  for (uint64_t w = (uint64_t)blockIdx.x * WG + threadIdx.x; w < nwords; w += (uint64_t)gridDim.x * WG) {
    uint32_t m = bits[w];
    uint32_t r = prefix[w];
    while (m) {
      const int j = __builtin_ctz(m);
      m &= m - 1u;
      const uint64_t i = 32 * w + j;
      key[r] = (unsigned long long)lab[i];
      idx[r] = r;
      start[r] = i;
      r++;
    }
  }
}

__global__ __launch_bounds__(WG) void k_runs_gather(const unsigned long long* skey, const uint32_t* sidx,
                                                    const unsigned long long* start, uint64_t nruns, uint64_t n,
                                                    unsigned long long* olab, unsigned long long* ostart,
                                                    unsigned long long* oend) {
  for (uint64_t k = (uint64_t)blockIdx.x * WG + threadIdx.x; k < nruns; k += (uint64_t)gridDim.x * WG) {
    const uint32_t r = sidx[k];
    olab[k] = skey[k];
    ostart[k] = start[r];
    oend[k] = r + 1 < nruns ? start[r + 1] : n;
  }
}

// First invalid run (end past the image, or start >= end) into *first.
__global__ __launch_bounds__(WG) void k_draw_check(const unsigned long long* st, const unsigned long long* en,
                                                   uint64_t nruns, uint64_t n, unsigned long long* first) {
  for (uint64_t k = (uint64_t)blockIdx.x * WG + threadIdx.x; k < nruns; k += (uint64_t)gridDim.x * WG)
    if (en[k] > n || st[k] >= en[k]) atomicMin(first, (unsigned long long)k);
}

// One workgroup per run (grid-stride over runs), threads striding its span.
template <class T>
__global__ __launch_bounds__(WG) void k_draw(T* img, const unsigned long long* st, const unsigned long long* en,
                                             const unsigned long long* limit, uint64_t nruns, T val) {
  const uint64_t lim = limit ? *limit : nruns;
  const uint64_t nr = lim < nruns ? lim : nruns;
  for (uint64_t k = blockIdx.x; k < nr; k += gridDim.x) {
    const uint64_t b = st[k], e = en[k];
    for (uint64_t i = b + threadIdx.x; i < e; i += WG) img[i] = val;
  }
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

static thread_local std::string g_rerr;
static int rfail(int code, const std::string& msg) {
  g_rerr = msg;
  return code;
}

#define RCHK(expr)                                                                                      \
  do {                                                                                                  \
    hipError_t _e = (expr);                                                                             \
    if (_e != hipSuccess) return rfail(CC3D_AMD_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

unsigned rgrid(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + WG - 1) / WG, 1u << 16)); }

}  // namespace
}  // namespace cc3d_amd

using namespace cc3d_amd;

struct cc3d_amd_runs {
  uint64_t n = 0;  // runs
  unsigned long long* label = nullptr;  // device, n each
  unsigned long long* start = nullptr;
  unsigned long long* end = nullptr;
  hipStream_t stream = nullptr;
};

namespace {

template <class T>
int runs_run(const T* lab, uint64_t n, cc3d_amd_runs* res, hipStream_t s) {
  const uint64_t nwords = (n + 31) / 32;
  const int64_t nblk = ((int64_t)nwords + 4095) / 4096;
  uint32_t* bits = nullptr;
  uint32_t* prefix = nullptr;
  uint32_t* bsum = nullptr;
  Counters* ctr = nullptr;
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&bits), nwords * 4, s));
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&prefix), nwords * 4, s));
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&bsum), (nblk + 1) * 4, s));
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&ctr), sizeof(Counters), s));
  hipLaunchKernelGGL((k_runs_mark<T>), dim3(rgrid(nwords)), dim3(WG), 0, s, lab, n, bits, nwords);
  RCHK(hipGetLastError());
  RCHK(scan_root_bits(bits, (int64_t)nwords, bsum, prefix, ctr, s));
  Counters h{};
  RCHK(hipMemcpyAsync(&h, ctr, sizeof(Counters), hipMemcpyDeviceToHost, s));
  RCHK(hipStreamSynchronize(s));
  const uint64_t nr = h.N;
  unsigned long long *key = nullptr, *skey = nullptr, *start = nullptr;
  uint32_t *idx = nullptr, *sidx = nullptr;
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&key), nr * 8, s));
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&skey), nr * 8, s));
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&start), nr * 8, s));
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&idx), nr * 4, s));
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&sidx), nr * 4, s));
  hipLaunchKernelGGL((k_runs_scatter<T>), dim3(rgrid(nwords)), dim3(WG), 0, s, lab, bits, prefix, nwords, key, idx,
                     start);
  RCHK(hipGetLastError());
  size_t tmp_bytes = 0;
  RCHK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, key, skey, idx, sidx, (size_t)nr, 0, 8 * (int)sizeof(T), s));
  void* tmp = nullptr;
  RCHK(hipMallocAsync(&tmp, std::max<size_t>(tmp_bytes, 8), s));
  RCHK(rocprim::radix_sort_pairs(tmp, tmp_bytes, key, skey, idx, sidx, (size_t)nr, 0, 8 * (int)sizeof(T), s));
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&res->label), std::max<uint64_t>(nr, 1) * 8, s));
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&res->start), std::max<uint64_t>(nr, 1) * 8, s));
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&res->end), std::max<uint64_t>(nr, 1) * 8, s));
  hipLaunchKernelGGL(k_runs_gather, dim3(rgrid(nr)), dim3(WG), 0, s, skey, sidx, start, nr, n, res->label, res->start,
                     res->end);
  RCHK(hipGetLastError());
  res->n = nr;
  for (void* p : {(void*)bits, (void*)prefix, (void*)bsum, (void*)ctr, (void*)key, (void*)skey, (void*)start,
                  (void*)idx, (void*)sidx, tmp})
    RCHK(hipFreeAsync(p, s));
  RCHK(hipStreamSynchronize(s));
  return CC3D_AMD_OK;
}

}  // namespace

extern "C" {

const char* cc3d_amd_runs_last_error(void) { return g_rerr.c_str(); }

int cc3d_amd_extract_runs(const void* labels, int label_size, uint64_t n, cc3d_amd_runs** result, uint64_t* n_runs,
                          void* stream) {
  if (!result || !n_runs || (n > 0 && !labels)) return rfail(CC3D_AMD_ERR_INVALID, "bad argument");
  if (n >= 0xFFFFFFFFull) return rfail(CC3D_AMD_ERR_INVALID, "runs: arrays of 2^32-1 elements or more are not supported");
  *result = nullptr;
  *n_runs = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  cc3d_amd_runs* res = new cc3d_amd_runs();
  res->stream = s;
  int rc = CC3D_AMD_OK;
  if (n > 0) {
    switch (label_size) {
      case 1: rc = runs_run<uint8_t>(static_cast<const uint8_t*>(labels), n, res, s); break;
      case 2: rc = runs_run<uint16_t>(static_cast<const uint16_t*>(labels), n, res, s); break;
      case 4: rc = runs_run<uint32_t>(static_cast<const uint32_t*>(labels), n, res, s); break;
      case 8: rc = runs_run<uint64_t>(static_cast<const uint64_t*>(labels), n, res, s); break;
      default: rc = rfail(CC3D_AMD_ERR_INVALID, "label_size must be 1, 2, 4 or 8");
    }
  }
  if (rc != CC3D_AMD_OK) {
    cc3d_amd_runs_free(res);
    return rc;
  }
  *result = res;
  *n_runs = res->n;
  return CC3D_AMD_OK;
}

int cc3d_amd_runs_device(cc3d_amd_runs* r, const uint64_t** label, const uint64_t** start, const uint64_t** end) {
  if (!r) return rfail(CC3D_AMD_ERR_INVALID, "bad argument");
  if (label) *label = reinterpret_cast<const uint64_t*>(r->label);
  if (start) *start = reinterpret_cast<const uint64_t*>(r->start);
  if (end) *end = reinterpret_cast<const uint64_t*>(r->end);
  return CC3D_AMD_OK;
}

int cc3d_amd_runs_fetch(cc3d_amd_runs* r, uint64_t* label, uint64_t* start, uint64_t* end) {
  if (!r) return rfail(CC3D_AMD_ERR_INVALID, "bad argument");
  if (r->n == 0) return CC3D_AMD_OK;
  if (!label || !start || !end) return rfail(CC3D_AMD_ERR_INVALID, "bad argument");
  RCHK(hipMemcpyAsync(label, r->label, r->n * 8, hipMemcpyDeviceToHost, r->stream));
  RCHK(hipMemcpyAsync(start, r->start, r->n * 8, hipMemcpyDeviceToHost, r->stream));
  RCHK(hipMemcpyAsync(end, r->end, r->n * 8, hipMemcpyDeviceToHost, r->stream));
  RCHK(hipStreamSynchronize(r->stream));
  return CC3D_AMD_OK;
}

void cc3d_amd_runs_free(cc3d_amd_runs* r) {
  if (!r) return;
  if (r->label) (void)hipFreeAsync(r->label, r->stream);
  if (r->start) (void)hipFreeAsync(r->start, r->stream);
  if (r->end) (void)hipFreeAsync(r->end, r->stream);
  (void)hipStreamSynchronize(r->stream);
  delete r;
}

int cc3d_amd_draw(void* image, int image_size, uint64_t n, uint64_t value, const uint64_t* start, const uint64_t* end,
                  uint64_t n_runs, void* stream) {
  if ((n_runs > 0 && (!start || !end)) || (n > 0 && !image)) return rfail(CC3D_AMD_ERR_INVALID, "bad argument");
  if (n_runs == 0) return CC3D_AMD_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned long long* first = nullptr;
  RCHK(hipMallocAsync(reinterpret_cast<void**>(&first), 8, s));
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, first, ~0ull);
  const auto* st = reinterpret_cast<const unsigned long long*>(start);
  const auto* en = reinterpret_cast<const unsigned long long*>(end);
  hipLaunchKernelGGL(k_draw_check, dim3(rgrid(n_runs)), dim3(WG), 0, s, st, en, n_runs, n, first);
  const dim3 grid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>(n_runs, 1u << 16)));
  switch (image_size) {
    case 1: hipLaunchKernelGGL((k_draw<uint8_t>), grid, dim3(WG), 0, s, (uint8_t*)image, st, en, first, n_runs, (uint8_t)value); break;
    case 2: hipLaunchKernelGGL((k_draw<uint16_t>), grid, dim3(WG), 0, s, (uint16_t*)image, st, en, first, n_runs, (uint16_t)value); break;
    case 4: hipLaunchKernelGGL((k_draw<uint32_t>), grid, dim3(WG), 0, s, (uint32_t*)image, st, en, first, n_runs, (uint32_t)value); break;
    case 8: hipLaunchKernelGGL((k_draw<uint64_t>), grid, dim3(WG), 0, s, (uint64_t*)image, st, en, first, n_runs, (uint64_t)value); break;
    default: (void)hipFreeAsync(first, s); return rfail(CC3D_AMD_ERR_INVALID, "image_size must be 1, 2, 4 or 8");
  }
  RCHK(hipGetLastError());
  unsigned long long bad = 0;
  RCHK(hipMemcpyAsync(&bad, first, 8, hipMemcpyDeviceToHost, s));
  RCHK(hipFreeAsync(first, s));
  RCHK(hipStreamSynchronize(s));
  if (bad != ~0ull) return rfail(CC3D_AMD_ERR_RUNTIME, "Invalid run.");
  return CC3D_AMD_OK;
}

}  // extern "C"
