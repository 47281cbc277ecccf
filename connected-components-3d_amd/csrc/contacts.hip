// Region adjacency and contact areas (cc3d/cc3d_graphs.hpp:299-468, bound at
// cc3d/fastcc3d.pyx:1180-1252 as contacts / region_graph).
//
// The reference walks the volume in raster order; for each nonzero voxel it
// visits the backward positions its neighbourhood offsets name
// (compute_neighborhood, cc3d/cc3d_graphs.hpp:238-297) and, for each
// neighbour holding another nonzero label, adds the position's area to the
// float32 accumulator of the label pair (smaller label first).  The offsets
// are reproduced as the reference computes them, quirks included: in 3D the
// corner offsets test only y and z, so on the x borders they name the edge
// neighbour a second time; in 2D the diagonal offsets test nothing, so on
// the borders they name a face neighbour a second time, or on row 0 the
// forward neighbour (x + 1).  Areas: faces wy*wz, wx*wz, wx*wy (3D) or wy,
// wx (2D), every other position 0; with surface_area off every position 1.
//
// Here each voxel is one thread.  Pairs go into an open-addressing hash
// table with, per pair, the number of contributions of each area class
// (the x, y and z faces, and the rest).  The float32 sum the reference
// forms sequentially is then recovered per pair:
//   * exactly, when every partial sum is representable: all contributions
//     are multiples of one power of two 2^g and their total is below 2^24 *
//     2^g (then order cannot matter), or all summands are 1 (the sum stops
//     at 2^24);
//   * by repeating the one summand, when a pair has a single nonzero area;
//   * otherwise by replaying the pair's contributions in raster order: they
//     are emitted as (pair rank, position * 13 + offset index) keys, radix
//     sorted, and summed one pair per thread.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/cc3d_amd.h"
#include "common.hpp"

namespace cc3d_amd {
namespace {

// Label dtype codes of this entry point (signed types order pairs signed).
enum : int { CT_U8 = 0, CT_U16 = 1, CT_U32 = 2, CT_U64 = 3, CT_I8 = 6, CT_I16 = 7, CT_I32 = 8, CT_I64 = 9 };

// Hash table.  Labels of at most 32 bits: one word per slot, the bit
// patterns of (lo, hi) as lo << 32 | hi.  64-bit labels: k0 = lo, k1 = hi.
// 0 marks an empty word: both labels of a pair are nonzero.
struct Table {
  unsigned long long* k0;
  unsigned long long* k1;
  unsigned long long* cnt;  // 4 per slot: classes 0, 1, 2 (faces), 3 (other)
  unsigned long long* used;  // slots claimed
  int* overflow;
  uint64_t cap;  // power of two
  uint64_t limit;  // claims beyond this set *overflow (the table is rebuilt larger)
};

__device__ inline uint64_t fmix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

template <class T>
__device__ inline uint64_t bits_of(T v) {
  if constexpr (sizeof(T) == 1) return (uint64_t)__builtin_bit_cast(uint8_t, v);
  else if constexpr (sizeof(T) == 2) return (uint64_t)__builtin_bit_cast(uint16_t, v);
  else if constexpr (sizeof(T) == 4) return (uint64_t)__builtin_bit_cast(uint32_t, v);
  else return __builtin_bit_cast(uint64_t, v);
}

// Slot of pair (lo, hi); claims an empty slot when `insert`.  Returns cap
// when the pair is absent (lookup) or the table is full (insert).
template <class T>
__device__ inline uint64_t table_slot(const Table& tb, T lo, T hi, bool insert) {
  constexpr bool WIDE = sizeof(T) == 8;
  const uint64_t a = bits_of(lo), b = bits_of(hi);
  const uint64_t key = WIDE ? a : ((a << 32) | b);
  const uint64_t mask = tb.cap - 1;
  uint64_t s = fmix64(WIDE ? (a ^ fmix64(b)) : key) & mask;
  for (uint64_t probe = 0; probe < tb.cap; probe++, s = (s + 1) & mask) {
    unsigned long long w = __hip_atomic_load(tb.k0 + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (w == 0) {
      if (!insert) return tb.cap;
      w = atomicCAS(tb.k0 + s, 0ull, (unsigned long long)key);
      if (w == 0) {
        w = key;
        if (atomicAdd(tb.used, 1ull) >= tb.limit) atomicOr(tb.overflow, 1);
      }
    }
    if (w != key) continue;
    if (WIDE) {
      unsigned long long v = __hip_atomic_load(tb.k1 + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (v == 0) {
        if (!insert) continue;  // a claim in progress is another pair's or ours; lookups run after inserts
        v = atomicCAS(tb.k1 + s, 0ull, (unsigned long long)b);
        if (v == 0) v = b;
      }
      if (v != b) continue;
    }
    return s;
  }
  if (insert) atomicOr(tb.overflow, 1);
  return tb.cap;
}

// Backward offsets exactly as the reference forms them (see the file
// comment); n = connectivity / 2 of them are visited.
__device__ inline void offsets3(int64_t* nb, int64_t x, int64_t y, int64_t z, int64_t sx, int64_t sy) {
  const int64_t sxy = sx * sy;
  const int64_t px = x < sx - 1 ? 1 : 0, mx = x > 0 ? -1 : 0;
  const int64_t py = y < sy - 1 ? sx : 0, my = y > 0 ? -sx : 0;
  const int64_t mz = z > 0 ? -sxy : 0;
  nb[0] = mx;
  nb[1] = my;
  nb[2] = mz;
  nb[3] = (mx && my) ? mx + my : 0;
  nb[4] = (px && my) ? px + my : 0;
  nb[5] = (mx && mz) ? mx + mz : 0;
  nb[6] = (px && mz) ? px + mz : 0;
  nb[7] = (my && mz) ? my + mz : 0;
  nb[8] = (py && mz) ? py + mz : 0;
  nb[9] = (my && mz) ? mx + my + mz : 0;
  nb[10] = (my && mz) ? px + my + mz : 0;
  nb[11] = (py && mz) ? mx + py + mz : 0;
  nb[12] = (py && mz) ? px + py + mz : 0;
}
__device__ inline void offsets2(int64_t* nb, int64_t x, int64_t y, int64_t sx, bool diag) {
  const int64_t px = x < sx - 1 ? 1 : 0, mx = x > 0 ? -1 : 0, my = y > 0 ? -sx : 0;
  nb[0] = mx;
  nb[1] = my;
  nb[2] = diag ? mx + my : 0;
  nb[3] = diag ? px + my : 0;
}

// Area class of offset index i: the faces keep their index, the rest 3.
__device__ inline int area_class(int i, bool d2) { return i < (d2 ? 2 : 3) ? i : 3; }

struct Geo {
  int64_t sx, sy, sz, V;
  int n;     // offsets visited (connectivity / 2)
  bool d2;   // 2D neighbourhood (connectivity 4 or 8)
};

// Counting pass: one thread per voxel, grid-stride.  A voxel's contributions
// to one neighbouring label are summed in registers (the two most recent
// labels are kept) before they reach the table.
template <class T>
__global__ __launch_bounds__(WG) void k_contacts_count(const T* __restrict__ lab, Geo g, Table tb) {
  for (int64_t loc = (int64_t)blockIdx.x * WG + threadIdx.x; loc < g.V; loc += (int64_t)gridDim.x * WG) {
    const T cur = lab[loc];
    if (cur == T(0)) continue;
    const int64_t x = loc % g.sx, r = loc / g.sx, y = r % g.sy, z = r / g.sy;
    int64_t nb[13];
    if (g.d2) offsets2(nb, x, y, g.sx, g.n == 4);
    else offsets3(nb, x, y, z, g.sx, g.sy);
    T pl[2] = {T(0), T(0)};
    uint32_t pc[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    int oldest = 0;
    auto flush = [&](int k) {
      if (pl[k] == T(0)) return;
      const T lo = cur < pl[k] ? cur : pl[k], hi = cur < pl[k] ? pl[k] : cur;
      const uint64_t s = table_slot<T>(tb, lo, hi, true);
      if (s < tb.cap) {
#pragma unroll
        for (int c = 0; c < 4; c++)
          if (pc[k][c]) atomicAdd(tb.cnt + 4 * s + c, (unsigned long long)pc[k][c]);
      }
      pl[k] = T(0);
#pragma unroll
      for (int c = 0; c < 4; c++) pc[k][c] = 0;
    };
    for (int i = 0; i < g.n; i++) {
      const T v = lab[loc + nb[i]];
      if (v == T(0) || v == cur) continue;
      const int c = area_class(i, g.d2);
      int k;
      if (v == pl[0]) k = 0;
      else if (v == pl[1]) k = 1;
      else {
        k = pl[0] == T(0) ? 0 : pl[1] == T(0) ? 1 : oldest;
        if (pl[k] != T(0)) flush(k);
        pl[k] = v;
        oldest = k ^ 1;
      }
      pc[k][c]++;
    }
    flush(0);
    flush(1);
  }
}

// Per pair: the float32 sum when it has a closed form (see the file
// comment), else a replay rank.  area[0..2]: face areas, area[3]: the rest.
struct Sums {
  float area[4];
  float* value;              // per slot
  uint32_t* rank_slot;       // replay rank -> slot
  unsigned long long* nrep;  // [0] ranks, [1] records
  uint32_t* rank;            // per slot: replay rank or 0xFFFFFFFF
};

__device__ inline bool low_bit(float a, int& e, uint32_t& m) {
  const uint32_t b = __float_as_uint(a) & 0x7FFFFFFFu;
  const uint32_t ex = b >> 23, mant = b & 0x7FFFFFu;
  if (ex == 0xFF) return false;  // inf / nan
  uint32_t mm = ex ? (mant | 0x800000u) : mant;
  int E = ex ? (int)ex - 150 : -149;
  const int tz = __builtin_ctz(mm);
  m = mm >> tz;
  e = E + tz;
  return true;
}

__global__ __launch_bounds__(WG) void k_contacts_sums(Table tb, Sums sm) {
  for (uint64_t s = (uint64_t)blockIdx.x * WG + threadIdx.x; s < tb.cap; s += (uint64_t)gridDim.x * WG) {
    sm.rank[s] = 0xFFFFFFFFu;
    if (tb.k0[s] == 0) continue;
    uint64_t n[4];
#pragma unroll
    for (int c = 0; c < 4; c++) n[c] = tb.cnt[4 * s + c];
    // the nonzero summands, as (area, count)
    int used = 0;
    float first = 0.0f;
    bool finite = true, mixed = false;
    int g = 1 << 30;
    for (int c = 0; c < 4; c++) {
      const float a = sm.area[c];
      if (n[c] == 0 || a == 0.0f) continue;
      if (used == 0) first = a;
      else if (__float_as_uint(a) != __float_as_uint(first)) mixed = true;
      used++;
      int e;
      uint32_t m;
      if (!low_bit(a, e, m)) finite = false;
      else g = e < g ? e : g;
    }
    float out = 0.0f;
    bool replay = false;
    if (used == 0) {
      out = 0.0f;
    } else {
      bool exact = finite;
      long long S = 0;
      uint64_t T = 0;
      if (exact) {
        for (int c = 0; c < 4 && exact; c++) {
          const float a = sm.area[c];
          if (n[c] == 0 || a == 0.0f) continue;
          int e;
          uint32_t m;
          low_bit(a, e, m);
          const int sh = e - g;
          if (sh >= 24 || n[c] >= (1ull << 24)) {
            exact = false;
            break;
          }
          const uint64_t M = (uint64_t)m << sh;
          if (M >= (1ull << 24)) {
            exact = false;
            break;
          }
          T += n[c] * M;
          S += (a < 0.0f ? -1ll : 1ll) * (long long)(n[c] * M);
        }
        exact = exact && T < (1ull << 24);
      }
      if (exact) {
        out = ldexpf((float)S, g);
      } else if (!mixed) {
        // one summand repeated: its sequential sum does not depend on order
        uint64_t cnt = 0;
        for (int c = 0; c < 4; c++)
          if (n[c] && sm.area[c] != 0.0f) cnt += n[c];
        float acc = 0.0f;
        if (first == 1.0f) {
          acc = cnt < (1ull << 24) ? (float)cnt : 16777216.0f;
        } else {
          for (uint64_t k = 0; k < cnt; k++) acc += first;
        }
        out = acc;
      } else {
        replay = true;
      }
    }
    sm.value[s] = out;
    if (replay) {
      const uint32_t r = (uint32_t)atomicAdd(sm.nrep, 1ull);
      sm.rank[s] = r;
      sm.rank_slot[r] = (uint32_t)s;
      uint64_t rec = 0;
      for (int c = 0; c < 4; c++)
        if (sm.area[c] != 0.0f) rec += n[c];
      atomicAdd(sm.nrep + 1, (unsigned long long)rec);
    }
  }
}

// Replay records: every contribution with a nonzero area to a pair that has
// a replay rank, keyed (rank << 36) | (position * 13 + offset index).
template <class T>
__global__ __launch_bounds__(WG) void k_contacts_emit(const T* __restrict__ lab, Geo g, Table tb, Sums sm,
                                                      unsigned long long* keys, unsigned long long* nkeys,
                                                      uint64_t cap_keys) {
  for (int64_t loc = (int64_t)blockIdx.x * WG + threadIdx.x; loc < g.V; loc += (int64_t)gridDim.x * WG) {
    const T cur = lab[loc];
    if (cur == T(0)) continue;
    const int64_t x = loc % g.sx, r = loc / g.sx, y = r % g.sy, z = r / g.sy;
    int64_t nb[13];
    if (g.d2) offsets2(nb, x, y, g.sx, g.n == 4);
    else offsets3(nb, x, y, z, g.sx, g.sy);
    for (int i = 0; i < g.n; i++) {
      const T v = lab[loc + nb[i]];
      if (v == T(0) || v == cur) continue;
      if (sm.area[area_class(i, g.d2)] == 0.0f) continue;
      const T lo = cur < v ? cur : v, hi = cur < v ? v : cur;
      const uint64_t s = table_slot<T>(tb, lo, hi, false);
      if (s >= tb.cap) continue;
      const uint32_t rk = sm.rank[s];
      if (rk == 0xFFFFFFFFu) continue;
      const unsigned long long at = atomicAdd(nkeys, 1ull);
      if (at < cap_keys) keys[at] = ((unsigned long long)rk << 36) | (unsigned long long)(loc * 13 + i);
    }
  }
}

__device__ inline uint64_t lower_bound_u64(const unsigned long long* a, uint64_t n, unsigned long long v) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One replayed pair per thread: its sorted records summed in raster order.
__global__ __launch_bounds__(WG) void k_contacts_replay(const unsigned long long* __restrict__ keys, uint64_t nkeys,
                                                        Sums sm, uint64_t nranks, bool d2) {
  for (uint64_t r = (uint64_t)blockIdx.x * WG + threadIdx.x; r < nranks; r += (uint64_t)gridDim.x * WG) {
    const uint64_t b = lower_bound_u64(keys, nkeys, (unsigned long long)r << 36);
    const uint64_t e = lower_bound_u64(keys, nkeys, (unsigned long long)(r + 1) << 36);
    float acc = 0.0f;
    for (uint64_t k = b; k < e; k++) {
      const int i = (int)((keys[k] & ((1ull << 36) - 1)) % 13ull);
      acc += sm.area[area_class(i, d2)];
    }
    sm.value[sm.rank_slot[r]] = acc;
  }
}

// Claimed slots into dense output arrays (any order; the caller sorts):
// labels widened to 64 bits (signed types sign-extended).
template <class T>
__global__ __launch_bounds__(WG) void k_contacts_pack(Table tb, const float* value, unsigned long long* lo,
                                                      unsigned long long* hi, float* val, unsigned long long* n,
                                                      uint64_t cap_out) {
  for (uint64_t s = (uint64_t)blockIdx.x * WG + threadIdx.x; s < tb.cap; s += (uint64_t)gridDim.x * WG) {
    const unsigned long long w = tb.k0[s];
    if (w == 0) continue;
    uint64_t a, b;
    if (sizeof(T) == 8) {
      a = w;
      b = tb.k1[s];
    } else {
      a = w >> 32;
      b = w & 0xFFFFFFFFull;
    }
    T ta, tb2;
    if constexpr (sizeof(T) == 1) {
      ta = __builtin_bit_cast(T, (uint8_t)a);
      tb2 = __builtin_bit_cast(T, (uint8_t)b);
    } else if constexpr (sizeof(T) == 2) {
      ta = __builtin_bit_cast(T, (uint16_t)a);
      tb2 = __builtin_bit_cast(T, (uint16_t)b);
    } else if constexpr (sizeof(T) == 4) {
      ta = __builtin_bit_cast(T, (uint32_t)a);
      tb2 = __builtin_bit_cast(T, (uint32_t)b);
    } else {
      ta = __builtin_bit_cast(T, a);
      tb2 = __builtin_bit_cast(T, b);
    }
    const unsigned long long at = atomicAdd(n, 1ull);
    if (at >= cap_out) continue;
    if (T(-1) < T(0)) {
      lo[at] = (unsigned long long)(long long)ta;
      hi[at] = (unsigned long long)(long long)tb2;
    } else {
      lo[at] = (unsigned long long)ta;
      hi[at] = (unsigned long long)tb2;
    }
    val[at] = value[s];
  }
}

__global__ void k_contacts_zero(unsigned long long* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) p[i] = 0;
}

static thread_local std::string g_cerr;
static int cfail(int code, const std::string& msg) {
  g_cerr = msg;
  return code;
}

#define CHK(expr)                                                                                       \
  do {                                                                                                  \
    hipError_t _e = (expr);                                                                             \
    if (_e != hipSuccess) return cfail(CC3D_AMD_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

unsigned grid_of(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + WG - 1) / WG, 1u << 16)); }

}  // namespace
}  // namespace cc3d_amd

using namespace cc3d_amd;

struct cc3d_amd_contacts {
  uint64_t n = 0;
  unsigned long long* lo = nullptr;  // device
  unsigned long long* hi = nullptr;
  float* val = nullptr;
  hipStream_t stream = nullptr;
};

namespace {

template <class T>
int contacts_run(const T* lab, const Geo& g, const float area[4], cc3d_amd_contacts* res, hipStream_t s) {
  // one word of counters for the run's scalars: used, overflow, nrep[2], npack, nkeys
  unsigned long long* sc = nullptr;
  CHK(hipMallocAsync(reinterpret_cast<void**>(&sc), 8 * sizeof(unsigned long long), s));
  uint64_t cap = 1ull << 16;
  while (cap < 4 * 1024 * 1024 && cap < (uint64_t)g.V / 4) cap <<= 1;
  Table tb{};
  unsigned long long* base = nullptr;
  for (;;) {
    const size_t words = cap * (sizeof(T) == 8 ? 6 : 5);
    CHK(hipMallocAsync(reinterpret_cast<void**>(&base), words * 8, s));
    hipLaunchKernelGGL(k_contacts_zero, dim3(grid_of(words)), dim3(WG), 0, s, base, (uint64_t)words);
    hipLaunchKernelGGL(k_contacts_zero, dim3(1), dim3(WG), 0, s, sc, (uint64_t)8);
    tb.cnt = base;
    tb.k0 = base + 4 * cap;
    tb.k1 = sizeof(T) == 8 ? base + 5 * cap : nullptr;
    tb.used = sc;
    tb.overflow = reinterpret_cast<int*>(sc + 1);
    tb.cap = cap;
    tb.limit = cap / 2;
    hipLaunchKernelGGL((k_contacts_count<T>), dim3(grid_of((uint64_t)g.V)), dim3(WG), 0, s, lab, g, tb);
    CHK(hipGetLastError());
    int ovf = 0;
    CHK(hipMemcpyAsync(&ovf, tb.overflow, sizeof(int), hipMemcpyDeviceToHost, s));
    CHK(hipStreamSynchronize(s));
    if (!ovf) break;
    CHK(hipFreeAsync(base, s));
    base = nullptr;
    cap <<= 1;
    if (cap > (1ull << 34)) return cfail(CC3D_AMD_ERR_NOMEM, "contacts: too many label pairs");
  }
  // sums
  float* value = nullptr;
  uint32_t* rank = nullptr;
  uint32_t* rank_slot = nullptr;
  CHK(hipMallocAsync(reinterpret_cast<void**>(&value), cap * sizeof(float), s));
  CHK(hipMallocAsync(reinterpret_cast<void**>(&rank), cap * sizeof(uint32_t), s));
  CHK(hipMallocAsync(reinterpret_cast<void**>(&rank_slot), cap * sizeof(uint32_t), s));
  Sums sm{};
  for (int c = 0; c < 4; c++) sm.area[c] = area[c];
  sm.value = value;
  sm.rank = rank;
  sm.rank_slot = rank_slot;
  sm.nrep = sc + 2;
  hipLaunchKernelGGL(k_contacts_sums, dim3(grid_of(cap)), dim3(WG), 0, s, tb, sm);
  CHK(hipGetLastError());
  unsigned long long nrep[2] = {0, 0};
  CHK(hipMemcpyAsync(nrep, sm.nrep, sizeof(nrep), hipMemcpyDeviceToHost, s));
  CHK(hipStreamSynchronize(s));
  if (nrep[0] > 0) {
    if (nrep[0] >= (1ull << 28)) return cfail(CC3D_AMD_ERR_NOMEM, "contacts: too many pairs to replay");
    unsigned long long* keys = nullptr;
    unsigned long long* sorted = nullptr;
    CHK(hipMallocAsync(reinterpret_cast<void**>(&keys), nrep[1] * 8, s));
    CHK(hipMallocAsync(reinterpret_cast<void**>(&sorted), nrep[1] * 8, s));
    hipLaunchKernelGGL((k_contacts_emit<T>), dim3(grid_of((uint64_t)g.V)), dim3(WG), 0, s, lab, g, tb, sm, keys,
                       sc + 5, (uint64_t)nrep[1]);
    CHK(hipGetLastError());
    size_t tmp_bytes = 0;
    CHK(rocprim::radix_sort_keys(nullptr, tmp_bytes, keys, sorted, (size_t)nrep[1], 0, 64, s));
    void* tmp = nullptr;
    CHK(hipMallocAsync(&tmp, std::max<size_t>(tmp_bytes, 8), s));
    CHK(rocprim::radix_sort_keys(tmp, tmp_bytes, keys, sorted, (size_t)nrep[1], 0, 64, s));
    hipLaunchKernelGGL(k_contacts_replay, dim3(grid_of(nrep[0])), dim3(WG), 0, s, sorted, (uint64_t)nrep[1], sm,
                       (uint64_t)nrep[0], g.d2);
    CHK(hipGetLastError());
    CHK(hipFreeAsync(tmp, s));
    CHK(hipFreeAsync(keys, s));
    CHK(hipFreeAsync(sorted, s));
  }
  // pack
  unsigned long long used = 0;
  CHK(hipMemcpyAsync(&used, tb.used, sizeof(used), hipMemcpyDeviceToHost, s));
  CHK(hipStreamSynchronize(s));
  const uint64_t n = used;
  CHK(hipMallocAsync(reinterpret_cast<void**>(&res->lo), std::max<uint64_t>(n, 1) * 8, s));
  CHK(hipMallocAsync(reinterpret_cast<void**>(&res->hi), std::max<uint64_t>(n, 1) * 8, s));
  CHK(hipMallocAsync(reinterpret_cast<void**>(&res->val), std::max<uint64_t>(n, 1) * 4, s));
  hipLaunchKernelGGL((k_contacts_pack<T>), dim3(grid_of(cap)), dim3(WG), 0, s, tb, value, res->lo, res->hi,
                     res->val, sc + 4, n);
  CHK(hipGetLastError());
  res->n = n;
  CHK(hipFreeAsync(value, s));
  CHK(hipFreeAsync(rank, s));
  CHK(hipFreeAsync(rank_slot, s));
  CHK(hipFreeAsync(base, s));
  CHK(hipFreeAsync(sc, s));
  CHK(hipStreamSynchronize(s));
  return CC3D_AMD_OK;
}

}  // namespace

extern "C" {

const char* cc3d_amd_contacts_last_error(void) { return g_cerr.c_str(); }

int cc3d_amd_contacts_run(const void* labels, int dtype, int64_t sx, int64_t sy, int64_t sz, int connectivity,
                          int surface_area, float wx, float wy, float wz, cc3d_amd_contacts** result, uint64_t* n_pairs,
                          void* stream) {
  if (!labels || !result || !n_pairs || sx <= 0 || sy <= 0 || sz <= 0)
    return cfail(CC3D_AMD_ERR_INVALID, "bad argument");
  *result = nullptr;
  *n_pairs = 0;
  Geo g;
  g.sx = sx;
  g.sy = sy;
  g.sz = sz;
  g.V = sx * sy * sz;
  float area[4];
  if (connectivity == 4 || connectivity == 8) {
    if (sz != 1) return cfail(CC3D_AMD_ERR_RUNTIME, "z thickness must be 1 for 2d region graph extraction.");
    g.d2 = true;
    area[0] = wy;
    area[1] = wx;
    area[2] = 0.0f;
    area[3] = 0.0f;
  } else if (connectivity == 6 || connectivity == 18 || connectivity == 26) {
    g.d2 = false;
    area[0] = wy * wz;
    area[1] = wx * wz;
    area[2] = wx * wy;
    area[3] = 0.0f;
  } else {
    return cfail(CC3D_AMD_ERR_RUNTIME, "Only (2d) 4, 8, or (3d) 6, 18, and 26 connectivities are supported.");
  }
  if (!surface_area) area[0] = area[1] = area[2] = area[3] = 1.0f;
  g.n = connectivity / 2;
  hipStream_t s = static_cast<hipStream_t>(stream);
  cc3d_amd_contacts* res = new cc3d_amd_contacts();
  res->stream = s;
  int rc;
  switch (dtype) {
    case CT_U8: rc = contacts_run<uint8_t>(static_cast<const uint8_t*>(labels), g, area, res, s); break;
    case CT_U16: rc = contacts_run<uint16_t>(static_cast<const uint16_t*>(labels), g, area, res, s); break;
    case CT_U32: rc = contacts_run<uint32_t>(static_cast<const uint32_t*>(labels), g, area, res, s); break;
    case CT_U64: rc = contacts_run<uint64_t>(static_cast<const uint64_t*>(labels), g, area, res, s); break;
    case CT_I8: rc = contacts_run<int8_t>(static_cast<const int8_t*>(labels), g, area, res, s); break;
    case CT_I16: rc = contacts_run<int16_t>(static_cast<const int16_t*>(labels), g, area, res, s); break;
    case CT_I32: rc = contacts_run<int32_t>(static_cast<const int32_t*>(labels), g, area, res, s); break;
    case CT_I64: rc = contacts_run<int64_t>(static_cast<const int64_t*>(labels), g, area, res, s); break;
    default: rc = cfail(CC3D_AMD_ERR_INVALID, "unsupported label dtype");
  }
  if (rc != CC3D_AMD_OK) {
    cc3d_amd_contacts_free(res);
    return rc;
  }
  *result = res;
  *n_pairs = res->n;
  return CC3D_AMD_OK;
}

int cc3d_amd_contacts_fetch(cc3d_amd_contacts* r, uint64_t* lo, uint64_t* hi, float* value) {
  if (!r) return cfail(CC3D_AMD_ERR_INVALID, "bad argument");
  if (r->n == 0) return CC3D_AMD_OK;
  if (!lo || !hi || !value) return cfail(CC3D_AMD_ERR_INVALID, "bad argument");
  CHK(hipMemcpyAsync(lo, r->lo, r->n * 8, hipMemcpyDeviceToHost, r->stream));
  CHK(hipMemcpyAsync(hi, r->hi, r->n * 8, hipMemcpyDeviceToHost, r->stream));
  CHK(hipMemcpyAsync(value, r->val, r->n * 4, hipMemcpyDeviceToHost, r->stream));
  CHK(hipStreamSynchronize(r->stream));
  return CC3D_AMD_OK;
}

void cc3d_amd_contacts_free(cc3d_amd_contacts* r) {
  if (!r) return;
  if (r->lo) (void)hipFreeAsync(r->lo, r->stream);
  if (r->hi) (void)hipFreeAsync(r->hi, r->stream);
  if (r->val) (void)hipFreeAsync(r->val, r->stream);
  (void)hipStreamSynchronize(r->stream);
  delete r;
}

}  // extern "C"
