// Global phases of the labelling pipeline and the C ABI (include/cc3d_amd.h).
//
// Pipeline for one volume (stream-ordered, no host round trip until stats):
//   k_init       reset counters, clear the root bitmask
//   [k_minmax]   continuous 2D-8 only: global min/max (cc3d/cc3d_continuous.hpp:298-303)
//   k_local      per tile: LDS union-find, local roots, J tables, requests; every
//                local root sets the bit of its key in the root bitmask
//   k_requests   per tile: cross-tile unions on the global forest GP over local
//                roots (entry tile_root0 + r per local root r), each tree rooted at
//                its smallest key; a link clears the bit of the root it buries,
//                so the bits left mark the global roots (kunion)
//   [k_periodic] wrap-around unions (cc3d/cc3d.hpp:1048-1073, :1265-1277, :1377-1418)
//   k_scan_*     exclusive popcount prefix over the bitmask -> N
//   --- stage 2 ---
//   k_finalize   per local root: its global root's key (a find on the settled
//                forest), final label = 1 + rank of that key;
//                per voxel: label of its local root, written in the output dtype
//   [k_q1_*]     binary 2D-8 only: provisional count and stale background labels
//
// Numbering: a component's label is 1 + the number of components whose
// minimum key is smaller -- the reference's first-appearance order
// (cc3d/cc3d.hpp:239-285), keys being voxel indices (or 2x2 block keys for the
// binary 2D-8 path, cc3d/cc3d_binary.hpp:1212-1231).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cc3d_amd.h"
#include "common.hpp"
#include "local.hpp"
#include "rows.hpp"

namespace cc3d_amd {

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
__global__ void k_init(Counters* c, uint32_t* bits, int64_t nwords) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    c->epl = 0;
    c->prov = 0;
    c->first_row = 0x7FFFFFFFFFFFFFFFll;
    c->last_row = -1;
    c->req_count = 0;
    c->req_overflow = 0;
    c->N = 0;
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if ((reinterpret_cast<uintptr_t>(bits) & 15) == 0) {  // 16-byte stores, then the tail
    const int64_t n4 = nwords >> 2;
    for (int64_t w = i; w < n4; w += stride) reinterpret_cast<uint4*>(bits)[w] = make_uint4(0u, 0u, 0u, 0u);
    for (int64_t w = 4 * n4 + i; w < nwords; w += stride) bits[w] = 0;
  } else {
    for (int64_t w = i; w < nwords; w += stride) bits[w] = 0;
  }
}

// Union tables (UTab in common.hpp), one copy per device, built once by
// k_build_utab: each thread grows one (border class, link mask) entry.
static __device__ uint16_t g_utab[UTAB_TOTAL];

template <int NB>
__global__ void k_build_utab(uint16_t* tab) {
  using U = UTab<NB>;
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= (U::ROWS << U::N)) return;
  const int cls = e >> U::N;
  const uint32_t lk = (uint32_t)e & ((1u << U::N) - 1u);
  const Border bd = class_border<Nbhd<NB>::IS3D>(cls);
  uint32_t conn = 0, u = 0;
  while (lk & ~conn) {
    const int i = __builtin_ctz(lk & ~conn);
    u |= 1u << i;
    uint32_t front = 1u << i;
    conn |= front;
    while (front) {
      uint32_t nb = 0;
      for (int j = 0; j < U::N; j++)
        if ((front >> j) & 1u) nb |= cover_from<NB>(j, bd);
      front = nb & lk & ~conn;
      conn |= front;
    }
  }
  tab[U::OFF + e] = (uint16_t)u;
}

// Batch: reset nc counters and clear the root bitmask (a fill kernel, so the
// sequence can be captured in a graph).
__global__ void k_batch_init(Counters* c, int nc, uint32_t* bits, int64_t nwords) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) {
    c[i].epl = 0;
    c[i].prov = 0;
    c[i].first_row = 0x7FFFFFFFFFFFFFFFll;
    c[i].last_row = -1;
    c[i].req_count = 0;
    c[i].req_overflow = 0;
    c[i].N = 0;
  }
  for (int64_t w = i; w < nwords; w += (int64_t)gridDim.x * blockDim.x) bits[w] = 0;
}

// Folds the local kernel's per-tile counters into Counters.
__global__ __launch_bounds__(WG) void k_tile_counters(const TileCnt* tc, int64_t ntiles, Counters* c) {
  __shared__ unsigned long long se[WG / 64], sp[WG / 64];
  __shared__ long long sf[WG / 64], sl[WG / 64];
  unsigned long long e = 0, pv = 0;
  long long f = 0x7FFFFFFFFFFFFFFFll, l = -1;
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < ntiles; i += (int64_t)gridDim.x * WG) {
    const TileCnt v = tc[i];
    e += v.epl;
    pv += v.prov;
    f = v.first_row < f ? v.first_row : f;
    l = v.last_row > l ? v.last_row : l;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    e += __shfl_xor(e, o);
    pv += __shfl_xor(pv, o);
    const long long f2 = __shfl_xor(f, o), l2 = __shfl_xor(l, o);
    f = f2 < f ? f2 : f;
    l = l2 > l ? l2 : l;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    se[wv] = e;
    sp[wv] = pv;
    sf[wv] = f;
    sl[wv] = l;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < WG / 64; w++) {
      e += se[w];
      pv += sp[w];
      f = sf[w] < f ? sf[w] : f;
      l = sl[w] > l ? sl[w] : l;
    }
    if (e) atomicAdd(&c->epl, e);
    if (pv) atomicAdd(&c->prov, pv);
    if (l >= 0) {
      atomicMin(&c->first_row, f);
      atomicMax(&c->last_row, l);
    }
  }
}

struct GlobalParams {
  int64_t sx, sy, sz, osx;
  int64_t ntx, nty, ntz, ntiles;
  uint32_t* GP;
  uint32_t* RL;
  const uint32_t* TR;
  const uint8_t* J8;
  const uint16_t* J16;
  const uint32_t* REQ;
  const uint32_t* TQ;
  uint32_t* bits;
  uint32_t* prefix;
  uint32_t* bsum;
  int64_t nwords;
  Counters* ctr;
  // batch of volumes (nvol > 0): per-volume table, first tile of each, and
  // the table entry of each tile
  const BatchVol* vols;
  const int64_t* vtile0;
  const uint32_t* vix;
  int nvol;
  // single volume (nvol == 0) placed at a tile and key offset: one volume of
  // a batch addressed on its own (its periodic unions)
  int64_t tile0;
  uint32_t key0;
};

// Geometry of the volume owning batch tile `slot`.
struct VolGeo {
  int64_t sx, sy, sz, osx, ntx, nty, ntz, tile0;
  uint32_t key0;
};

__device__ inline VolGeo batch_geo(const BatchVol& v) {
  VolGeo g;
  g.sx = v.sx; g.sy = v.sy; g.sz = v.sz; g.osx = v.osx;
  g.ntx = v.ntx; g.nty = v.nty; g.ntz = v.ntz;
  g.tile0 = v.tile0;
  g.key0 = v.key0;
  return g;
}

__device__ inline VolGeo vol_geo(const GlobalParams& p, int64_t slot) {
  if (p.nvol != 0) return batch_geo(p.vols[p.vix[slot]]);
  VolGeo g;
  g.sx = p.sx; g.sy = p.sy; g.sz = p.sz; g.osx = p.osx;
  g.ntx = p.ntx; g.nty = p.nty; g.ntz = p.ntz;
  g.tile0 = p.tile0;
  g.key0 = p.key0;
  return g;
}

// Forest entry of the local root of voxel (gx, gy, gz); 0xFFFFFFFF for background.
template <class G>
__device__ inline uint32_t lroot_index(const GlobalParams& p, const VolGeo& v, int64_t gx, int64_t gy, int64_t gz) {
  const int64_t tx = gx / G::TX, ty = gy / G::TY, tz = G::IS3D ? gz / G::TZ : 0;
  const int64_t slot = v.tile0 + tx + v.ntx * (ty + v.nty * tz);
  const int ji = G::jidx((int)(gx - tx * G::TX), (int)(gy - ty * G::TY), (int)(gz - tz * G::TZ));
  const uint32_t R = p.TR[slot];
  const uint32_t j = R <= 255 ? p.J8[slot * G::TV + ji] : p.J16[slot * G::TV + ji];
  if (j == 0) return 0xFFFFFFFFu;
  return tile_root0<G>(v.sx, v.sy, v.sz, tx, ty, tz, v.key0) + j - 1;
}

// First forest entry of tile `slot` of volume v.
template <class G>
__device__ inline uint32_t slot_root0(const VolGeo& v, int64_t slot) {
  const int64_t tile = slot - v.tile0;
  const int64_t tz = tile / (v.ntx * v.nty);
  const int64_t tr = tile - tz * v.ntx * v.nty;
  const int64_t ty = tr / v.ntx, tx = tr - ty * v.ntx;
  return tile_root0<G>(v.sx, v.sy, v.sz, tx, ty, tz, v.key0);
}

// Cross-tile unions: LANES lanes per tile walk that tile's requests, each
// (local root r, halo index h): the forest entry of r and that of the local
// root owning halo voxel h in its own tile.  A tile holds a few tens of
// requests, so a whole wave per tile would leave most lanes idle.
template <class G, int QK, int LANES>
__global__ __launch_bounds__(WG) void k_requests(GlobalParams p) {
  const int lane = threadIdx.x & 63, sub = lane % LANES;
  const int64_t group = ((int64_t)blockIdx.x * WG + threadIdx.x) / LANES;
  const int64_t ngroups = ((int64_t)gridDim.x * WG) / LANES;
  for (int64_t slot = group; slot < p.ntiles; slot += ngroups) {
    const uint32_t nq = p.TQ[slot];
    if (nq == 0) continue;
    const VolGeo v = vol_geo(p, slot);
    const int64_t tile = slot - v.tile0;
    const int64_t tz = tile / (v.ntx * v.nty);
    const int64_t tr = tile - tz * v.ntx * v.nty;
    const int64_t ty = tr / v.ntx, tx = tr - ty * v.ntx;
    const int64_t x0 = tx * G::TX, y0 = ty * G::TY, z0 = tz * G::TZ;
    const uint32_t root0 = tile_root0<G>(v.sx, v.sy, v.sz, tx, ty, tz, v.key0);
    for (uint32_t i = sub; i < nq; i += LANES) {
      const uint32_t rq = p.REQ[slot * G::HALO + i];
      const uint32_t a = root0 + (rq >> 16);
      int bx, by, bz;
      G::halo_coords(rq & 0xFFFF, bx, by, bz);
      const int64_t gx = x0 + bx - 1, gy = y0 + by - 1, gz = G::IS3D ? z0 + bz - 1 : 0;
      const uint32_t b = lroot_index<G>(p, v, gx, gy, gz);
      if (b != 0xFFFFFFFFu) kunion(p.GP, p.RL, p.bits, a, b);
    }
  }
}

// Periodic boundary unions.  NB: 6 (x, y, z wraps), 4 (x, y wraps) or 8 (y wrap
// with diagonals, the two cross corners, x wrap with diagonals), matching the
// pairs the reference visits.  Binary links need both voxels set; multilabel
// links need equal nonzero values.  Skipped when the reference would have taken
// its single-foreground-row shortcut (cc3d/fastcc3d.pyx:469-470), which ignores
// periodicity.
template <class T, int MODE, int NB, class G>
__global__ __launch_bounds__(WG) void k_periodic(GlobalParams p, const T* in, int special_row_rule) {
  if (special_row_rule && p.ctr->first_row == p.ctr->last_row && p.ctr->first_row >= 0) return;
  const int64_t sx = p.sx, sy = p.sy, sz = p.sz;
  int64_t n0, n1, n2;
  if (NB == 6) {
    n0 = sy * sz;
    n1 = sx * sz;
    n2 = sx * sy;
  } else if (NB == 4) {
    n0 = sx;
    n1 = sy;
    n2 = 0;
  } else {
    n0 = 3 * sx;
    n1 = 2;
    n2 = 3 * sy;
  }
  const int64_t total = n0 + n1 + n2;
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < total; i += (int64_t)gridDim.x * WG) {
    int64_t ax, ay, az = 0, bx, by, bz = 0;
    bool valid = true;
    if (NB == 6) {
      if (i < n0) {
        ay = i % sy;
        az = i / sy;
        ax = 0;
        bx = sx - 1;
        by = ay;
        bz = az;
      } else if (i < n0 + n1) {
        const int64_t k = i - n0;
        ax = k % sx;
        az = k / sx;
        ay = 0;
        bx = ax;
        by = sy - 1;
        bz = az;
      } else {
        const int64_t k = i - n0 - n1;
        ax = k % sx;
        ay = k / sx;
        az = 0;
        bx = ax;
        by = ay;
        bz = sz - 1;
      }
    } else if (NB == 4) {
      if (i < n0) {
        ax = i;
        ay = 0;
        bx = ax;
        by = sy - 1;
      } else {
        ay = i - n0;
        ax = 0;
        bx = sx - 1;
        by = ay;
      }
    } else {
      if (i < n0) {
        const int64_t x = i / 3, d = i % 3 - 1;
        ax = x;
        ay = 0;
        bx = x + d;
        by = sy - 1;
        valid = bx >= 0 && bx < sx;
      } else if (i < n0 + n1) {
        if (i == n0) {
          ax = 0;
          ay = 0;
          bx = sx - 1;
          by = sy - 1;
        } else {
          ax = sx - 1;
          ay = 0;
          bx = 0;
          by = sy - 1;
        }
      } else {
        const int64_t k = i - n0 - n1;
        const int64_t y = k / 3, d = k % 3 - 1;
        ax = 0;
        ay = y;
        bx = sx - 1;
        by = y + d;
        valid = by >= 0 && by < sy;
      }
    }
    if (!valid) continue;
    const T va = in[ax + sx * (ay + sy * az)];
    const T vb = in[bx + sx * (by + sy * bz)];
    if (!fg(va)) continue;
    const bool link = MODE == M_BINARY ? fg(vb) : (va == vb);
    if (!link) continue;
    const VolGeo v = vol_geo(p, 0);
    const uint32_t ka = lroot_index<G>(p, v, ax, ay, az), kb = lroot_index<G>(p, v, bx, by, bz);
    if (ka != 0xFFFFFFFFu && kb != 0xFFFFFFFFu) kunion(p.GP, p.RL, p.bits, ka, kb);
  }
}

// Exclusive popcount prefix over the root bitmask, 4096 words per block.
constexpr int SCAN_WPT = 16;
constexpr int SCAN_WPB = SCAN_WPT * WG;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* tmp, uint32_t& total) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) tmp[wv] = x;
  __syncthreads();
  uint32_t base = 0, tot = 0;
  for (int w = 0; w < WG / 64; w++) {
    if (w < wv) base += tmp[w];
    tot += tmp[w];
  }
  __syncthreads();
  total = tot;
  return base + x - v;
}

// SCAN_WPT consecutive words from base (zeros past nwords), with 16-byte
// loads when the whole group is in range.
__device__ inline void load_words(const uint32_t* bits, int64_t base, int64_t nwords, uint32_t* w) {
  if (base + SCAN_WPT <= nwords) {
#pragma unroll
    for (int i = 0; i < SCAN_WPT; i += 4) {
      const uint4 q = *reinterpret_cast<const uint4*>(bits + base + i);
      w[i] = q.x;
      w[i + 1] = q.y;
      w[i + 2] = q.z;
      w[i + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_WPT; i++) w[i] = base + i < nwords ? bits[base + i] : 0u;
  }
}

__global__ __launch_bounds__(WG) void k_scan_a(const uint32_t* bits, int64_t nwords, uint32_t* bsum) {
  __shared__ uint32_t tmp[WG / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_WPB + (int64_t)threadIdx.x * SCAN_WPT;
  uint32_t w[SCAN_WPT];
  load_words(bits, base, nwords, w);
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < SCAN_WPT; i++) c += __popc(w[i]);
  uint32_t total;
  block_exclusive_scan(c, tmp, total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(WG) void k_scan_b(uint32_t* bsum, int64_t nblk, Counters* ctr) {
  __shared__ uint32_t tmp[WG / 64];
  uint32_t carry = 0;
  for (int64_t off = 0; off < nblk; off += WG) {
    const int64_t i = off + threadIdx.x;
    const uint32_t v = i < nblk ? bsum[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, tmp, total);
    if (i < nblk) bsum[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0 && ctr) ctr->N = carry;
}

__global__ __launch_bounds__(WG) void k_scan_c(const uint32_t* bits, int64_t nwords, const uint32_t* bsum,
                                               uint32_t* prefix) {
  __shared__ uint32_t tmp[WG / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_WPB + (int64_t)threadIdx.x * SCAN_WPT;
  uint32_t w[SCAN_WPT];
  load_words(bits, base, nwords, w);
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < SCAN_WPT; i++) {
    w[i] = __popc(w[i]);
    c += w[i];
  }
  uint32_t total;
  uint32_t run = bsum[blockIdx.x] + block_exclusive_scan(c, tmp, total);
#pragma unroll
  for (int i = 0; i < SCAN_WPT; i++) {
    const uint32_t v = w[i];
    w[i] = run;
    run += v;
  }
  if (base + SCAN_WPT <= nwords) {  // 16-byte stores for whole groups
#pragma unroll
    for (int i = 0; i < SCAN_WPT; i += 4)
      *reinterpret_cast<uint4*>(prefix + base + i) = make_uint4(w[i], w[i + 1], w[i + 2], w[i + 3]);
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_WPT; i++)
      if (base + i < nwords) prefix[base + i] = w[i];
  }
}

// k_scan_b and k_scan_c in one launch, for up to SCAN_FUSE_MAX blocks: each
// block adds up the (L2-resident) sums of the blocks before it instead of
// waiting for a single-block pass over them; block 0 adds up all of them,
// the total.
constexpr int64_t SCAN_FUSE_MAX = 4096;
__global__ __launch_bounds__(WG) void k_scan_bc(const uint32_t* bits, int64_t nwords, const uint32_t* bsum,
                                                int64_t nblk, uint32_t* prefix, Counters* ctr) {
  __shared__ uint32_t tmp[WG / 64];
  const int64_t lim = blockIdx.x == 0 ? nblk : (int64_t)blockIdx.x;
  uint32_t part = 0;
  for (int64_t i = threadIdx.x; i < lim; i += WG) part += bsum[i];
  uint32_t sum;
  block_exclusive_scan(part, tmp, sum);
  if (blockIdx.x == 0 && threadIdx.x == 0 && ctr) ctr->N = sum;
  const uint32_t carry = blockIdx.x == 0 ? 0u : sum;
  const int64_t base = (int64_t)blockIdx.x * SCAN_WPB + (int64_t)threadIdx.x * SCAN_WPT;
  uint32_t w[SCAN_WPT];
  load_words(bits, base, nwords, w);
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < SCAN_WPT; i++) {
    w[i] = __popc(w[i]);
    c += w[i];
  }
  uint32_t total;
  uint32_t run = carry + block_exclusive_scan(c, tmp, total);
#pragma unroll
  for (int i = 0; i < SCAN_WPT; i++) {
    const uint32_t v = w[i];
    w[i] = run;
    run += v;
  }
  if (base + SCAN_WPT <= nwords) {  // 16-byte stores for whole groups
#pragma unroll
    for (int i = 0; i < SCAN_WPT; i += 4)
      *reinterpret_cast<uint4*>(prefix + base + i) = make_uint4(w[i], w[i + 1], w[i + 2], w[i + 3]);
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_WPT; i++)
      if (base + i < nwords) prefix[base + i] = w[i];
  }
}

// Exclusive popcount prefix of nwords bitmask words (nblk = blocks of
// SCAN_WPB words); the total goes to ctr->N when ctr is given.
static void launch_scan(const uint32_t* bits, int64_t nwords, uint32_t* bsum, int64_t nblk, uint32_t* prefix,
                        Counters* ctr, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_a, dim3((unsigned)nblk), dim3(WG), 0, s, bits, nwords, bsum);
  if (nblk <= SCAN_FUSE_MAX) {
    hipLaunchKernelGGL(k_scan_bc, dim3((unsigned)nblk), dim3(WG), 0, s, bits, nwords, bsum, nblk, prefix, ctr);
  } else {
    hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(WG), 0, s, bsum, nblk, ctr);
    hipLaunchKernelGGL(k_scan_c, dim3((unsigned)nblk), dim3(WG), 0, s, bits, nwords, bsum, prefix);
  }
}

// Stage 2: every voxel of the tile gets its local root's label.  Thread t
// owns the column the local kernel gave it: its 8 root indices are one 8- or
// 16-byte load, issued before the tile's label table is built in LDS (one
// entry per local root, from its global root's key), so the per-voxel
// lookups hit LDS instead of scattered global reads.
// Final label of a local root whose global root has key r: 1 + the rank of r
// among the global roots' keys, the bits left set in the bitmask (prefix: the
// bitmask's exclusive popcount prefix per word).  base: the prefix at the
// volume's first key word (a batch numbers each volume from 1).
__device__ inline uint32_t root_label(const GlobalParams& p, uint32_t r, uint32_t base) {
  const uint32_t wd = r >> 5;
  return p.prefix[wd] - base + __popc(p.bits[wd] & ((1u << (r & 31)) - 1u)) + 1u;
}

// J entries finalized per thread (a tile's entries over one workgroup), and
// tiles per workgroup: all loads of a workgroup's tiles (root indices and
// label tables) are issued before any store, so each thread keeps 16 entries
// in flight.
template <class G>
__host__ __device__ constexpr int fin_ept() {
  return G::TV / WG;
}
template <class G>
__host__ __device__ constexpr int fin_tiles() {
  return fin_ept<G>() >= 16 ? 1 : 16 / fin_ept<G>();
}

// Thread t finalizes the 8-voxel columns t, t + WG, ... of the tile (J
// entries 8c..8c+7 of column c, the local kernel's thread-major order), so
// for each voxel of a column the workgroup's stores cover whole x-rows.
// fin_load: the root indices of those entries (vector loads) and the tile's
// label table: entry 0 background, entry 1 + i the label of local root i.
template <class G>
__device__ inline void fin_load(const GlobalParams& p, int64_t slot, uint32_t root0, uint32_t base,
                                uint32_t (&j)[fin_ept<G>()], uint32_t* lab) {
  constexpr int EPT = fin_ept<G>();
  static_assert(EPT * WG == G::TV && G::PER_THREAD == 8 && EPT % 8 == 0,
                "each thread finalizes whole 8-voxel columns");
  const int t = threadIdx.x;
  const uint32_t R = p.TR[slot];
  if (R <= 255) {
#pragma unroll
    for (int h = 0; h < EPT / 8; h++) {
      const uint2 w = reinterpret_cast<const uint2*>(p.J8 + slot * G::TV)[t + WG * h];
#pragma unroll
      for (int k = 0; k < 8; k++) j[8 * h + k] = ((k < 4 ? w.x : w.y) >> (8 * (k & 3))) & 0xFFu;
    }
  } else {
#pragma unroll
    for (int h = 0; h < EPT / 8; h++) {
      const uint4 w = reinterpret_cast<const uint4*>(p.J16 + slot * G::TV)[t + WG * h];
      const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int k = 0; k < 8; k++) j[8 * h + k] = (ws[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
    }
  }
  lab[0] = 0u;
  for (uint32_t i = t; i < R; i += WG) lab[i + 1] = root_label(p, kroot_key(p.GP, p.RL, root0 + i), base);
}

// Thread t's EPT output voxels of tile `tile` of volume v (after a barrier
// behind fin_load).
template <class OUT, class G>
__device__ inline void fin_store(const VolGeo& v, int64_t tile, OUT* __restrict__ out,
                                 const uint32_t (&j)[fin_ept<G>()], const uint32_t* lab) {
  constexpr int EPT = fin_ept<G>();
  const int t = threadIdx.x;
  const int64_t tz = tile / (v.ntx * v.nty);
  const int64_t tr = tile - tz * v.ntx * v.nty;
  const int64_t ty = tr / v.ntx, tx = tr - ty * v.ntx;
#pragma unroll
  for (int k = 0; k < EPT; k++) {
    int vx, vy, vz;
    G::entry_voxel(8 * (t + WG * (k >> 3)) + (k & 7), vx, vy, vz);
    const int64_t gx = tx * G::TX + vx, gy = ty * G::TY + vy, gz = tz * G::TZ + vz;
    if (gx < v.sx && gy < v.sy && gz < v.sz) out[gx + v.sx * (gy + v.sy * gz)] = (OUT)lab[j[k]];
  }
}

template <class OUT, class G>
__global__ __launch_bounds__(WG) void k_finalize(GlobalParams p, OUT* __restrict__ out) {
  constexpr int FT = fin_tiles<G>();
  __shared__ uint32_t lab[FT][G::TV + 1];
  const int64_t ngroups = (p.ntiles + FT - 1) / FT;
  const int64_t first = xcd_tile(blockIdx.x, ngroups) * FT;
  const VolGeo v = vol_geo(p, 0);
  uint32_t j[FT][fin_ept<G>()];
#pragma unroll
  for (int n = 0; n < FT; n++) {
    if (first + n >= p.ntiles) break;
    fin_load<G>(p, first + n, slot_root0<G>(v, first + n), 0u, j[n], lab[n]);
  }
  __syncthreads();
#pragma unroll
  for (int n = 0; n < FT; n++) {
    if (first + n >= p.ntiles) break;
    fin_store<OUT, G>(v, first + n, out, j[n], lab[n]);
  }
}

// Batch: fin_tiles consecutive tiles per workgroup, each writing its own
// volume's output in that volume's dtype, numbered from the prefix at the
// volume's first key word.
template <class G>
__global__ __launch_bounds__(WG) void k_finalize_batch(GlobalParams p) {
  constexpr int FT = fin_tiles<G>();
  __shared__ uint32_t lab[FT][G::TV + 1];
  const int64_t ngroups = (p.ntiles + FT - 1) / FT;
  const int64_t first = xcd_tile(blockIdx.x, ngroups) * FT;
  uint32_t j[FT][fin_ept<G>()];
  int vi[FT];
#pragma unroll
  for (int n = 0; n < FT; n++) {
    const int64_t slot = first + n;
    if (slot >= p.ntiles) break;
    vi[n] = (int)p.vix[slot];
    fin_load<G>(p, slot, slot_root0<G>(batch_geo(p.vols[vi[n]]), slot), p.prefix[p.vols[vi[n]].key0 >> 5], j[n],
                lab[n]);
  }
  __syncthreads();
#pragma unroll
  for (int n = 0; n < FT; n++) {
    const int64_t slot = first + n;
    if (slot >= p.ntiles) break;
    const BatchVol& bv = p.vols[vi[n]];
    const VolGeo v = batch_geo(bv);
    const int64_t tile = slot - bv.tile0;
    if (bv.out_dtype == DT_U16) fin_store<uint16_t, G>(v, tile, static_cast<uint16_t*>(bv.out), j[n], lab[n]);
    else if (bv.out_dtype == DT_U64) fin_store<uint64_t, G>(v, tile, static_cast<uint64_t*>(bv.out), j[n], lab[n]);
    else fin_store<uint32_t, G>(v, tile, static_cast<uint32_t*>(bv.out), j[n], lab[n]);
  }
}

// Exclusive popcount prefix of a root bitmask (also used by the graph
// colouring in color.hip); the total goes to ctr->N.
hipError_t scan_root_bits(const uint32_t* bits, int64_t nwords, uint32_t* bsum, uint32_t* prefix, Counters* ctr,
                          hipStream_t s) {
  const int64_t nblk = (nwords + SCAN_WPB - 1) / SCAN_WPB;
  launch_scan(bits, nwords, bsum, nblk, prefix, ctr, s);
  return hipGetLastError();
}

// Row tiles (GeoR): one workgroup per tile; thread t owns J entries
// 16t .. 16t + 15, i.e. 16 consecutive voxels of one row, loaded with one
// (J8) or two (J16) 16-byte loads and written as one contiguous run.
constexpr int RFIN_PT = GeoR::TV / WG;
static_assert(RFIN_PT == 16, "16 voxels per thread");

template <class OUT>
__device__ inline void rfin_tile(const GlobalParams& p, const VolGeo& v, int64_t tile, int64_t slot, uint32_t base,
                                 OUT* __restrict__ out, uint32_t* lab) {
  const int t = threadIdx.x;
  const uint32_t R = p.TR[slot];
  uint32_t j[RFIN_PT];
  if (R <= 255) {
    const uint4 w = reinterpret_cast<const uint4*>(p.J8 + slot * GeoR::TV)[t];
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int k = 0; k < RFIN_PT; k++) j[k] = (ws[k >> 2] >> (8 * (k & 3))) & 0xFFu;
  } else {
    const uint4 w0 = reinterpret_cast<const uint4*>(p.J16 + slot * GeoR::TV)[2 * t];
    const uint4 w1 = reinterpret_cast<const uint4*>(p.J16 + slot * GeoR::TV)[2 * t + 1];
    const uint32_t ws[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
    for (int k = 0; k < RFIN_PT; k++) j[k] = (ws[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
  }
  const uint32_t root0 = slot_root0<GeoR>(v, slot);
  if (t == 0) lab[0] = 0u;
  for (uint32_t i = t; i < R; i += WG) lab[i + 1] = root_label(p, kroot_key(p.GP, p.RL, root0 + i), base);
  __syncthreads();
  const int e0 = RFIN_PT * t;
  const int lx = e0 & 63, ly = (e0 >> 6) & (GeoR::TY - 1), lz = e0 >> 9;
  const int64_t tz = tile / (v.ntx * v.nty);
  const int64_t tr = tile - tz * v.ntx * v.nty;
  const int64_t ty = tr / v.ntx, tx = tr - ty * v.ntx;
  const int64_t gx = tx * GeoR::TX + lx, gy = ty * GeoR::TY + ly, gz = tz * GeoR::TZ + lz;
  if (gy >= v.sy || gz >= v.sz || gx >= v.sx) return;
  OUT* o = out + gx + v.sx * (gy + v.sy * gz);
  OUT vals[RFIN_PT];
#pragma unroll
  for (int k = 0; k < RFIN_PT; k++) vals[k] = (OUT)lab[j[k]];
  if (gx + RFIN_PT <= v.sx && (reinterpret_cast<uintptr_t>(o) & 15) == 0) {
    constexpr int NV = RFIN_PT * (int)sizeof(OUT) / 16;
#pragma unroll
    for (int i = 0; i < NV; i++) {
      uint4 w;
      __builtin_memcpy(&w, reinterpret_cast<const char*>(vals) + 16 * i, 16);
      reinterpret_cast<uint4*>(o)[i] = w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < RFIN_PT; k++)
      if (gx + k < v.sx) o[k] = vals[k];
  }
}

template <class OUT>
__global__ __launch_bounds__(WG) void k_rfinalize(GlobalParams p, OUT* __restrict__ out) {
  __shared__ uint32_t lab[GeoR::TV + 1];
  const int64_t tile = xcd_tile(blockIdx.x, p.ntiles);
  rfin_tile<OUT>(p, vol_geo(p, 0), tile, tile, 0u, out, lab);
}

__global__ __launch_bounds__(WG) void k_rfinalize_batch(GlobalParams p) {
  __shared__ uint32_t lab[GeoR::TV + 1];
  const int64_t slot = xcd_tile(blockIdx.x, p.ntiles);
  const BatchVol& bv = p.vols[p.vix[slot]];
  const VolGeo v = batch_geo(bv);
  const uint32_t base = p.prefix[bv.key0 >> 5];
  if (bv.out_dtype == DT_U16) rfin_tile<uint16_t>(p, v, slot - bv.tile0, slot, base, static_cast<uint16_t*>(bv.out), lab);
  else if (bv.out_dtype == DT_U64) rfin_tile<uint64_t>(p, v, slot - bv.tile0, slot, base, static_cast<uint64_t*>(bv.out), lab);
  else rfin_tile<uint32_t>(p, v, slot - bv.tile0, slot, base, static_cast<uint32_t*>(bv.out), lab);
}

// Batch: the table entry of every tile (one block per volume).
__global__ __launch_bounds__(WG) void k_fill_vix(const int64_t* vtile0, uint32_t* vix) {
  const int v = blockIdx.x;
  for (int64_t i = vtile0[v] + threadIdx.x; i < vtile0[v + 1]; i += WG) vix[i] = (uint32_t)v;
}

// Batch: per-volume counters from the per-tile records (one block per volume).
__global__ __launch_bounds__(WG) void k_tile_counters_batch(const TileCnt* tc, const int64_t* vtile0, Counters* ctrs) {
  __shared__ unsigned long long se[WG / 64], sp[WG / 64];
  __shared__ long long sf[WG / 64], sl[WG / 64];
  const int v = blockIdx.x;
  unsigned long long e = 0, pv = 0;
  long long f = 0x7FFFFFFFFFFFFFFFll, l = -1;
  for (int64_t i = vtile0[v] + threadIdx.x; i < vtile0[v + 1]; i += WG) {
    const TileCnt c = tc[i];
    e += c.epl;
    pv += c.prov;
    f = c.first_row < f ? c.first_row : f;
    l = c.last_row > l ? c.last_row : l;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    e += __shfl_xor(e, o);
    pv += __shfl_xor(pv, o);
    const long long f2 = __shfl_xor(f, o), l2 = __shfl_xor(l, o);
    f = f2 < f ? f2 : f;
    l = l2 > l ? l2 : l;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    se[wv] = e;
    sp[wv] = pv;
    sf[wv] = f;
    sl[wv] = l;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < WG / 64; w++) {
      e += se[w];
      pv += sp[w];
      f = sf[w] < f ? sf[w] : f;
      l = sl[w] > l ? sl[w] : l;
    }
    Counters& c = ctrs[v];
    c.epl = e;
    c.prov = pv;
    c.first_row = f;
    c.last_row = l;
  }
}

// Batch: N of every volume from the prefix at its first key word.
__global__ void k_batch_N(const BatchVol* vols, int n, const uint32_t* prefix, int64_t nwords, const Counters* total,
                          Counters* ctrs) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int64_t w0 = vols[v].key0 >> 5;
  const int64_t w1 = v + 1 < n ? (int64_t)(vols[v + 1].key0 >> 5) : nwords;
  const uint64_t hi = w1 < nwords ? prefix[w1] : total->N;
  ctrs[v].N = hi - prefix[w0];
}

// estimate_provisional_label_count (cc3d/cc3d.hpp:287-315): one wave per row.
template <class T>
__global__ __launch_bounds__(WG) void k_epl(const T* in, int64_t sx, int64_t rows, Counters* c) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * WG + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * WG) >> 6;
  unsigned long long tot = 0;
  long long first = 0x7FFFFFFFFFFFFFFFll, last = -1;
  for (int64_t row = wave; row < rows; row += nwaves) {
    const T* r = in + row * sx;
    unsigned long long cnt = 0;
    for (int64_t x = lane; x < sx; x += 64) {
      const T v = r[x];
      if (fg(v) && (x == 0 || v != r[x - 1])) cnt++;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    tot += cnt;
    if (cnt) {
      first = row < first ? row : first;
      last = row > last ? row : last;
    }
  }
  if (lane == 0) {
    if (tot) atomicAdd(&c->epl, tot);
    if (last >= 0) {
      atomicMin(&c->first_row, first);
      atomicMax(&c->last_row, last);
    }
  }
}

// Global min/max with the reference's sequential fold semantics
// (cc3d/cc3d_continuous.hpp:298-303): std::max(a, b) = a < b ? b : a folds a
// NaN in and lets the next value replace it, so the result is the plain
// min/max of the elements after the last NaN (NaN if the last element is NaN).
template <class T>
__global__ __launch_bounds__(WG) void k_lastnan(const T* in, int64_t n, long long* lastnan) {
  long long m = -1;
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG)
    if (in[i] != in[i]) m = i;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const long long y = __shfl_xor(m, o);
    m = y > m ? y : m;
  }
  if ((threadIdx.x & 63) == 0 && m >= 0) atomicMax(lastnan, m);
}

template <class T>
__device__ inline uint64_t ordkey(T v) {
  if constexpr (sizeof(T) == 4 && T(-1) < T(0)) {
    uint32_t b;
    __builtin_memcpy(&b, &v, 4);
    return (b & 0x80000000u) ? (uint64_t)(~b) : (uint64_t)(b | 0x80000000u);
  } else if constexpr (sizeof(T) == 8 && T(-1) < T(0)) {
    uint64_t b;
    __builtin_memcpy(&b, &v, 8);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
  } else {
    return (uint64_t)v;
  }
}
template <class T>
__host__ __device__ inline uint64_t unordkey(uint64_t k) {
  if constexpr (sizeof(T) == 4 && T(-1) < T(0)) {
    const uint32_t b = (k & 0x80000000u) ? (uint32_t)(k & 0x7FFFFFFFu) : (uint32_t)(~k);
    return b;
  } else if constexpr (sizeof(T) == 8 && T(-1) < T(0)) {
    return (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  } else {
    return k;
  }
}

// mm[0] = min key, mm[1] = max key over indices > *lastnan
template <class T>
__global__ __launch_bounds__(WG) void k_minmax(const T* in, int64_t n, const long long* lastnan,
                                               unsigned long long* mm) {
  const int64_t start = *lastnan + 1;
  unsigned long long lo = ~0ull, hi = 0;
  for (int64_t i = start + (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) {
    const unsigned long long k = ordkey(in[i]);
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(mm, lo);
    atomicMax(mm + 1, hi);
  }
}

template <class T>
__global__ void k_minmax_final(const long long* lastnan, int64_t n, unsigned long long* mm) {
  // converts ordered keys back to bit patterns; a trailing NaN makes both NaN
  uint64_t lo, hi;
  if (*lastnan == n - 1) {
    lo = hi = (sizeof(T) == 4) ? 0x7FC00000ull : 0x7FF8000000000000ull;
  } else {
    lo = unordkey<T>(mm[0]);
    hi = unordkey<T>(mm[1]);
  }
  mm[0] = lo;
  mm[1] = hi;
}

// ---------------------------------------------------------------------------
// binary 2D-8 (2x2 block) provisional labels: cc3d/cc3d_binary.hpp:1029-1161
// ---------------------------------------------------------------------------
// Source block of the block whose top-left pixel is (x, y): the block index
// the reference copies its provisional label from, -1 for a new label, -2 for
// a block without foreground.
template <class T>
__device__ inline int64_t q1_source(const T* in, int64_t x, int64_t y, int64_t sx, int64_t sy) {
  const int64_t osx = (sx + 1) >> 1;
  const int64_t ob = (x >> 1) + osx * (y >> 1);
  auto px = [&](int64_t xx, int64_t yy) -> bool { return fg(in[xx + sx * yy]); };
  const bool Y = px(x, y);
  const bool Z = x < sx - 1 && px(x + 1, y);
  const bool X = y < sy - 1 && px(x, y + 1);
  const bool W = x < sx - 1 && y < sy - 1 && px(x + 1, y + 1);
  const int64_t up = ob - osx, left = ob - 1, upleft = ob - 1 - osx, upright = ob + 1 - osx;
  if (Y) {
    if (y > 0 && px(x, y - 1)) return up;
    if (y > 0 && x < sx - 1 && px(x + 1, y - 1)) return up;
    if (x > 0 && px(x - 1, y)) return left;
    if (x > 0 && y > 0 && px(x - 1, y - 1)) return upleft;
    if (x > 0 && y < sy - 1 && px(x - 1, y + 1)) return left;
    if (x < sx - 2 && y > 0 && Z && px(x + 2, y - 1)) return upright;
    return -1;
  }
  if (Z) {
    if (y > 0 && px(x + 1, y - 1)) return up;
    if (y > 0 && px(x, y - 1)) return up;
    if (x < sx - 2 && y > 0 && px(x + 2, y - 1)) return upright;
    if (X) {
      if (x > 0 && px(x - 1, y)) return left;
      if (x > 0 && px(x - 1, y + 1)) return left;
    }
    return -1;
  }
  if (X) {
    if (x > 0 && px(x - 1, y)) return left;
    if (x > 0 && px(x - 1, y + 1)) return left;
    return -1;
  }
  return W ? -1 : -2;
}

// Marks blocks taking a new provisional label in `bits` and counts them.
template <class T>
__global__ __launch_bounds__(WG) void k_q1_newblocks(const T* in, int64_t sx, int64_t sy, uint32_t* bits,
                                                     Counters* c) {
  const int64_t osx = (sx + 1) >> 1, osy = (sy + 1) >> 1, nb = osx * osy;
  unsigned long long cnt = 0;
  for (int64_t b = (int64_t)blockIdx.x * WG + threadIdx.x; b < nb; b += (int64_t)gridDim.x * WG) {
    const int64_t by = b / osx, bx = b - by * osx;
    if (q1_source(in, bx << 1, by << 1, sx, sy) == -1) {
      atomicOr(bits + (b >> 5), 1u << (b & 31));
      cnt++;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(&c->prov, cnt);
}

// Stale background labels of the reference's in-place block expansion
// (cc3d/cc3d_binary.hpp:1233-1260): rows with y*sx >= voxels/4 are rewritten
// only on [xstart, xend) of rows y and y+1; pixels left of xstart or right of
// xend whose index lies inside the block scratch (< osx*osy) keep the
// provisional label of the block with that index.
template <class T, class OUT>
__global__ __launch_bounds__(WG) void k_q1_stale(const T* in, int64_t sx, int64_t sy, const uint32_t* bits,
                                                 const uint32_t* prefix, OUT* out) {
  const int64_t osx = (sx + 1) >> 1, osy = (sy + 1) >> 1, nb = osx * osy;
  const int64_t V = sx * sy;
  const int64_t target = V >> 2;
  const int64_t ystart = (target + sx - 1) / sx;  // first row with y*sx >= target
  const int64_t lo = ystart * sx;
  for (int64_t loc = lo + (int64_t)blockIdx.x * WG + threadIdx.x; loc < nb; loc += (int64_t)gridDim.x * WG) {
    const int64_t y = loc / sx, x = loc - y * sx;
    const int64_t ymax = y < sy - 1 ? y + 1 : y;
    // xstart/xend of compute_foreground_index (cc3d/cc3d.hpp:317-342) for y and ymax
    int64_t s0 = 0, e0 = 0, s1 = 0, e1 = 0;
    for (int64_t xx = 0; xx < sx; xx++)
      if (fg(in[xx + sx * y])) { s0 = xx; break; }
    for (int64_t xx = sx - 1; xx >= s0; xx--)
      if (fg(in[xx + sx * y])) { e0 = xx + 1; break; }
    for (int64_t xx = 0; xx < sx; xx++)
      if (fg(in[xx + sx * ymax])) { s1 = xx; break; }
    for (int64_t xx = sx - 1; xx >= s1; xx--)
      if (fg(in[xx + sx * ymax])) { e1 = xx + 1; break; }
    const int64_t xs = (s0 < s1 ? s0 : s1) & ~(int64_t)1;
    const int64_t xe = e0 > e1 ? e0 : e1;
    if (x >= xs && x < xe) continue;
    // provisional label of block `loc`: follow copies back to a new-label block
    const int64_t bby = loc / osx, bbx = loc - bby * osx;
    int64_t src = q1_source(in, bbx << 1, bby << 1, sx, sy);
    if (src == -2) continue;  // no foreground: the scratch held 0
    int64_t b = loc;
    while (src >= 0) {
      b = src;
      const int64_t yy = b / osx, xx = b - yy * osx;
      src = q1_source(in, xx << 1, yy << 1, sx, sy);
    }
    const uint32_t w = (uint32_t)(b >> 5);
    const uint32_t lab = prefix[w] + __popc(bits[w] & ((1u << (b & 31)) - 1u)) + 1u;
    out[loc] = (OUT)lab;
  }
}

__global__ void k_zero_words(uint32_t* w, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    w[i] = 0;
}

template <class OUT>
__global__ void k_zero_out(OUT* out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = OUT(0);
}

__global__ void k_mm_init(unsigned long long* mm) {
  mm[0] = ~0ull;
  mm[1] = 0ull;
  reinterpret_cast<long long*>(mm)[2] = -1;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHECK(expr)                                                                                  \
  do {                                                                                                  \
    hipError_t _e = (expr);                                                                             \
    if (_e != hipSuccess) return fail(CC3D_AMD_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

// Device address of the current device's union tables; the first call on a
// device builds them on a private stream and waits for it.
static std::mutex g_utab_mu;
static std::map<int, const uint16_t*> g_utab_dev;

// on: build on this stream (and wait for it) instead of a private one.
static int utab_ready(const uint16_t** out, hipStream_t on = nullptr) {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_utab_mu);
  auto it = g_utab_dev.find(dev);
  if (it != g_utab_dev.end()) {
    *out = it->second;
    return CC3D_AMD_OK;
  }
  void* ptr = nullptr;
  HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(g_utab)));
  uint16_t* tab = static_cast<uint16_t*>(ptr);
  hipStream_t s = on;
  if (!on) HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipLaunchKernelGGL(k_build_utab<26>, dim3((UTab<26>::ROWS << UTab<26>::N) / 256), dim3(256), 0, s, tab);
  hipLaunchKernelGGL(k_build_utab<18>, dim3((UTab<18>::ROWS << UTab<18>::N) / 256), dim3(256), 0, s, tab);
  hipLaunchKernelGGL(k_build_utab<8>, dim3(1), dim3(256), 0, s, tab);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (!on) (void)hipStreamDestroy(s);  // the build status above is what matters
  if (e != hipSuccess) return fail(CC3D_AMD_ERR_HIP, std::string("building the union tables: ") + hipGetErrorString(e));
  g_utab_dev[dev] = tab;
  *out = tab;
  return CC3D_AMD_OK;
}

// Scratch of the no-allocation entry point: one copy per device, held until
// the process ends.  Host code never passes the array's name to a kernel
// (in host code it is only a stand-in for the device address): every use goes
// through the address hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) unsigned char scratch[CC3D_AMD_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static std::map<int, char*> g_scratch_dev;

static int scratch_ready(char** out) {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_dev.find(dev);
  if (it != g_scratch_dev.end()) {
    *out = it->second;
    return CC3D_AMD_OK;
  }
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(CC3D_AMD_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
  g_scratch_dev[dev] = static_cast<char*>(ptr);
  *out = static_cast<char*>(ptr);
  return CC3D_AMD_OK;
}

// Optional per-kernel timing with HIP events on the launch stream (enabled by
// cc3d_amd_profile_enable; read back by cc3d_amd_profile_collect).
struct ProfRec {
  hipEvent_t a, b;
  const char* name;
};
static std::vector<ProfRec> g_prof_recs;
static std::vector<hipEvent_t> g_prof_pool;
static int g_prof_on = 0;  // 0 off, 1 every kernel, 2 the local kernel only
static thread_local bool g_prof_suspended = false;  // inside a graph capture
static std::mutex g_prof_mu;

static hipEvent_t prof_event() {
  if (!g_prof_pool.empty()) {
    hipEvent_t e = g_prof_pool.back();
    g_prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

struct ProfScope {
  hipEvent_t a = nullptr;
  hipStream_t s;
  const char* name;
  ProfScope(const char* n, hipStream_t st) : s(st), name(n) {
    if (!g_prof_on || g_prof_suspended) return;
    if (g_prof_on == 2 && std::strcmp(n, "k_local") != 0) return;  // dominant kernel only
    std::lock_guard<std::mutex> g(g_prof_mu);
    a = prof_event();
    (void)hipEventRecord(a, s);
  }
  ~ProfScope() {
    if (!a) return;
    std::lock_guard<std::mutex> g(g_prof_mu);
    hipEvent_t b = prof_event();
    (void)hipEventRecord(b, s);
    g_prof_recs.push_back({a, b, name});
  }
};

static inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Pinned host buffers for the jobs' counter read-back, reused across calls
// (pinned allocation and release cost far more than a small labelling call).
static std::mutex g_pin_mu;
static std::vector<Counters*> g_pin_pool;

static Counters* pinned_counters() {
  {
    std::lock_guard<std::mutex> g(g_pin_mu);
    if (!g_pin_pool.empty()) {
      Counters* c = g_pin_pool.back();
      g_pin_pool.pop_back();
      return c;
    }
  }
  Counters* c = nullptr;
  if (hipHostMalloc(reinterpret_cast<void**>(&c), sizeof(Counters), hipHostMallocDefault) != hipSuccess) return nullptr;
  return c;
}

static void release_pinned_counters(Counters* c) {
  if (!c) return;
  std::lock_guard<std::mutex> g(g_pin_mu);
  if (g_pin_pool.size() < 64) g_pin_pool.push_back(c);
  else (void)hipHostFree(c);
}
static inline unsigned grid_for(int64_t items, int64_t per_block, int64_t cap) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((items + per_block - 1) / per_block, cap));
}

struct Geometry {
  int64_t sx = 0, sy = 0, sz = 0, V = 0, K = 0, osx = 0;
  int nb = 0, qk = Q_NONE, mode = M_MULTI;
  bool is3d = true;
  bool rows = false;  // row tiles (GeoR): 3D multilabel and binary
  bool tall = false;  // tall voxel tiles (Geo3T) instead of Geo3
  int64_t ntx = 0, nty = 0, ntz = 0, ntiles = 0;
  int TV = 0, HALO = 0;
  int64_t nwords = 0, nblk = 0;
  int64_t qblocks = 0, qwords = 0, qnblk = 0;  // binary 2D-8 block bitmask
  size_t oGP, oRL, oTR, oTQ, oTC, oJ8, oJ16, oREQ, oBITS, oPRE, oBSUM, oCTR, oMM, oQB, oQP, oQS, total;
};

// tiles: CC3D_AMD_TILES_ROWS / _VOXELS / _TALL bits (else the default family)
static int plan(Geometry& g, int64_t sx, int64_t sy, int64_t sz, int64_t connectivity, int binary,
                int continuous, int tiles = 0) {
  if (sx <= 0 || sy <= 0 || sz <= 0) return fail(CC3D_AMD_ERR_INVALID, "volume must be non-empty");
  if (connectivity != 26 && connectivity != 18 && connectivity != 6 && connectivity != 8 && connectivity != 4)
    return fail(CC3D_AMD_ERR_RUNTIME, "Only 4 and 8 2D and 6, 18, and 26 3D connectivities are supported.");
  if ((connectivity == 8 || connectivity == 4) && sz != 1)
    return fail(CC3D_AMD_ERR_RUNTIME, "sz must be 1 for 2D connectivities.");
  g.sx = sx;
  g.sy = sy;
  g.sz = sz;
  g.V = sx * sy * sz;
  g.mode = binary ? M_BINARY : (continuous ? M_CONT : M_MULTI);
  g.osx = (sx + 1) >> 1;
  g.is3d = sz > 1;
  g.qk = Q_NONE;
  if (g.is3d) {
    g.nb = (int)connectivity;
  } else {
    // a single plane: 26/18 reduce to the 8-neighbourhood, 6 to the 4-neighbourhood
    g.nb = (connectivity == 26 || connectivity == 18) ? 8 : connectivity == 6 ? 4 : (int)connectivity;
    if (connectivity == 8 && g.mode == M_BINARY) g.qk = Q_BLOCK8;
    if (connectivity == 8 && g.mode == M_CONT) g.qk = Q_CONT8;
    if (connectivity == 4 && g.mode == M_BINARY) g.qk = Q_BIN4;
  }
  g.K = g.qk == Q_BLOCK8 ? 4 * g.osx * ((sy + 1) >> 1) : g.V;
  if (g.K >= 0xFFFFFFFFll || g.V >= 0xFFFFFFFFll)
    return fail(CC3D_AMD_ERR_INVALID, "volumes of 2^32-1 voxels or more are not supported");
  int TX, TY, TZ;
  // Row tiles for 6-connected 3D volumes (multilabel or binary): measured
  // faster there than the voxel tiles; the voxel tiles stay faster at 26/18.
  g.rows = g.is3d && g.mode != M_CONT &&
           ((tiles & CC3D_AMD_TILES_ROWS)                            ? true
            : (tiles & (CC3D_AMD_TILES_VOXELS | CC3D_AMD_TILES_TALL)) ? false
                                                                      : connectivity == 6);
  g.tall = g.is3d && !g.rows && (tiles & CC3D_AMD_TILES_TALL) && !(tiles & CC3D_AMD_TILES_VOXELS);
  if (g.rows) {
    TX = GeoR::TX; TY = GeoR::TY; TZ = GeoR::TZ; g.TV = GeoR::TV; g.HALO = GeoR::HALO;
  } else if (g.tall) {
    TX = Geo3T::TX; TY = Geo3T::TY; TZ = Geo3T::TZ; g.TV = Geo3T::TV; g.HALO = Geo3T::HALO;
  } else if (g.is3d) {
    TX = Geo3::TX; TY = Geo3::TY; TZ = Geo3::TZ; g.TV = Geo3::TV; g.HALO = Geo3::HALO;
  } else {
    TX = Geo2::TX; TY = Geo2::TY; TZ = 1; g.TV = Geo2::TV; g.HALO = Geo2::HALO;
  }
  g.ntx = (sx + TX - 1) / TX;
  g.nty = (sy + TY - 1) / TY;
  g.ntz = (sz + TZ - 1) / TZ;
  g.ntiles = g.ntx * g.nty * g.ntz;
  g.nwords = (g.K + 31) / 32;
  g.nblk = (g.nwords + SCAN_WPB - 1) / SCAN_WPB;
  if (g.qk == Q_BLOCK8) {
    g.qblocks = g.osx * ((sy + 1) >> 1);
    g.qwords = (g.qblocks + 31) / 32;
    g.qnblk = (g.qwords + SCAN_WPB - 1) / SCAN_WPB;
  }
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t r = o;
    o += align256(bytes);
    return r;
  };
  const size_t slots = (size_t)g.ntiles * g.TV;
  g.oGP = take((size_t)g.K * 4);
  g.oRL = take((size_t)g.K * 4);
  g.oTR = take((size_t)g.ntiles * 4);
  g.oTQ = take((size_t)g.ntiles * 4);
  g.oTC = take((size_t)g.ntiles * sizeof(TileCnt));
  g.oJ8 = take(slots);
  g.oJ16 = take(slots * 2);
  g.oREQ = take((size_t)g.ntiles * g.HALO * 4);
  g.oBITS = take((size_t)g.nwords * 4);
  g.oPRE = take((size_t)g.nwords * 4);
  g.oBSUM = take((size_t)(g.nblk + 1) * 4);
  g.oCTR = take(sizeof(Counters));
  g.oMM = take(64);
  g.oQB = take((size_t)g.qwords * 4);
  g.oQP = take((size_t)g.qwords * 4);
  g.oQS = take((size_t)(g.qnblk + 1) * 4);
  g.total = o;
  return CC3D_AMD_OK;
}

}  // namespace cc3d_amd

using namespace cc3d_amd;

struct cc3d_amd_job {
  Geometry g;
  int dt = 0;
  uint64_t delta_bits = 0;
  const void* in = nullptr;
  hipStream_t stream = nullptr;
  char* ws = nullptr;
  bool own_ws = false;
  int periodic = 0;
  int special_row_rule = 0;
  bool count_prov = false;  // binary 2D-8: count provisional block labels
  bool stale = false;       // binary 2D-8 with an odd side: stale background labels
  uint32_t* qbits = nullptr;
  uint32_t* qprefix = nullptr;
  uint32_t* qbsum = nullptr;
  Counters* h_ctr = nullptr;
  Counters h_ctr_local{};  // host copy of the counters when nothing may be allocated
  bool ctr_local = false;
  bool have_stats = false;
  GlobalParams gp{};
};

struct cc3d_amd_batch {
  int n = 0, dt = 0, mode = M_MULTI;
  hipStream_t stream = nullptr;
  char* ws = nullptr;
  bool own_ws = false;
  size_t total = 0;
  int64_t ntiles = 0, nkeys = 0, nwords = 0, nblk = 0;
  size_t oVOL = 0, oGP = 0, oRL = 0, oBITS = 0, oPRE = 0, oBSUM = 0, oCTR = 0;
  BatchVol* h_tab = nullptr;  // pinned host copy of the volume table
  int64_t* h_t0 = nullptr;    // pinned: first tile of each volume, per tile group (n + tile groups)
  Counters* h_ctr = nullptr;  // pinned: entry 0 whole batch, 1 + k table entry k
  // The table holds the volumes grouped by connectivity (one local launch
  // per group, 26 and 18 first); perm[k] is the caller's index of table
  // entry k.  The volumes of a tile group share one tile geometry and their
  // per-tile arrays: 26- and 18-connected volumes the voxel tiles (Geo3),
  // 6-connected ones the row tiles (GeoR).  Tiles are numbered from 0 within
  // their tile group; the key space, bitmask and counters are shared.
  struct Group {
    int nb, tg;
    int64_t tile0, tiles;  // within tile group tg
  };
  // tile families of a batch (one geometry and set of per-tile arrays each)
  enum { FAM_VOXELS = 0, FAM_TALL = 1, FAM_ROWS = 2, FAM_PLANE = 3 };
  struct TileGroup {
    int fam = FAM_VOXELS;
    bool rows() const { return fam == FAM_ROWS; }
    int v0 = 0, v1 = 0;  // table entries
    int64_t tiles = 0;
    size_t oT0 = 0, oVIX = 0, oTR = 0, oTQ = 0, oTC = 0, oJ8 = 0, oJ16 = 0, oREQ = 0;
    GlobalParams gp{};
    LocalParams lp{};
  };
  // volumes with a periodic boundary: table entry, wrap neighbourhood (6, 4
  // or 8, as the single-volume path picks it) and tile group
  struct Periodic {
    int k, nbp, tg;
  };
  std::vector<int> perm;
  std::vector<Group> groups;
  std::vector<TileGroup> tgroups;
  std::vector<Periodic> periodic;
  // continuous 2D-8 volumes: table entry and offset of their {min, max} block
  struct MinMax {
    int k;
    size_t oMM;
  };
  std::vector<MinMax> minmax;
  // binary 2D-8 volumes with an odd side: table entry, their block bitmask,
  // prefix and block sums (the reference's stale background labels)
  struct Stale {
    int k;
    int64_t qblocks, qwords, qnblk;
    size_t oQB, oQP, oQS;
  };
  std::vector<Stale> stale;
  bool delta_set = false;
  bool outputs_set = false, capturing = false;
  GlobalParams gp{};  // whole batch: the key space, bitmask, counters
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

namespace {

// k_requests over `ntiles` tiles.  Binary volumes usually hold components
// that span most tiles, whose unions all meet at one root: there 4 lanes per
// tile contend less than 8 (measured on 512^3 uint8 Bernoulli volumes:
// p = 0.5 26-connected 0.42-0.47 -> 0.14-0.16 ms, p = 0.5 6-connected 0.30 ->
// 0.21 ms, p = 0.3 26-connected 0.20-0.23 -> 0.19-0.21 ms), while the many
// separate components of multilabel volumes keep more requests in flight with
// 8 (512^3 uint32 Voronoi 26-connected: 0.110 ms with 8 lanes, 0.132 with 4).
template <class G, int QK>
void launch_requests(const GlobalParams& p, int64_t ntiles, int mode, hipStream_t s) {
  if (mode == M_BINARY)
    hipLaunchKernelGGL((k_requests<G, QK, 4>), dim3(grid_for(ntiles, WG / 4, 65536)), dim3(WG), 0, s, p);
  else hipLaunchKernelGGL((k_requests<G, QK, 8>), dim3(grid_for(ntiles, WG / 8, 65536)), dim3(WG), 0, s, p);
}

template <class G>
int run_global(cc3d_amd_job* j) {
  const GlobalParams& p = j->gp;
  hipStream_t s = j->stream;
  {
  ProfScope ps_("k_requests", s);
  if (j->g.qk == Q_BLOCK8) launch_requests<G, Q_BLOCK8>(p, j->g.ntiles, j->g.mode, s);
  else launch_requests<G, Q_NONE>(p, j->g.ntiles, j->g.mode, s);
  }
  HIPCHECK(hipGetLastError());
  if (j->periodic && j->g.mode != M_CONT) {
    // a 6-connected single plane wraps like the 4-neighbourhood (its z wrap pairs a voxel with itself)
    const int nbp = j->g.is3d ? 6 : j->g.nb;
    const int64_t npairs = nbp == 6 ? (j->g.sy * j->g.sz + j->g.sx * j->g.sz + j->g.sx * j->g.sy)
                           : nbp == 4 ? (j->g.sx + j->g.sy) : (3 * j->g.sx + 2 + 3 * j->g.sy);
    const dim3 grid(grid_for(npairs, WG, 65536));
    const int sr = j->special_row_rule;
    ProfScope ps_("k_periodic", s);
#define CC3D_P(T)                                                                                             \
  do {                                                                                                        \
    const T* in = static_cast<const T*>(j->in);                                                               \
    if (j->g.mode == M_BINARY) {                                                                              \
      if (nbp == 6) hipLaunchKernelGGL((k_periodic<T, M_BINARY, 6, G>), grid, dim3(WG), 0, s, p, in, 0);     \
      else if (nbp == 4) hipLaunchKernelGGL((k_periodic<T, M_BINARY, 4, G>), grid, dim3(WG), 0, s, p, in, 0); \
      else hipLaunchKernelGGL((k_periodic<T, M_BINARY, 8, G>), grid, dim3(WG), 0, s, p, in, 0);              \
    } else {                                                                                                  \
      if (nbp == 6) hipLaunchKernelGGL((k_periodic<T, M_MULTI, 6, G>), grid, dim3(WG), 0, s, p, in, sr);     \
      else if (nbp == 4) hipLaunchKernelGGL((k_periodic<T, M_MULTI, 4, G>), grid, dim3(WG), 0, s, p, in, sr); \
      else hipLaunchKernelGGL((k_periodic<T, M_MULTI, 8, G>), grid, dim3(WG), 0, s, p, in, sr);              \
    }                                                                                                         \
  } while (0)
    switch (j->dt) {
      case DT_U8: CC3D_P(uint8_t); break;
      case DT_U16: CC3D_P(uint16_t); break;
      case DT_U32: CC3D_P(uint32_t); break;
      case DT_U64: CC3D_P(uint64_t); break;
      case DT_F32: CC3D_P(float); break;
      case DT_F64: CC3D_P(double); break;
    }
#undef CC3D_P
    HIPCHECK(hipGetLastError());
  }
  {
  ProfScope ps_("k_scan", s);
  launch_scan(p.bits, p.nwords, p.bsum, (int64_t)j->g.nblk, p.prefix, p.ctr, s);
  }
  HIPCHECK(hipGetLastError());
  return CC3D_AMD_OK;
}

template <class T>
int launch_minmax_on(const T* in, int64_t V, unsigned long long* mm, hipStream_t s) {
  long long* lastnan = reinterpret_cast<long long*>(mm + 2);
  const dim3 grid(grid_for(V, WG, 4096));
  hipLaunchKernelGGL(k_mm_init, dim3(1), dim3(1), 0, s, mm);
  if (T(-1) < T(0)) hipLaunchKernelGGL((k_lastnan<T>), grid, dim3(WG), 0, s, in, V, lastnan);
  hipLaunchKernelGGL((k_minmax<T>), grid, dim3(WG), 0, s, in, V, lastnan, mm);
  hipLaunchKernelGGL((k_minmax_final<T>), dim3(1), dim3(1), 0, s, lastnan, V, mm);
  HIPCHECK(hipGetLastError());
  return CC3D_AMD_OK;
}
template <class T>
int launch_minmax(cc3d_amd_job* j, unsigned long long* mm) {
  return launch_minmax_on<T>(static_cast<const T*>(j->in), j->g.V, mm, j->stream);
}

template <class T>
int launch_q1(cc3d_amd_job* j) {
  const Geometry& g = j->g;
  hipStream_t s = j->stream;
  hipLaunchKernelGGL(k_zero_words, dim3(grid_for(g.qwords, WG, 8192)), dim3(WG), 0, s, j->qbits, g.qwords);
  hipLaunchKernelGGL((k_q1_newblocks<T>), dim3(grid_for(g.qblocks, WG, 65536)), dim3(WG), 0, s,
                     static_cast<const T*>(j->in), g.sx, g.sy, j->qbits, j->gp.ctr);
  HIPCHECK(hipGetLastError());
  if (j->stale) {
    launch_scan(j->qbits, g.qwords, j->qbsum, g.qnblk, j->qprefix, nullptr, s);
    HIPCHECK(hipGetLastError());
  }
  return CC3D_AMD_OK;
}

template <class T, class OUT>
int launch_stale(cc3d_amd_job* j, OUT* out) {
  const Geometry& g = j->g;
  const int64_t n = g.qblocks - ((g.V >> 2) / g.sx) * g.sx;
  hipLaunchKernelGGL((k_q1_stale<T, OUT>), dim3(grid_for(std::max<int64_t>(n, 1), WG, 4096)), dim3(WG), 0, j->stream,
                     static_cast<const T*>(j->in), g.sx, g.sy, j->qbits, j->qprefix, out);
  HIPCHECK(hipGetLastError());
  return CC3D_AMD_OK;
}

int delta_bits_of(int dt, const void* delta, uint64_t& bits) {
  bits = 0;
  if (!delta) return CC3D_AMD_OK;
  const size_t sz = dt == DT_U8 ? 1 : dt == DT_U16 ? 2 : (dt == DT_U32 || dt == DT_F32) ? 4 : 8;
  std::memcpy(&bits, delta, sz);
  return CC3D_AMD_OK;
}

bool delta_is_zero(int dt, uint64_t bits) {
  if (dt == DT_F32) return from_bits<float>(bits) == 0.0f;
  if (dt == DT_F64) return from_bits<double>(bits) == 0.0;
  return bits == 0;
}

// Continuous 2D-8 volumes of a batch: each volume's global {min, max}
// (cc3d/cc3d_continuous.hpp:298-303), before the local kernels read it.
template <class T>
hipError_t batch_minmax(const cc3d_amd_batch* b, hipStream_t s) {
  for (const auto& m : b->minmax) {
    const BatchVol& v = b->h_tab[m.k];
    if (launch_minmax_on<T>(static_cast<const T*>(v.in), v.sx * v.sy * v.sz,
                            reinterpret_cast<unsigned long long*>(b->ws + m.oMM), s) != CC3D_AMD_OK)
      return hipErrorLaunchFailure;
  }
  return hipGetLastError();
}

// Binary 2D-8 volumes with an odd side: each volume's new-label blocks and
// their prefix, as the single-volume path's stage 1 forms them (the block
// count also goes to the volume's provisional counter there).
template <class T>
hipError_t batch_q1(const cc3d_amd_batch* b, hipStream_t s) {
  for (const auto& st : b->stale) {
    const BatchVol& v = b->h_tab[st.k];
    uint32_t* qb = reinterpret_cast<uint32_t*>(b->ws + st.oQB);
    uint32_t* qp = reinterpret_cast<uint32_t*>(b->ws + st.oQP);
    uint32_t* qs = reinterpret_cast<uint32_t*>(b->ws + st.oQS);
    hipLaunchKernelGGL(k_zero_words, dim3(grid_for(st.qwords, WG, 8192)), dim3(WG), 0, s, qb, st.qwords);
    hipLaunchKernelGGL((k_q1_newblocks<T>), dim3(grid_for(st.qblocks, WG, 65536)), dim3(WG), 0, s,
                       static_cast<const T*>(v.in), v.sx, v.sy, qb, b->gp.ctr + 1 + st.k);
    launch_scan(qb, st.qwords, qs, st.qnblk, qp, nullptr, s);
  }
  return hipGetLastError();
}

// Their stale background labels, after the volume's output is written.
template <class T, class OUT>
void batch_stale_one(const cc3d_amd_batch* b, const cc3d_amd_batch::Stale& st, hipStream_t s) {
  const BatchVol& v = b->h_tab[st.k];
  const int64_t V = v.sx * v.sy;
  const int64_t n = st.qblocks - ((V >> 2) / v.sx) * v.sx;
  hipLaunchKernelGGL((k_q1_stale<T, OUT>), dim3(grid_for(std::max<int64_t>(n, 1), WG, 4096)), dim3(WG), 0, s,
                     static_cast<const T*>(v.in), v.sx, v.sy, reinterpret_cast<const uint32_t*>(b->ws + st.oQB),
                     reinterpret_cast<const uint32_t*>(b->ws + st.oQP), static_cast<OUT*>(v.out));
}
template <class T>
hipError_t batch_stale(const cc3d_amd_batch* b, hipStream_t s) {
  for (const auto& st : b->stale) {
    const int od = b->h_tab[st.k].out_dtype;
    if (od == DT_U16) batch_stale_one<T, uint16_t>(b, st, s);
    else if (od == DT_U64) batch_stale_one<T, uint64_t>(b, st, s);
    else batch_stale_one<T, uint32_t>(b, st, s);
  }
  return hipGetLastError();
}

// Wrap-around unions of one periodic volume of a batch: the single-volume
// kernel on that volume's tiles and keys (table entry k, tile group tg).
template <class G>
hipError_t batch_periodic(const cc3d_amd_batch* b, const cc3d_amd_batch::Periodic& pv, hipStream_t s) {
  const BatchVol& v = b->h_tab[pv.k];
  const auto& tg = b->tgroups[pv.tg];
  GlobalParams q = tg.gp;
  q.nvol = 0;
  q.vols = nullptr;
  q.vix = nullptr;
  q.sx = v.sx;
  q.sy = v.sy;
  q.sz = v.sz;
  q.osx = v.osx;
  q.ntx = v.ntx;
  q.nty = v.nty;
  q.ntz = v.ntz;
  q.tile0 = v.tile0;
  q.key0 = v.key0;
  q.ctr = b->gp.ctr + 1 + pv.k;
  const int64_t npairs = pv.nbp == 6 ? (v.sy * v.sz + v.sx * v.sz + v.sx * v.sy)
                         : pv.nbp == 4 ? (v.sx + v.sy) : (3 * v.sx + 2 + 3 * v.sy);
  const dim3 grid(grid_for(npairs, WG, 65536));
  // a multilabel volume whose foreground is one row skips its wrap unions, as
  // cc3d_amd_label_begin does by default (the volume's own counters decide)
  const int sr = b->mode == M_MULTI;
#define CC3D_PB(T, M)                                                                                                \
  do {                                                                                                               \
    const T* in = static_cast<const T*>(v.in);                                                                       \
    if (pv.nbp == 6) hipLaunchKernelGGL((k_periodic<T, M, 6, G>), grid, dim3(WG), 0, s, q, in, sr);                  \
    else if (pv.nbp == 4) hipLaunchKernelGGL((k_periodic<T, M, 4, G>), grid, dim3(WG), 0, s, q, in, sr);             \
    else hipLaunchKernelGGL((k_periodic<T, M, 8, G>), grid, dim3(WG), 0, s, q, in, sr);                              \
  } while (0)
#define CC3D_PBT(T)                           \
  do {                                        \
    if (b->mode == M_BINARY) CC3D_PB(T, M_BINARY); \
    else CC3D_PB(T, M_MULTI);                 \
  } while (0)
  switch (b->dt) {
    case DT_U8: CC3D_PBT(uint8_t); break;
    case DT_U16: CC3D_PBT(uint16_t); break;
    case DT_U32: CC3D_PBT(uint32_t); break;
    case DT_U64: CC3D_PBT(uint64_t); break;
    case DT_F32: CC3D_PBT(float); break;
    case DT_F64: CC3D_PBT(double); break;
    default: return hipErrorInvalidValue;
  }
#undef CC3D_PBT
#undef CC3D_PB
  return hipGetLastError();
}

}  // namespace

extern "C" {

const char* cc3d_amd_last_error(void) { return g_err.c_str(); }
const char* cc3d_amd_version(void) { return "cc3d_amd 0.1.0 (gfx950)"; }

size_t cc3d_amd_workspace_bytes(int64_t sx, int64_t sy, int64_t sz, int64_t connectivity, int binary_image,
                                int flags) {
  Geometry g;
  if (plan(g, sx, sy, sz, connectivity, binary_image, (flags & CC3D_AMD_CONTINUOUS) != 0, flags) != CC3D_AMD_OK)
    return 0;
  return g.total;
}

// Stage 1 into the job object j (constructed by the caller, who releases it
// with job_release on failure too).  no_alloc: the caller's workspace is used
// and nothing is allocated -- counters are read back through the job's own
// host copy, and the union tables (first use on a device) are built on the
// job's stream.
}  // extern "C"

static int job_begin(cc3d_amd_job* j, const void* in, int in_dtype, int64_t sx, int64_t sy, int64_t sz,
                     int64_t connectivity, const void* delta, int flags, int periodic_boundary, int binary_image,
                     void* workspace, size_t workspace_bytes, void* stream, bool no_alloc) {
  if (in_dtype < DT_U8 || in_dtype > DT_F64) return fail(CC3D_AMD_ERR_INVALID, "unsupported input dtype");
  if (!in) return fail(CC3D_AMD_ERR_INVALID, "input pointer is NULL");
  const bool continuous = (flags & CC3D_AMD_CONTINUOUS) && !binary_image;
  int rc = plan(j->g, sx, sy, sz, connectivity, binary_image, continuous, flags);
  if (rc != CC3D_AMD_OK) return rc;
  if (periodic_boundary && j->g.mode == M_CONT)
    return fail(CC3D_AMD_ERR_RUNTIME, "periodic_boundary is not currently supported for continuous data.");
  const Geometry& g = j->g;
  j->dt = in_dtype;
  j->in = in;
  j->stream = static_cast<hipStream_t>(stream);
  j->periodic = periodic_boundary && (connectivity == 6 || connectivity == 4 || connectivity == 8);
  j->special_row_rule = (g.mode == M_MULTI) && !(flags & CC3D_AMD_NO_SINGLE_ROW_RULE);
  j->count_prov = g.qk == Q_BLOCK8 && (flags & CC3D_AMD_COUNT_PROVISIONAL);
  j->stale = g.qk == Q_BLOCK8 && ((g.sx & 1) || (g.sy & 1));
  delta_bits_of(in_dtype, delta, j->delta_bits);
  if (workspace) {
    if (workspace_bytes < g.total) return fail(CC3D_AMD_ERR_INVALID, "workspace too small");
    if (reinterpret_cast<uintptr_t>(workspace) & 255)
      return fail(CC3D_AMD_ERR_INVALID, "workspace must be 256-byte aligned");
    j->ws = static_cast<char*>(workspace);
  } else {
    if (no_alloc) return fail(CC3D_AMD_ERR_INVALID, "no workspace");
    void* p = nullptr;
    if (hipMallocAsync(&p, g.total, j->stream) != hipSuccess || !p)
      return fail(CC3D_AMD_ERR_NOMEM, "could not allocate workspace");
    j->ws = static_cast<char*>(p);
    j->own_ws = true;
  }
  if (no_alloc) {
    j->h_ctr = &j->h_ctr_local;
    j->ctr_local = true;
  } else {
    j->h_ctr = pinned_counters();
    if (!j->h_ctr) return fail(CC3D_AMD_ERR_NOMEM, "could not allocate pinned counters");
  }
  char* ws = j->ws;
  GlobalParams& gp = j->gp;
  gp.sx = g.sx; gp.sy = g.sy; gp.sz = g.sz; gp.osx = g.osx;
  gp.ntx = g.ntx; gp.nty = g.nty; gp.ntz = g.ntz; gp.ntiles = g.ntiles;
  gp.GP = reinterpret_cast<uint32_t*>(ws + g.oGP);
  gp.RL = reinterpret_cast<uint32_t*>(ws + g.oRL);
  gp.TR = reinterpret_cast<uint32_t*>(ws + g.oTR);
  gp.J8 = reinterpret_cast<uint8_t*>(ws + g.oJ8);
  gp.J16 = reinterpret_cast<uint16_t*>(ws + g.oJ16);
  gp.REQ = reinterpret_cast<uint32_t*>(ws + g.oREQ);
  gp.TQ = reinterpret_cast<uint32_t*>(ws + g.oTQ);
  gp.bits = reinterpret_cast<uint32_t*>(ws + g.oBITS);
  gp.prefix = reinterpret_cast<uint32_t*>(ws + g.oPRE);
  gp.bsum = reinterpret_cast<uint32_t*>(ws + g.oBSUM);
  gp.nwords = g.nwords;
  gp.ctr = reinterpret_cast<Counters*>(ws + g.oCTR);
  j->qbits = reinterpret_cast<uint32_t*>(ws + g.oQB);
  j->qprefix = reinterpret_cast<uint32_t*>(ws + g.oQP);
  j->qbsum = reinterpret_cast<uint32_t*>(ws + g.oQS);
  unsigned long long* mm = reinterpret_cast<unsigned long long*>(ws + g.oMM);
  hipStream_t s = j->stream;

  {
    ProfScope ps_("k_init", s);
    hipLaunchKernelGGL(k_init, dim3(grid_for((g.nwords + 3) / 4, WG, 8192)), dim3(WG), 0, s, gp.ctr, gp.bits, g.nwords);
  }
  if (hipGetLastError() != hipSuccess) return fail(CC3D_AMD_ERR_HIP, "k_init launch failed");

#define CC3D_DT_SWITCH(FN)                      \
  switch (in_dtype) {                           \
    case DT_U8: rc = FN<uint8_t>; break;        \
    case DT_U16: rc = FN<uint16_t>; break;      \
    case DT_U32: rc = FN<uint32_t>; break;      \
    case DT_U64: rc = FN<uint64_t>; break;      \
    case DT_F32: rc = FN<float>; break;         \
    default: rc = FN<double>; break;            \
  }
  if (g.qk == Q_CONT8) {
#define CC3D_MM launch_minmax
    switch (in_dtype) {
      case DT_U8: rc = launch_minmax<uint8_t>(j, mm); break;
      case DT_U16: rc = launch_minmax<uint16_t>(j, mm); break;
      case DT_U32: rc = launch_minmax<uint32_t>(j, mm); break;
      case DT_U64: rc = launch_minmax<uint64_t>(j, mm); break;
      case DT_F32: rc = launch_minmax<float>(j, mm); break;
      default: rc = launch_minmax<double>(j, mm); break;
    }
#undef CC3D_MM
    if (rc != CC3D_AMD_OK) return rc;
  }
#undef CC3D_DT_SWITCH

  LocalParams lp;
  lp.utab = nullptr;
  if (!g.rows) {  // the row kernels need no union tables
    rc = utab_ready(&lp.utab, no_alloc ? s : nullptr);
    if (rc != CC3D_AMD_OK) return rc;
  }
  lp.in = in;
  lp.sx = g.sx; lp.sy = g.sy; lp.sz = g.sz; lp.osx = g.osx;
  lp.ntx = g.ntx; lp.nty = g.nty; lp.ntz = g.ntz; lp.ntiles = g.ntiles;
  lp.delta_bits = j->delta_bits;
  lp.minmax = reinterpret_cast<const uint64_t*>(mm);
  lp.GP = gp.GP;
  lp.RL = gp.RL;
  lp.bits = gp.bits;
  lp.TR = reinterpret_cast<uint32_t*>(ws + g.oTR);
  lp.J8 = reinterpret_cast<uint8_t*>(ws + g.oJ8);
  lp.J16 = reinterpret_cast<uint16_t*>(ws + g.oJ16);
  lp.REQ = reinterpret_cast<uint32_t*>(ws + g.oREQ);
  lp.TQ = reinterpret_cast<uint32_t*>(ws + g.oTQ);
  lp.tcnt = reinterpret_cast<TileCnt*>(ws + g.oTC);
  hipError_t e;
  {
  ProfScope ps_("k_local", s);
  if (g.rows && g.mode == M_MULTI) e = launch_rows_multi(in_dtype, g.nb, lp, s);
  else if (g.rows) e = launch_rows_binary(in_dtype, g.nb, lp, s);
  else if (g.mode == M_MULTI) e = launch_local_multi(in_dtype, g.nb, g.qk, g.tall, lp, s);
  else if (g.mode == M_BINARY) e = launch_local_binary(in_dtype, g.nb, g.qk, g.tall, lp, s);
  else e = launch_local_cont(in_dtype, g.nb, g.qk, g.tall, lp, s);
  }
  if (e != hipSuccess) return fail(CC3D_AMD_ERR_HIP, std::string("k_local: ") + hipGetErrorString(e));
  {
    ProfScope ps_("k_tile_counters", s);
    hipLaunchKernelGGL(k_tile_counters, dim3(grid_for(g.ntiles, WG * 2, 256)), dim3(WG), 0, s, lp.tcnt, g.ntiles,
                       gp.ctr);
  }
  if (hipGetLastError() != hipSuccess) return fail(CC3D_AMD_ERR_HIP, "k_tile_counters launch failed");
  rc = g.rows ? run_global<GeoR>(j)
       : g.tall ? run_global<Geo3T>(j)
       : g.is3d ? run_global<Geo3>(j)
                : run_global<Geo2>(j);
  if (rc == CC3D_AMD_OK && (j->count_prov || j->stale)) {
    switch (in_dtype) {
      case DT_U8: rc = launch_q1<uint8_t>(j); break;
      case DT_U16: rc = launch_q1<uint16_t>(j); break;
      case DT_U32: rc = launch_q1<uint32_t>(j); break;
      case DT_U64: rc = launch_q1<uint64_t>(j); break;
      case DT_F32: rc = launch_q1<float>(j); break;
      default: rc = launch_q1<double>(j); break;
    }
  }
  return rc;
}

extern "C" {

int cc3d_amd_label_begin(cc3d_amd_job** out_job, const void* in, int in_dtype, int64_t sx, int64_t sy,
                         int64_t sz, int64_t connectivity, const void* delta, int flags, int periodic_boundary,
                         int binary_image, void* workspace, size_t workspace_bytes, void* stream) {
  if (!out_job) return fail(CC3D_AMD_ERR_INVALID, "job pointer is NULL");
  *out_job = nullptr;
  cc3d_amd_job* j = new cc3d_amd_job();
  const int rc = job_begin(j, in, in_dtype, sx, sy, sz, connectivity, delta, flags, periodic_boundary, binary_image,
                           workspace, workspace_bytes, stream, false);
  if (rc != CC3D_AMD_OK) {
    cc3d_amd_job_free(j);
    return rc;
  }
  *out_job = j;
  return CC3D_AMD_OK;
}

int cc3d_amd_job_stats(cc3d_amd_job* j, cc3d_amd_stats* st) {
  if (!j || !st) return fail(CC3D_AMD_ERR_INVALID, "NULL job or stats");
  if (!j->have_stats) {
    HIPCHECK(hipMemcpyAsync(j->h_ctr, j->gp.ctr, sizeof(Counters), hipMemcpyDeviceToHost, j->stream));
    HIPCHECK(hipStreamSynchronize(j->stream));
    j->have_stats = true;
  }
  st->N = j->h_ctr->N;
  st->provisional = j->h_ctr->prov;
  st->epl = j->h_ctr->epl;
  st->first_foreground_row = j->h_ctr->last_row < 0 ? -1 : j->h_ctr->first_row;
  st->last_foreground_row = j->h_ctr->last_row;
  return CC3D_AMD_OK;
}

int cc3d_amd_label_finish(cc3d_amd_job* j, void* out, int out_dtype) {
  if (!j || !out) return fail(CC3D_AMD_ERR_INVALID, "NULL job or output");
  const Geometry& g = j->g;
  hipStream_t s = j->stream;
  const dim3 grid3((unsigned)((g.ntiles + fin_tiles<Geo3>() - 1) / fin_tiles<Geo3>()));
  const dim3 grid3t((unsigned)((g.ntiles + fin_tiles<Geo3T>() - 1) / fin_tiles<Geo3T>()));
  const dim3 grid2((unsigned)((g.ntiles + fin_tiles<Geo2>() - 1) / fin_tiles<Geo2>())), block(WG);
  int rc = CC3D_AMD_OK;
  ProfScope ps_("k_finalize", s);
#define CC3D_F(OUT)                                                                                \
  do {                                                                                             \
    if (g.rows) hipLaunchKernelGGL((k_rfinalize<OUT>), dim3((unsigned)g.ntiles), block, 0, s, j->gp, (OUT*)out); \
    else if (g.tall) hipLaunchKernelGGL((k_finalize<OUT, Geo3T>), grid3t, block, 0, s, j->gp, (OUT*)out); \
    else if (g.is3d) hipLaunchKernelGGL((k_finalize<OUT, Geo3>), grid3, block, 0, s, j->gp, (OUT*)out); \
    else hipLaunchKernelGGL((k_finalize<OUT, Geo2>), grid2, block, 0, s, j->gp, (OUT*)out);        \
    if (j->stale) {                                                                                \
      switch (j->dt) {                                                                             \
        case DT_U8: rc = launch_stale<uint8_t, OUT>(j, (OUT*)out); break;                          \
        case DT_U16: rc = launch_stale<uint16_t, OUT>(j, (OUT*)out); break;                        \
        case DT_U32: rc = launch_stale<uint32_t, OUT>(j, (OUT*)out); break;                        \
        case DT_U64: rc = launch_stale<uint64_t, OUT>(j, (OUT*)out); break;                        \
        case DT_F32: rc = launch_stale<float, OUT>(j, (OUT*)out); break;                           \
        default: rc = launch_stale<double, OUT>(j, (OUT*)out); break;                              \
      }                                                                                            \
    }                                                                                              \
  } while (0)
  switch (out_dtype) {
    case DT_U16: CC3D_F(uint16_t); break;
    case DT_U32: CC3D_F(uint32_t); break;
    case DT_U64: CC3D_F(uint64_t); break;
    default: return fail(CC3D_AMD_ERR_INVALID, "output dtype must be uint16, uint32 or uint64");
  }
#undef CC3D_F
  if (rc != CC3D_AMD_OK) return rc;
  HIPCHECK(hipGetLastError());
  return CC3D_AMD_OK;
}

void cc3d_amd_job_free(cc3d_amd_job* j) {
  if (!j) return;
  if (j->own_ws && j->ws) (void)hipFreeAsync(j->ws, j->stream);
  if (!j->ctr_local) release_pinned_counters(j->h_ctr);
  delete j;
}

int cc3d_amd_estimate_provisional_label_count(const void* in, int in_dtype, int64_t sx, int64_t voxels,
                                              uint64_t* epl, int64_t* first_row, int64_t* last_row,
                                              void* stream) {
  if (!in || !epl || !first_row || !last_row) return fail(CC3D_AMD_ERR_INVALID, "NULL argument");
  if (sx <= 0 || voxels <= 0 || voxels % sx) return fail(CC3D_AMD_ERR_INVALID, "bad geometry");
  if (in_dtype < DT_U8 || in_dtype > DT_F64) return fail(CC3D_AMD_ERR_INVALID, "unsupported input dtype");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Counters* d = nullptr;
  Counters* h = nullptr;
  HIPCHECK(hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(Counters), s));
  hipLaunchKernelGGL(k_init, dim3(1), dim3(WG), 0, s, d, (uint32_t*)nullptr, (int64_t)0);
  const int64_t rows = voxels / sx;
  const dim3 grid(grid_for(rows, 4, 16384));
  switch (in_dtype) {
    case DT_U8: hipLaunchKernelGGL((k_epl<uint8_t>), grid, dim3(WG), 0, s, (const uint8_t*)in, sx, rows, d); break;
    case DT_U16: hipLaunchKernelGGL((k_epl<uint16_t>), grid, dim3(WG), 0, s, (const uint16_t*)in, sx, rows, d); break;
    case DT_U32: hipLaunchKernelGGL((k_epl<uint32_t>), grid, dim3(WG), 0, s, (const uint32_t*)in, sx, rows, d); break;
    case DT_U64: hipLaunchKernelGGL((k_epl<uint64_t>), grid, dim3(WG), 0, s, (const uint64_t*)in, sx, rows, d); break;
    case DT_F32: hipLaunchKernelGGL((k_epl<float>), grid, dim3(WG), 0, s, (const float*)in, sx, rows, d); break;
    default: hipLaunchKernelGGL((k_epl<double>), grid, dim3(WG), 0, s, (const double*)in, sx, rows, d); break;
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&h), sizeof(Counters), hipHostMallocDefault));
  HIPCHECK(hipMemcpyAsync(h, d, sizeof(Counters), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  *epl = h->epl;
  *first_row = h->last_row < 0 ? -1 : h->first_row;
  *last_row = h->last_row;
  (void)hipHostFree(h);
  (void)hipFreeAsync(d, s);
  return CC3D_AMD_OK;
}

// Union-find length the reference's first pass allocates for these arguments
// (cc3d/cc3d.hpp:360-366 and its 18/6/2D twins; cc3d/cc3d_binary.hpp:47-51,
// :350-352, :527-530, :794-797, :964-967; cc3d/cc3d_continuous.hpp:105-107,
// :224-226, :289-291), or 0 when that pass returns before labelling.
uint64_t cc3d_amd_union_find_length(int64_t max_labels, int64_t voxels, int64_t connectivity, int binary_image,
                                    int out_dtype) {
  const uint64_t omax = out_dtype == DT_U16 ? 0xFFFFull : out_dtype == DT_U32 ? 0xFFFFFFFFull : ~0ull;
  const uint64_t V = (uint64_t)voxels;
  uint64_t ml = (uint64_t)max_labels + 1;
  if (binary_image && connectivity == 8) {
    ml = std::max<uint64_t>(std::min<uint64_t>(ml, (V >> 1) + 1), 1) + 1;
    return std::min(ml, omax);
  }
  if (max_labels == 0) return 0;
  if (binary_image && connectivity == 26) ml = std::min<uint64_t>(ml, (V >> 2) + 1) + 1;
  else if (binary_image && (connectivity == 6 || connectivity == 4)) ml = std::min<uint64_t>(ml, (V >> 1) + 1) + 1;
  else ml = std::min<uint64_t>(ml, V + 1);
  return std::min(ml, omax);
}

}  // extern "C"

// cc3d_amd_connected_components3d; scratch != NULL: the no-allocation form
// (the job lives on the stack and its workspace is `scratch`).
static int cc_core(const void* in, int in_dtype, int64_t sx, int64_t sy, int64_t sz, int64_t max_labels,
                   int64_t connectivity, const void* delta, void* out, int out_dtype, uint64_t* N,
                   int periodic_boundary, int binary_image, void* stream, char* scratch, size_t scratch_bytes) {
  if (!out || !N) return fail(CC3D_AMD_ERR_INVALID, "NULL output");
  if (out_dtype != DT_U16 && out_dtype != DT_U32 && out_dtype != DT_U64)
    return fail(CC3D_AMD_ERR_INVALID, "output dtype must be uint16, uint32 or uint64");
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint64_t db = 0;
  delta_bits_of(in_dtype, delta, db);
  const bool continuous = !binary_image && !delta_is_zero(in_dtype, db);
  const int64_t V = sx * sy * sz;
  *N = 0;
  // The reference dispatcher's own checks come before any first pass
  // (cc3d/cc3d_continuous.hpp:407-454, cc3d/cc3d.hpp:1432-1470).
  if (continuous && periodic_boundary)
    return fail(CC3D_AMD_ERR_RUNTIME, "periodic_boundary is not currently supported for continuous data.");
  if (connectivity != 26 && connectivity != 18 && connectivity != 6 && connectivity != 8 && connectivity != 4)
    return fail(CC3D_AMD_ERR_RUNTIME, "Only 4 and 8 2D and 6, 18, and 26 3D connectivities are supported.");
  if ((connectivity == 8 || connectivity == 4) && sz != 1)
    return fail(CC3D_AMD_ERR_RUNTIME, "sz must be 1 for 2D connectivities.");
  const uint64_t L = cc3d_amd_union_find_length(max_labels, V, connectivity, binary_image, out_dtype);
  if (L == 0) {
    // the reference returns before its first pass: no foreground is labelled
    if (sx <= 0 || sy <= 0 || sz <= 0) return fail(CC3D_AMD_ERR_INVALID, "volume must be non-empty");
    const dim3 grid(grid_for(V, WG, 65536));
    if (out_dtype == DT_U16) hipLaunchKernelGGL((k_zero_out<uint16_t>), grid, dim3(WG), 0, s, (uint16_t*)out, V);
    else if (out_dtype == DT_U32) hipLaunchKernelGGL((k_zero_out<uint32_t>), grid, dim3(WG), 0, s, (uint32_t*)out, V);
    else hipLaunchKernelGGL((k_zero_out<uint64_t>), grid, dim3(WG), 0, s, (uint64_t*)out, V);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s));
    return CC3D_AMD_OK;
  }
  const bool b8 = binary_image && connectivity == 8 && sz == 1;
  int flags = (continuous ? CC3D_AMD_CONTINUOUS : 0) | CC3D_AMD_NO_SINGLE_ROW_RULE;
  if (b8 && (uint64_t)(((sx + 1) >> 1) * ((sy + 1) >> 1)) >= L) flags |= CC3D_AMD_COUNT_PROVISIONAL;
  cc3d_amd_job local_job;
  cc3d_amd_job* j = scratch ? &local_job : new cc3d_amd_job();
  int rc = job_begin(j, in, in_dtype, sx, sy, sz, connectivity, delta, flags, periodic_boundary, binary_image,
                     scratch, scratch_bytes, stream, scratch != nullptr);
  cc3d_amd_stats st;
  if (rc == CC3D_AMD_OK) rc = cc3d_amd_job_stats(j, &st);
  if (rc == CC3D_AMD_OK && st.provisional >= L && (!b8 || (flags & CC3D_AMD_COUNT_PROVISIONAL)))
    rc = fail(CC3D_AMD_ERR_RUNTIME, "maximum length exception");
  if (rc == CC3D_AMD_OK) rc = cc3d_amd_label_finish(j, out, out_dtype);
  if (rc == CC3D_AMD_OK) *N = st.N;
  if (scratch) {
    // stream-ordered work on the shared scratch ends before the call returns
    const hipError_t e = hipStreamSynchronize(s);
    if (rc == CC3D_AMD_OK && e != hipSuccess) rc = fail(CC3D_AMD_ERR_HIP, hipGetErrorString(e));
    return rc;
  }
  cc3d_amd_job_free(j);
  if (rc == CC3D_AMD_OK) HIPCHECK(hipStreamSynchronize(s));
  return rc;
}

extern "C" {

int cc3d_amd_connected_components3d(const void* in, int in_dtype, int64_t sx, int64_t sy, int64_t sz,
                                    int64_t max_labels, int64_t connectivity, const void* delta, void* out,
                                    int out_dtype, uint64_t* N, int periodic_boundary, int binary_image,
                                    void* stream) {
  return cc_core(in, in_dtype, sx, sy, sz, max_labels, connectivity, delta, out, out_dtype, N, periodic_boundary,
                 binary_image, stream, nullptr, 0);
}

int cc3d_amd_connected_components3d_noalloc(const void* in, int in_dtype, int64_t sx, int64_t sy, int64_t sz,
                                            int64_t max_labels, int64_t connectivity, const void* delta, void* out,
                                            int out_dtype, uint64_t* N, int periodic_boundary, int binary_image,
                                            void* stream) {
  Geometry g;
  uint64_t db = 0;
  delta_bits_of(in_dtype, delta, db);
  const bool continuous = !binary_image && !delta_is_zero(in_dtype, db);
  if (sx > 0 && sy > 0 && sz > 0 && plan(g, sx, sy, sz, connectivity, binary_image, continuous) == CC3D_AMD_OK &&
      g.total > CC3D_AMD_NOALLOC_SCRATCH_BYTES)
    return fail(CC3D_AMD_ERR_INVALID, "cc3d_amd_connected_components3d_noalloc: the volume needs " +
                                          std::to_string(g.total) + " bytes of scratch; the limit is " +
                                          std::to_string((unsigned long long)CC3D_AMD_NOALLOC_SCRATCH_BYTES));
  char* ptr = nullptr;
  int rc = scratch_ready(&ptr);
  if (rc != CC3D_AMD_OK) return rc;
  // per-kernel timing would create events: off for this call
  const bool was_suspended = g_prof_suspended;
  g_prof_suspended = true;
  rc = cc_core(in, in_dtype, sx, sy, sz, max_labels, connectivity, delta, out, out_dtype, N, periodic_boundary,
               binary_image, stream, ptr, CC3D_AMD_NOALLOC_SCRATCH_BYTES);
  g_prof_suspended = was_suspended;
  return rc;
}

void cc3d_amd_profile_enable(int on) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof_on = on == 2 ? 2 : (on != 0 ? 1 : 0);
}

// Waits for the recorded events and returns up to cap (name, total ms, count)
// rows aggregated by kernel name, clearing the records.  Returns the row count.
int cc3d_amd_profile_collect(char* names, int name_len, double* total_ms, int* counts, int cap) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  std::map<std::string, std::pair<double, int>> agg;
  for (auto& r : g_prof_recs) {
    (void)hipEventSynchronize(r.b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    auto& e = agg[r.name];
    e.first += ms;
    e.second += 1;
    g_prof_pool.push_back(r.a);
    g_prof_pool.push_back(r.b);
  }
  g_prof_recs.clear();
  int n = 0;
  for (auto& kv : agg) {
    if (n >= cap) break;
    std::strncpy(names + (size_t)n * name_len, kv.first.c_str(), name_len - 1);
    names[(size_t)n * name_len + name_len - 1] = 0;
    total_ms[n] = kv.second.first;
    counts[n] = kv.second.second;
    n++;
  }
  return n;
}


// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
// Volumes of a batch run the single-volume path's kernels on their own tiles.
// A single plane is labelled on the 3D tiles: its 26/18-neighbourhood within
// the plane is the 8-neighbourhood and its 6-neighbourhood the
// 4-neighbourhood, so 8 and 4 map to 26 and 6.  Not covered: binary 2D-8 and
// 2D-4 (their block and first-row numberings) and continuous 2D-8 (its global
// min/max rule).
static void batch_tile_dims(int fam, int& TX, int& TY, int& TZ) {
  if (fam == cc3d_amd_batch::FAM_PLANE) {
    TX = Geo2::TX; TY = Geo2::TY; TZ = 1;
  } else if (fam == cc3d_amd_batch::FAM_ROWS) {
    TX = GeoR::TX; TY = GeoR::TY; TZ = GeoR::TZ;
  } else if (fam == cc3d_amd_batch::FAM_TALL) {
    TX = Geo3T::TX; TY = Geo3T::TY; TZ = Geo3T::TZ;
  } else {
    TX = Geo3::TX; TY = Geo3::TY; TZ = Geo3::TZ;
  }
}

static int batch_layout(const cc3d_amd_volume* vols, int n, int mode, std::vector<BatchVol>& tab,
                        std::vector<int64_t>& t0, cc3d_amd_batch& b) {
  if (!vols || n <= 0) return fail(CC3D_AMD_ERR_INVALID, "a batch needs at least one volume");
  if (mode != CC3D_AMD_BATCH_MULTILABEL && mode != CC3D_AMD_BATCH_BINARY && mode != CC3D_AMD_BATCH_CONTINUOUS)
    return fail(CC3D_AMD_ERR_INVALID, "batch mode must be multilabel, binary or continuous");
  tab.resize(n);
  b.perm.resize(n);
  b.groups.clear();
  std::vector<int> nbs(n), nbps(n), qks(n);
  std::vector<int> famv(n);
  for (int i = 0; i < n; i++) {  // checked in the caller's order
    const cc3d_amd_volume& v = vols[i];
    if (!v.in) return fail(CC3D_AMD_ERR_INVALID, "NULL input in batch");
    if (v.sx <= 0 || v.sy <= 0 || v.sz <= 0) return fail(CC3D_AMD_ERR_INVALID, "volume must be non-empty");
    const int conn = (int)(v.connectivity & 0xFF);
    int nb = conn;
    if ((v.connectivity &
         ~(int64_t)((CC3D_AMD_TILES_ROWS | CC3D_AMD_TILES_VOXELS | CC3D_AMD_TILES_TALL) << 8 |
                    CC3D_AMD_VOLUME_PERIODIC | 0xFF)) != 0)
      return fail(CC3D_AMD_ERR_RUNTIME, "Only 4 and 8 2D and 6, 18, and 26 3D connectivities are supported.");
    qks[i] = Q_NONE;
    if (nb == 8 || nb == 4) {
      if (v.sz != 1) return fail(CC3D_AMD_ERR_RUNTIME, "sz must be 1 for 2D connectivities.");
      if (mode == CC3D_AMD_BATCH_BINARY) qks[i] = nb == 8 ? Q_BLOCK8 : Q_BIN4;
      else if (mode == CC3D_AMD_BATCH_CONTINUOUS && nb == 8) qks[i] = Q_CONT8;
      else nb = nb == 8 ? 26 : 6;  // a single plane: the same links on the 3D tiles
    } else if (nb != 26 && nb != 18 && nb != 6) {
      return fail(CC3D_AMD_ERR_RUNTIME, "Only 4 and 8 2D and 6, 18, and 26 3D connectivities are supported.");
    }
    nbps[i] = 0;
    if (v.connectivity & CC3D_AMD_VOLUME_PERIODIC) {
      if (mode == CC3D_AMD_BATCH_CONTINUOUS)
        return fail(CC3D_AMD_ERR_INVALID, "periodic_boundary is not implemented for continuous data");
      // the single-volume path's wrap neighbourhood: 6 in 3D, else the plane's
      nbps[i] = v.sz > 1 ? 6 : (conn == 26 || conn == 18) ? 8 : conn == 6 ? 4 : conn;
    }
    nbs[i] = nb;
    // the single-volume path's choice (plan): rows unless continuous, then
    // voxels, then tall; by default rows at 6-connectivity, else voxels
    const int pref = (int)((v.connectivity >> 8) & 0xFF);
    const bool cont = mode == CC3D_AMD_BATCH_CONTINUOUS;
    famv[i] = qks[i] != Q_NONE                           ? cc3d_amd_batch::FAM_PLANE
              : (!cont && (pref & CC3D_AMD_TILES_ROWS))  ? cc3d_amd_batch::FAM_ROWS
              : (pref & CC3D_AMD_TILES_VOXELS)          ? cc3d_amd_batch::FAM_VOXELS
              : (pref & CC3D_AMD_TILES_TALL)            ? cc3d_amd_batch::FAM_TALL
              : (!cont && nb == 6)                      ? cc3d_amd_batch::FAM_ROWS
                                                        : cc3d_amd_batch::FAM_VOXELS;
    b.perm[i] = i;
  }
  // by tile family (voxels, tall, rows), then 26, 18, 6 within each
  std::stable_sort(b.perm.begin(), b.perm.end(), [&](int x, int y) {
    return famv[x] != famv[y] ? famv[x] < famv[y] : nbs[x] > nbs[y];
  });
  uint64_t key = 0;
  b.tgroups.clear();
  b.periodic.clear();
  b.minmax.clear();
  b.stale.clear();
  for (int k = 0; k < n; k++) {
    const int i = b.perm[k], nb = nbs[i], fam = famv[i];
    if (b.tgroups.empty() || b.tgroups.back().fam != fam) {
      cc3d_amd_batch::TileGroup tg;
      tg.fam = fam;
      tg.v0 = k;
      b.tgroups.push_back(tg);
    }
    cc3d_amd_batch::TileGroup& tg = b.tgroups.back();
    int TXg, TYg, TZg;
    batch_tile_dims(fam, TXg, TYg, TZg);
    const cc3d_amd_volume& v = vols[i];
    BatchVol& t = tab[k];
    t.in = v.in;
    t.out = nullptr;
    t.sx = v.sx;
    t.sy = v.sy;
    t.sz = v.sz;
    t.osx = (v.sx + 1) >> 1;
    t.ntx = (v.sx + TXg - 1) / TXg;
    t.nty = (v.sy + TYg - 1) / TYg;
    t.ntz = (v.sz + TZg - 1) / TZg;
    t.ntiles = t.ntx * t.nty * t.ntz;
    t.tile0 = tg.tiles;
    const uint64_t V = (uint64_t)v.sx * (uint64_t)v.sy * (uint64_t)v.sz;
    // binary 2D-8 keys are 2x2 block keys: 4 per block (plan's K)
    const uint64_t K = qks[i] == Q_BLOCK8 ? 4ull * (uint64_t)t.osx * (uint64_t)((v.sy + 1) >> 1) : V;
    t.key0 = (uint32_t)key;
    t.nkeys = (uint32_t)std::min<uint64_t>(K, 0xFFFFFFFFull);
    t.nb = nb;
    t.out_dtype = DT_U32;
    t.qk = qks[i];
    t.minmax = nullptr;
    if (qks[i] == Q_CONT8) b.minmax.push_back({k, 0});
    if (qks[i] == Q_BLOCK8 && ((v.sx & 1) || (v.sy & 1))) {
      cc3d_amd_batch::Stale st{};
      st.k = k;
      st.qblocks = t.osx * ((v.sy + 1) >> 1);
      st.qwords = (st.qblocks + 31) / 32;
      st.qnblk = (st.qwords + SCAN_WPB - 1) / SCAN_WPB;
      b.stale.push_back(st);
    }
    if (b.groups.empty() || b.groups.back().nb != nb || b.groups.back().tg != (int)b.tgroups.size() - 1)
      b.groups.push_back({nb, (int)b.tgroups.size() - 1, tg.tiles, 0});
    b.groups.back().tiles += t.ntiles;
    if (nbps[i]) b.periodic.push_back({k, nbps[i], (int)b.tgroups.size() - 1});
    tg.tiles += t.ntiles;
    tg.v1 = k + 1;
    key += (K + 31) & ~(uint64_t)31;
    if (key >= 0xFFFFFFFFull) return fail(CC3D_AMD_ERR_INVALID, "a batch must hold fewer than 2^32 - 1 voxels");
  }
  // first tile of each volume within its tile group; tile group g's entries
  // start at index v0 + g (one closing entry per group)
  t0.assign(n + b.tgroups.size(), 0);
  for (size_t g = 0; g < b.tgroups.size(); g++) {
    const auto& tg = b.tgroups[g];
    for (int k = tg.v0; k < tg.v1; k++) t0[k + g] = tab[k].tile0;
    t0[tg.v1 + g] = tg.tiles;
  }
  b.n = n;
  b.ntiles = 0;
  for (const auto& tg : b.tgroups) b.ntiles += tg.tiles;
  b.nkeys = (int64_t)key;
  b.nwords = (b.nkeys + 31) / 32;
  b.nblk = (b.nwords + SCAN_WPB - 1) / SCAN_WPB;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t r = o;
    o += align256(bytes);
    return r;
  };
  b.oVOL = take((size_t)n * sizeof(BatchVol));
  b.oGP = take((size_t)b.nkeys * 4);
  b.oRL = take((size_t)b.nkeys * 4);
  b.oBITS = take((size_t)b.nwords * 4);
  b.oPRE = take((size_t)b.nwords * 4);
  b.oBSUM = take((size_t)(b.nblk + 1) * 4);
  b.oCTR = take((size_t)(n + 1) * sizeof(Counters));
  for (auto& tg : b.tgroups) {
    int TXg, TYg, TZg;
    batch_tile_dims(tg.fam, TXg, TYg, TZg);
    const size_t TVg = (size_t)TXg * TYg * TZg;
    const size_t HALOg = tg.rows()                              ? GeoR::HALO
                         : tg.fam == cc3d_amd_batch::FAM_TALL  ? Geo3T::HALO
                         : tg.fam == cc3d_amd_batch::FAM_PLANE ? Geo2::HALO
                                                               : Geo3::HALO;
    const size_t slots = (size_t)tg.tiles * TVg;
    tg.oT0 = take((size_t)(tg.v1 - tg.v0 + 1) * sizeof(int64_t));
    tg.oVIX = take((size_t)tg.tiles * 4);
    tg.oTR = take((size_t)tg.tiles * 4);
    tg.oTQ = take((size_t)tg.tiles * 4);
    tg.oTC = take((size_t)tg.tiles * sizeof(TileCnt));
    tg.oJ8 = take(slots);
    tg.oJ16 = take(slots * 2);
    tg.oREQ = take((size_t)tg.tiles * HALOg * 4);
  }
  for (auto& m : b.minmax) m.oMM = take(64);  // {min, max} and the last NaN index
  for (auto& st : b.stale) {
    st.oQB = take((size_t)st.qwords * 4);
    st.oQP = take((size_t)st.qwords * 4);
    st.oQS = take((size_t)(st.qnblk + 1) * 4);
  }
  b.total = o;
  return CC3D_AMD_OK;
}

size_t cc3d_amd_batch_workspace_bytes(const cc3d_amd_volume* vols, int n, int mode) {
  std::vector<BatchVol> tab;
  std::vector<int64_t> t0;
  cc3d_amd_batch b;
  if (batch_layout(vols, n, mode, tab, t0, b) != CC3D_AMD_OK) return 0;
  return b.total;
}

int cc3d_amd_batch_create(cc3d_amd_batch** out_batch, const cc3d_amd_volume* vols, int n, int in_dtype, int mode,
                          void* workspace, size_t workspace_bytes, void* stream) {
  if (!out_batch) return fail(CC3D_AMD_ERR_INVALID, "NULL batch");
  *out_batch = nullptr;
  if (in_dtype < DT_U8 || in_dtype > DT_F64) return fail(CC3D_AMD_ERR_INVALID, "unsupported input dtype");
  std::vector<BatchVol> tab;
  std::vector<int64_t> t0;
  cc3d_amd_batch* b = new cc3d_amd_batch();
  int rc = batch_layout(vols, n, mode, tab, t0, *b);
  if (rc != CC3D_AMD_OK) {
    delete b;
    return rc;
  }
  b->dt = in_dtype;
  b->mode = mode == CC3D_AMD_BATCH_BINARY ? M_BINARY : mode == CC3D_AMD_BATCH_CONTINUOUS ? M_CONT : M_MULTI;
  b->stream = static_cast<hipStream_t>(stream);
  if (workspace) {
    if (workspace_bytes < b->total) {
      delete b;
      return fail(CC3D_AMD_ERR_INVALID, "workspace too small");
    }
    if (reinterpret_cast<uintptr_t>(workspace) & 255) {
      delete b;
      return fail(CC3D_AMD_ERR_INVALID, "workspace must be 256-byte aligned");
    }
    b->ws = static_cast<char*>(workspace);
  } else {
    void* p = nullptr;
    if (hipMallocAsync(&p, b->total, b->stream) != hipSuccess || !p) {
      delete b;
      return fail(CC3D_AMD_ERR_NOMEM, "could not allocate batch workspace");
    }
    b->ws = static_cast<char*>(p);
    b->own_ws = true;
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&b->h_tab), n * sizeof(BatchVol), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&b->h_t0), t0.size() * sizeof(int64_t), hipHostMallocDefault) !=
        hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&b->h_ctr), (n + 1) * sizeof(Counters), hipHostMallocDefault) !=
        hipSuccess) {
    cc3d_amd_batch_free(b);
    return fail(CC3D_AMD_ERR_NOMEM, "could not allocate pinned batch tables");
  }
  for (const auto& m : b->minmax) tab[m.k].minmax = reinterpret_cast<const uint64_t*>(b->ws + m.oMM);
  std::memcpy(b->h_tab, tab.data(), n * sizeof(BatchVol));
  std::memcpy(b->h_t0, t0.data(), t0.size() * sizeof(int64_t));
  char* ws = b->ws;
  if (hipMemcpyAsync(ws + b->oVOL, b->h_tab, n * sizeof(BatchVol), hipMemcpyHostToDevice, b->stream) != hipSuccess) {
    cc3d_amd_batch_free(b);
    return fail(CC3D_AMD_ERR_HIP, "could not upload the batch table");
  }
  GlobalParams& gp = b->gp;
  gp.GP = reinterpret_cast<uint32_t*>(ws + b->oGP);
  gp.bits = reinterpret_cast<uint32_t*>(ws + b->oBITS);
  gp.prefix = reinterpret_cast<uint32_t*>(ws + b->oPRE);
  gp.bsum = reinterpret_cast<uint32_t*>(ws + b->oBSUM);
  gp.nwords = b->nwords;
  gp.ctr = reinterpret_cast<Counters*>(ws + b->oCTR);
  gp.vols = reinterpret_cast<const BatchVol*>(ws + b->oVOL);
  gp.nvol = n;
  gp.ntiles = b->ntiles;
  bool need_utab = false;
  for (size_t g = 0; g < b->tgroups.size(); g++) {
    auto& tg = b->tgroups[g];
    const int ng = tg.v1 - tg.v0;
    if (hipMemcpyAsync(ws + tg.oT0, b->h_t0 + tg.v0 + g, (ng + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                       b->stream) != hipSuccess) {
      cc3d_amd_batch_free(b);
      return fail(CC3D_AMD_ERR_HIP, "could not upload the batch table");
    }
    GlobalParams& q = tg.gp;
    q = gp;
    q.ntiles = tg.tiles;
    q.RL = reinterpret_cast<uint32_t*>(ws + b->oRL);
    q.TR = reinterpret_cast<uint32_t*>(ws + tg.oTR);
    q.J8 = reinterpret_cast<uint8_t*>(ws + tg.oJ8);
    q.J16 = reinterpret_cast<uint16_t*>(ws + tg.oJ16);
    q.REQ = reinterpret_cast<uint32_t*>(ws + tg.oREQ);
    q.TQ = reinterpret_cast<uint32_t*>(ws + tg.oTQ);
    q.vols = gp.vols + tg.v0;
    q.vtile0 = reinterpret_cast<const int64_t*>(ws + tg.oT0);
    q.vix = reinterpret_cast<const uint32_t*>(ws + tg.oVIX);
    q.nvol = ng;
    hipLaunchKernelGGL(k_fill_vix, dim3((unsigned)ng), dim3(WG), 0, b->stream, q.vtile0,
                       reinterpret_cast<uint32_t*>(ws + tg.oVIX));
    if (hipGetLastError() != hipSuccess) {
      cc3d_amd_batch_free(b);
      return fail(CC3D_AMD_ERR_HIP, "could not index the batch tiles");
    }
    LocalParams& lp = tg.lp;
    lp.utab = nullptr;
    lp.ntiles = tg.tiles;
    lp.delta_bits = 0;
    lp.minmax = nullptr;
    lp.GP = gp.GP;
    lp.RL = q.RL;
    lp.bits = gp.bits;
    lp.TR = reinterpret_cast<uint32_t*>(ws + tg.oTR);
    lp.J8 = reinterpret_cast<uint8_t*>(ws + tg.oJ8);
    lp.J16 = reinterpret_cast<uint16_t*>(ws + tg.oJ16);
    lp.REQ = reinterpret_cast<uint32_t*>(ws + tg.oREQ);
    lp.TQ = reinterpret_cast<uint32_t*>(ws + tg.oTQ);
    lp.tcnt = reinterpret_cast<TileCnt*>(ws + tg.oTC);
    need_utab = need_utab || !tg.rows();
  }
  if (need_utab) {
    const uint16_t* ut = nullptr;
    rc = utab_ready(&ut);
    if (rc != CC3D_AMD_OK) {
      cc3d_amd_batch_free(b);
      return rc;
    }
    for (auto& tg : b->tgroups) tg.lp.utab = ut;
  }
  *out_batch = b;
  return CC3D_AMD_OK;
}

int cc3d_amd_batch_set_outputs(cc3d_amd_batch* b, void* const* outs, const int* out_dtypes) {
  if (!b || !outs || !out_dtypes) return fail(CC3D_AMD_ERR_INVALID, "NULL argument");
  for (int i = 0; i < b->n; i++) {
    if (!outs[i]) return fail(CC3D_AMD_ERR_INVALID, "NULL output in batch");
    if (out_dtypes[i] != DT_U16 && out_dtypes[i] != DT_U32 && out_dtypes[i] != DT_U64)
      return fail(CC3D_AMD_ERR_INVALID, "output dtype must be uint16, uint32 or uint64");
  }
  // the pinned table may still feed an earlier upload
  HIPCHECK(hipStreamSynchronize(b->stream));
  for (int k = 0; k < b->n; k++) {
    b->h_tab[k].out = outs[b->perm[k]];
    b->h_tab[k].out_dtype = out_dtypes[b->perm[k]];
  }
  HIPCHECK(hipMemcpyAsync(b->ws + b->oVOL, b->h_tab, b->n * sizeof(BatchVol), hipMemcpyHostToDevice, b->stream));
  b->outputs_set = true;
  return CC3D_AMD_OK;
}

int cc3d_amd_batch_set_delta(cc3d_amd_batch* b, const void* delta) {
  if (!b || !delta) return fail(CC3D_AMD_ERR_INVALID, "NULL argument");
  if (b->mode != M_CONT) return fail(CC3D_AMD_ERR_INVALID, "delta applies to continuous batches only");
  uint64_t bits = 0;
  const int rc = delta_bits_of(b->dt, delta, bits);
  if (rc != CC3D_AMD_OK) return rc;
  for (auto& tg : b->tgroups) tg.lp.delta_bits = bits;
  b->delta_set = true;
  return CC3D_AMD_OK;
}

int cc3d_amd_batch_begin(cc3d_amd_batch* b) {
  if (!b) return fail(CC3D_AMD_ERR_INVALID, "NULL batch");
  if (b->mode == M_CONT && !b->delta_set)
    return fail(CC3D_AMD_ERR_INVALID, "a continuous batch needs cc3d_amd_batch_set_delta before it runs");
  hipStream_t s = b->stream;
  const GlobalParams& p = b->gp;
  {
    ProfScope ps_("k_init", s);
    hipLaunchKernelGGL(k_batch_init, dim3(grid_for(std::max<int64_t>(b->nwords, b->n + 1), WG, 8192)), dim3(WG), 0, s,
                       p.ctr, b->n + 1, p.bits, b->nwords);
  }
  HIPCHECK(hipGetLastError());
  hipError_t e = hipSuccess;
#define CC3D_BDT(FN)                                 \
  switch (b->dt) {                                   \
    case DT_U8: e = FN<uint8_t>(b, s); break;        \
    case DT_U16: e = FN<uint16_t>(b, s); break;      \
    case DT_U32: e = FN<uint32_t>(b, s); break;      \
    case DT_U64: e = FN<uint64_t>(b, s); break;      \
    case DT_F32: e = FN<float>(b, s); break;         \
    default: e = FN<double>(b, s); break;            \
  }
  if (!b->minmax.empty()) {
    ProfScope ps_("k_minmax", s);
    CC3D_BDT(batch_minmax);
    if (e != hipSuccess) return fail(CC3D_AMD_ERR_HIP, std::string("k_minmax: ") + hipGetErrorString(e));
  }
  {
    ProfScope ps_("k_local", s);
    for (const auto& g : b->groups) {
      const auto& tg = b->tgroups[g.tg];
      const bool tall = tg.fam == cc3d_amd_batch::FAM_TALL;
      if (tg.fam == cc3d_amd_batch::FAM_PLANE && b->mode == M_BINARY)
        e = launch_local_batch_plane_binary(b->dt, g.nb, tg.lp, tg.gp.vols, tg.gp.vix, g.tile0, g.tiles, s);
      else if (tg.fam == cc3d_amd_batch::FAM_PLANE)
        e = launch_local_batch_plane_cont(b->dt, g.nb, tg.lp, tg.gp.vols, tg.gp.vix, g.tile0, g.tiles, s);
      else if (tg.rows() && b->mode == M_BINARY)
        e = launch_rows_batch_binary(b->dt, g.nb, tg.lp, tg.gp.vols, tg.gp.vix, g.tile0, g.tiles, s);
      else if (tg.rows())
        e = launch_rows_batch_multi(b->dt, g.nb, tg.lp, tg.gp.vols, tg.gp.vix, g.tile0, g.tiles, s);
      else if (b->mode == M_BINARY)
        e = launch_local_batch_binary(b->dt, g.nb, tall, tg.lp, tg.gp.vols, tg.gp.vix, g.tile0, g.tiles, s);
      else if (b->mode == M_CONT)
        e = launch_local_batch_cont(b->dt, g.nb, tall, tg.lp, tg.gp.vols, tg.gp.vix, g.tile0, g.tiles, s);
      else
        e = launch_local_batch_multi(b->dt, g.nb, tall, tg.lp, tg.gp.vols, tg.gp.vix, g.tile0, g.tiles, s);
      if (e != hipSuccess) break;
    }
  }
  if (e != hipSuccess) return fail(CC3D_AMD_ERR_HIP, std::string("k_local_batch: ") + hipGetErrorString(e));
  {
    ProfScope ps_("k_tile_counters", s);
    for (const auto& tg : b->tgroups)
      hipLaunchKernelGGL(k_tile_counters_batch, dim3(tg.v1 - tg.v0), dim3(WG), 0, s, tg.lp.tcnt, tg.gp.vtile0,
                         p.ctr + 1 + tg.v0);
  }
  HIPCHECK(hipGetLastError());
  {
    ProfScope ps_("k_requests", s);
    for (const auto& tg : b->tgroups) {
      if (tg.rows()) launch_requests<GeoR, Q_NONE>(tg.gp, tg.tiles, b->mode, s);
      else if (tg.fam == cc3d_amd_batch::FAM_TALL) launch_requests<Geo3T, Q_NONE>(tg.gp, tg.tiles, b->mode, s);
      else if (tg.fam == cc3d_amd_batch::FAM_PLANE) launch_requests<Geo2, Q_NONE>(tg.gp, tg.tiles, b->mode, s);
      else launch_requests<Geo3, Q_NONE>(tg.gp, tg.tiles, b->mode, s);
    }
  }
  HIPCHECK(hipGetLastError());
  if (!b->periodic.empty()) {
    ProfScope ps_("k_periodic", s);
    for (const auto& pv : b->periodic) {
      const int fam = b->tgroups[pv.tg].fam;
      e = fam == cc3d_amd_batch::FAM_ROWS    ? batch_periodic<GeoR>(b, pv, s)
          : fam == cc3d_amd_batch::FAM_TALL  ? batch_periodic<Geo3T>(b, pv, s)
          : fam == cc3d_amd_batch::FAM_PLANE ? batch_periodic<Geo2>(b, pv, s)
                                             : batch_periodic<Geo3>(b, pv, s);
      if (e != hipSuccess) return fail(CC3D_AMD_ERR_HIP, std::string("k_periodic: ") + hipGetErrorString(e));
    }
  }
  {
    ProfScope ps_("k_scan", s);
    launch_scan(p.bits, p.nwords, p.bsum, (int64_t)b->nblk, p.prefix, p.ctr, s);
  }
  {
    hipLaunchKernelGGL(k_batch_N, dim3((b->n + 63) / 64), dim3(64), 0, s, p.vols, b->n, p.prefix, p.nwords, p.ctr,
                       p.ctr + 1);
  }
  HIPCHECK(hipGetLastError());
  if (!b->stale.empty()) {
    CC3D_BDT(batch_q1);
    if (e != hipSuccess) return fail(CC3D_AMD_ERR_HIP, std::string("k_q1: ") + hipGetErrorString(e));
  }
#undef CC3D_BDT
  return CC3D_AMD_OK;
}

int cc3d_amd_batch_stats(cc3d_amd_batch* b, cc3d_amd_stats* st) {
  if (!b || !st) return fail(CC3D_AMD_ERR_INVALID, "NULL argument");
  HIPCHECK(hipMemcpyAsync(b->h_ctr, b->ws + b->oCTR, (b->n + 1) * sizeof(Counters), hipMemcpyDeviceToHost, b->stream));
  HIPCHECK(hipStreamSynchronize(b->stream));
  for (int k = 0; k < b->n; k++) {
    const Counters& c = b->h_ctr[k + 1];
    cc3d_amd_stats& o = st[b->perm[k]];
    o.N = c.N;
    o.provisional = c.prov;
    o.epl = c.epl;
    o.first_foreground_row = c.last_row < 0 ? -1 : c.first_row;
    o.last_foreground_row = c.last_row;
  }
  return CC3D_AMD_OK;
}

int cc3d_amd_batch_finish(cc3d_amd_batch* b) {
  if (!b) return fail(CC3D_AMD_ERR_INVALID, "NULL batch");
  if (!b->outputs_set) return fail(CC3D_AMD_ERR_INVALID, "batch outputs are not set");
  ProfScope ps_("k_finalize", b->stream);
  for (const auto& tg : b->tgroups) {
    if (tg.rows())
      hipLaunchKernelGGL(k_rfinalize_batch, dim3((unsigned)tg.tiles), dim3(WG), 0, b->stream, tg.gp);
    else if (tg.fam == cc3d_amd_batch::FAM_TALL)
      hipLaunchKernelGGL((k_finalize_batch<Geo3T>), dim3((unsigned)((tg.tiles + fin_tiles<Geo3T>() - 1) / fin_tiles<Geo3T>())),
                         dim3(WG), 0, b->stream, tg.gp);
    else if (tg.fam == cc3d_amd_batch::FAM_PLANE)
      hipLaunchKernelGGL((k_finalize_batch<Geo2>), dim3((unsigned)((tg.tiles + fin_tiles<Geo2>() - 1) / fin_tiles<Geo2>())),
                         dim3(WG), 0, b->stream, tg.gp);
    else
      hipLaunchKernelGGL((k_finalize_batch<Geo3>), dim3((unsigned)((tg.tiles + fin_tiles<Geo3>() - 1) / fin_tiles<Geo3>())),
                         dim3(WG), 0, b->stream, tg.gp);
  }
  HIPCHECK(hipGetLastError());
  if (!b->stale.empty()) {
    hipError_t e;
    hipStream_t s = b->stream;
    switch (b->dt) {
      case DT_U8: e = batch_stale<uint8_t>(b, s); break;
      case DT_U16: e = batch_stale<uint16_t>(b, s); break;
      case DT_U32: e = batch_stale<uint32_t>(b, s); break;
      case DT_U64: e = batch_stale<uint64_t>(b, s); break;
      case DT_F32: e = batch_stale<float>(b, s); break;
      default: e = batch_stale<double>(b, s); break;
    }
    if (e != hipSuccess) return fail(CC3D_AMD_ERR_HIP, std::string("k_q1_stale: ") + hipGetErrorString(e));
  }
  return CC3D_AMD_OK;
}

int cc3d_amd_batch_graph_capture(cc3d_amd_batch* b) {
  if (!b) return fail(CC3D_AMD_ERR_INVALID, "NULL batch");
  if (b->own_ws) return fail(CC3D_AMD_ERR_INVALID, "graph capture needs a caller-provided workspace");
  if (!b->outputs_set) return fail(CC3D_AMD_ERR_INVALID, "batch outputs are not set");
  if (b->stream == nullptr) return fail(CC3D_AMD_ERR_INVALID, "graph capture needs a non-default stream");
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  HIPCHECK(hipStreamSynchronize(b->stream));
  HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  g_prof_suspended = true;
  int rc = cc3d_amd_batch_begin(b);
  if (rc == CC3D_AMD_OK) rc = cc3d_amd_batch_finish(b);
  g_prof_suspended = false;
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &graph);
  if (rc != CC3D_AMD_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(CC3D_AMD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  b->graph = graph;
  HIPCHECK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  return CC3D_AMD_OK;
}

int cc3d_amd_batch_graph_launch(cc3d_amd_batch* b) {
  if (!b || !b->gexec) return fail(CC3D_AMD_ERR_INVALID, "no captured graph");
  HIPCHECK(hipGraphLaunch(b->gexec, b->stream));
  return CC3D_AMD_OK;
}

void cc3d_amd_batch_free(cc3d_amd_batch* b) {
  if (!b) return;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->h_tab || b->h_t0 || b->h_ctr) (void)hipStreamSynchronize(b->stream);
  if (b->own_ws && b->ws) (void)hipFreeAsync(b->ws, b->stream);
  if (b->h_tab) (void)hipHostFree(b->h_tab);
  if (b->h_t0) (void)hipHostFree(b->h_t0);
  if (b->h_ctr) (void)hipHostFree(b->h_ctr);
  delete b;
}

}  // extern "C"
