// Shared definitions for the cc3d MI355X labelling kernels.
//
// Layout conventions (all kernels):
//   * volumes are dense, x fastest: index = x + sx * (y + sy * z), the layout
//     cc3d/fastcc3d.pyx:352-359 establishes for both C and F inputs;
//   * the volume is cut into tiles; one 256-thread workgroup owns one tile in
//     the local kernel and the finalize kernel;
//   * union-find "keys" order components for numbering: the voxel index, or for
//     the binary 2D 8-connected path the 2x2 block key
//     ((y>>1)*osx + (x>>1))*4 + (y&1)*2 + (x&1) (cc3d/cc3d_binary.hpp:1016-1161).
//     The root of every component is its minimum key, which is the key of the
//     voxel that receives the first provisional label in the reference scan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cc3d_amd {

enum DType : int { DT_U8 = 0, DT_U16 = 1, DT_U32 = 2, DT_U64 = 3, DT_F32 = 4, DT_F64 = 5 };
enum Mode : int { M_MULTI = 0, M_BINARY = 1, M_CONT = 2 };
// Quirk variants of the 2D paths (see the kernel headers for the rules).
enum Quirk : int { Q_NONE = 0, Q_BLOCK8 = 1, Q_BIN4 = 2, Q_CONT8 = 3 };

constexpr int WG = 256;

// Wave-wide inclusive scans over the 64 lanes with DPP: shifts by 1, 2, 4 and
// 8 lanes within each row of 16, then lane 15 of rows 0 and 2 into rows 1 and
// 3, then lane 31 into rows 2 and 3.  A lane whose source lies outside its
// row, or whose row the row mask leaves out, takes the identity.  No LDS
// traffic and no per-lane address registers, unlike shuffles.  Every lane of
// the wave must be active.
template <int CTRL, int ROWS>
__device__ inline uint32_t dpp_or(uint32_t ident, uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)ident, (int)x, CTRL, ROWS, 0xF, false);
}
template <class Op>
__device__ inline uint32_t wave_scan(uint32_t x, uint32_t ident, Op op) {
  x = op(x, dpp_or<0x111, 0xF>(ident, x));  // row_shr:1
  x = op(x, dpp_or<0x112, 0xF>(ident, x));  // row_shr:2
  x = op(x, dpp_or<0x114, 0xF>(ident, x));  // row_shr:4
  x = op(x, dpp_or<0x118, 0xF>(ident, x));  // row_shr:8
  x = op(x, dpp_or<0x142, 0xA>(ident, x));  // row_bcast:15
  x = op(x, dpp_or<0x143, 0xC>(ident, x));  // row_bcast:31
  return x;
}
__device__ inline uint32_t wave_incl_add(uint32_t x) {
  return wave_scan(x, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}
__device__ inline int wave_incl_min(int x) {
  return (int)wave_scan((uint32_t)x, 0x7FFFFFFFu, [](uint32_t a, uint32_t b) { return (int)a < (int)b ? a : b; });
}
__device__ inline int wave_incl_max(int x) {
  return (int)wave_scan((uint32_t)x, 0x80000000u, [](uint32_t a, uint32_t b) { return (int)a > (int)b ? a : b; });
}
// Lane 63's value (the wave total of an inclusive scan), as a wave-uniform value.
__device__ inline uint32_t wave_lane63(uint32_t x) { return (uint32_t)__builtin_amdgcn_readlane((int)x, 63); }

// Device-side counters of one labelling job.
struct Counters {
  unsigned long long epl;        // provisional-label estimate (transitions)
  unsigned long long prov;       // voxels (or blocks) taking a new provisional label
  long long first_row;           // first row holding foreground (INT64_MAX if none)
  long long last_row;            // last row holding foreground (-1 if none)
  unsigned int req_count;        // cross-tile union requests emitted
  unsigned int req_overflow;     // 1 when the request buffer overflowed
  unsigned long long N;          // number of components
  unsigned long long pad[2];
};

// Per-tile partial counters written by the local kernel (plain stores; one
// reduction kernel folds them into Counters).
struct TileCnt {
  uint32_t epl, prov;
  long long first_row, last_row;
};

// One volume of a batch (device-side table).
struct BatchVol {
  const void* in;
  void* out;
  int64_t sx, sy, sz, osx;
  int64_t ntx, nty, ntz, ntiles;
  int64_t tile0;      // first tile of this volume in the batch's tile numbering
  uint32_t key0;      // first key of this volume in the batch's key space (multiple of 32)
  uint32_t nkeys;
  int nb, out_dtype;
  int qk;                   // numbering / link variant of a single plane on the 2D tiles (Q_*)
  const uint64_t* minmax;   // continuous 2D-8: the volume's global {min, max} bit patterns
};

// Tile geometry.  A tile is TX x TY x TZ voxels; the local kernel's
// workgroup of NT threads (one column of PER_THREAD voxels each) stages it in
// LDS together with a one-voxel halo on the sides that backward neighbours
// reach (x-1..x+1, y-1..y+1 and z-1 in 3D; x-1..x+1 and y-1 in 2D).
template <int TX_, int TY_, int TZ_, bool IS3D_, int NT_>
struct TileGeo {
  static constexpr int TX = TX_, TY = TY_, TZ = TZ_;
  static constexpr bool IS3D = IS3D_;
  static constexpr int NT = NT_;
  static constexpr int TV = TX * TY * TZ;
  static constexpr int PER_THREAD = TV / NT;
  static constexpr int BX = TX + 2;
  static constexpr int BY = IS3D ? TY + 2 : TY + 1;
  static constexpr int BZ = IS3D ? TZ + 1 : 1;
  static constexpr int BOXN = BX * BY * BZ;
  static constexpr int RING = 2 * BX + 2 * (BY - 2);
  static constexpr int HALO = IS3D ? BX * BY + TZ * RING : BX + 2 * TY;
  static constexpr int NN = TV + HALO;
  static_assert(TV % NT == 0 && NT % 64 == 0, "tile must be a multiple of the workgroup");
  static_assert(IS3D ? (TX * TY == NT && TZ == PER_THREAD) : (TX * TY == NT * PER_THREAD && TX <= NT),
                "one column of PER_THREAD voxels per thread");
  // J entry e (thread-major: column e / PER_THREAD, voxel e % PER_THREAD)
  __device__ static inline void entry_voxel(int e, int& lx, int& ly, int& lz) {
    voxel(e / PER_THREAD, e % PER_THREAD, lx, ly, lz);
  }
  static_assert((TX & (TX - 1)) == 0 && (TY & (TY - 1)) == 0 && (TZ & (TZ - 1)) == 0,
                "tile sides are powers of two");

  __host__ __device__ static constexpr int box(int bx, int by, int bz) { return bx + BX * (by + BY * bz); }
  __device__ static inline bool is_tile(int bx, int by, int bz) {
    return bx >= 1 && bx <= TX && by >= 1 && by <= TY && (!IS3D || bz >= 1);
  }
  __device__ static inline int halo_index(int bx, int by, int bz) {
    if (IS3D) {
      if (bz == 0) return bx + BX * by;
      int r;
      if (by == 0) r = bx;
      else if (by == BY - 1) r = BX + bx;
      else r = 2 * BX + 2 * (by - 1) + (bx == 0 ? 0 : 1);
      return BX * BY + (bz - 1) * RING + r;
    } else {
      if (by == 0) return bx;
      return BX + 2 * (by - 1) + (bx == 0 ? 0 : 1);
    }
  }
  __device__ static inline void halo_coords(int h, int& bx, int& by, int& bz) {
    if (IS3D) {
      if (h < BX * BY) {
        bz = 0;
        by = h / BX;
        bx = h - by * BX;
        return;
      }
      h -= BX * BY;
      bz = h / RING + 1;
      int r = h - (bz - 1) * RING;
      if (r < BX) {
        by = 0;
        bx = r;
      } else if (r < 2 * BX) {
        by = BY - 1;
        bx = r - BX;
      } else {
        r -= 2 * BX;
        by = (r >> 1) + 1;
        bx = (r & 1) ? BX - 1 : 0;
      }
    } else {
      bz = 0;
      if (h < BX) {
        by = 0;
        bx = h;
        return;
      }
      h -= BX;
      by = (h >> 1) + 1;
      bx = (h & 1) ? BX - 1 : 0;
    }
  }
  // Position of a tile voxel inside its tile (raster order).
  __device__ static inline int pos(int lx, int ly, int lz) { return lx + TX * (ly + TY * lz); }
  // Union-find node of a box position (tile or halo voxel).  Node order is
  // global key order, so the root of every LDS tree is its member with the
  // smallest key: the box raster index, or for the binary 2D 8-connected
  // numbering the box-local 2x2 block key (x and y shifted by one so that
  // block parity matches the global coordinates; tile origins are even).
  static constexpr int BW = (BX + 2) / 2;
  template <int QK>
  static constexpr int NIDS = QK == Q_BLOCK8 ? 4 * BW * ((BY >> 1) + 1) : BOXN;
  static constexpr int NIDS_MAX = NIDS<Q_BLOCK8> > BOXN ? NIDS<Q_BLOCK8> : BOXN;
  template <int QK>
  __host__ __device__ static constexpr int nid(int bx, int by, int bz) {
    if (QK == Q_BLOCK8) {
      const int xb = bx + 1, yb = by + 1;
      return (((yb >> 1) * BW + (xb >> 1)) << 2) | ((yb & 1) << 1) | (xb & 1);
    }
    return box(bx, by, bz);
  }
  template <int QK>
  __device__ static inline void nid_box(int n, int& bx, int& by, int& bz) {
    if (QK == Q_BLOCK8) {
      const int blk = n >> 2;
      bx = (((blk % BW) << 1) | (n & 1)) - 1;
      by = (((blk / BW) << 1) | ((n >> 1) & 1)) - 1;
      bz = 0;
      return;
    }
    bz = n / (BX * BY);
    const int r = n - bz * (BX * BY);
    by = r / BX;
    bx = r - by * BX;
  }
  // Thread t owns a column of PER_THREAD voxels along the column axis (z in
  // 3D, y in 2D); voxel k of the column.
  __device__ static inline void voxel(int t, int k, int& lx, int& ly, int& lz) {
    lx = t & (TX - 1);
    if (IS3D) {
      ly = t / TX;
      lz = k;
    } else {
      ly = (t / TX) * PER_THREAD + k;
      lz = 0;
    }
  }
  // Index of a tile voxel in the per-voxel tables J8/J16: thread-major, so a
  // thread's column is PER_THREAD consecutive entries.
  __device__ static inline int jidx(int lx, int ly, int lz) {
    if (IS3D) return (lx + TX * ly) * PER_THREAD + lz;
    return (lx + TX * (ly / PER_THREAD)) * PER_THREAD + (ly & (PER_THREAD - 1));
  }
  // Union-find node of a tile voxel: its raster position, or the 2x2 block
  // order used by the binary 2D 8-connected numbering.
  template <int QK>
  __device__ static inline int node(int lx, int ly, int lz) {
    if (QK == Q_BLOCK8)
      return ((((ly >> 1) * (TX / 2) + (lx >> 1)) << 2) | ((ly & 1) << 1) | (lx & 1));
    return pos(lx, ly, lz);
  }
  template <int QK>
  __device__ static inline int node_box(int bx, int by, int bz) {
    if (is_tile(bx, by, bz)) return node<QK>(bx - 1, by - 1, IS3D ? bz - 1 : 0);
    return TV + halo_index(bx, by, bz);
  }
  template <int QK>
  __device__ static inline void node_to_local(int n, int& lx, int& ly, int& lz) {
    if (QK == Q_BLOCK8) {
      const int blk = n >> 2;
      const int by = blk / (TX / 2), bxx = blk - by * (TX / 2);
      lx = (bxx << 1) | (n & 1);
      ly = (by << 1) | ((n >> 1) & 1);
      lz = 0;
    } else {
      lx = n & (TX - 1);
      ly = (n / TX) & (TY - 1);
      lz = n / (TX * TY);
    }
  }
};

// 3D: 32 x 8 x 8 voxels; 2D: 64 x 32 pixels; 256 threads.  (A 32 x 16 x 4
// tile with 512 threads puts fewer columns on the tile border but holds
// fewer tiles per CU in flight: measured slower.)
// Index of a tile's first local root in the per-volume root arrays (global
// forest GP and root keys RL): tiles in raster order, each taking as many
// entries as it holds voxels of the volume (a tile has at most one local root
// per voxel), so the entries of a volume are exactly [key0, key0 + V) and a
// tile's roots sit together.
template <class G>
__host__ __device__ inline uint32_t tile_root0(int64_t sx, int64_t sy, int64_t sz, int64_t tx, int64_t ty, int64_t tz,
                                               uint32_t key0) {
  const int64_t z0 = tz * G::TZ, y0 = ty * G::TY, x0 = tx * G::TX;
  const int64_t wz = G::IS3D ? (sz - z0 < G::TZ ? sz - z0 : G::TZ) : 1;
  const int64_t wy = sy - y0 < G::TY ? sy - y0 : G::TY;
  return key0 + (uint32_t)(sx * sy * z0 + wz * sx * y0 + wz * wy * x0);
}

using Geo3 = TileGeo<32, 8, 8, true, 256>;
// Tall voxel tiles: twice the rows, 512 threads.  Fewer tile borders (18 % of
// the columns instead of 30 %) and a smaller halo per voxel (34 % instead of
// 49 %), for a larger LDS footprint per workgroup.
using Geo3T = TileGeo<32, 16, 8, true, 512>;
using Geo2 = TileGeo<64, 32, 1, false, 256>;

// Global key of a voxel.
template <int QK>
__device__ __host__ inline uint32_t gkey(int64_t gx, int64_t gy, int64_t gz, int64_t sx,
                                         int64_t sy, int64_t osx) {
  if (QK == Q_BLOCK8)
    return (uint32_t)(((((gy >> 1) * osx + (gx >> 1)) << 2)) | ((gy & 1) << 1) | (gx & 1));
  return (uint32_t)(gx + sx * (gy + sy * gz));
}

// XCD-aware tile order: MI355X dispatches workgroups round-robin over 8 XCDs,
// each with a private L2.  Give every XCD a contiguous range of tiles so the
// halos a tile re-reads were recently touched by the same L2.
__device__ inline int64_t xcd_tile(int64_t bid, int64_t ntiles) {
  const int64_t q = ntiles >> 3, r = ntiles & 7;
  const int64_t x = bid & 7, k = bid >> 3;
  return x * q + (x < r ? x : r) + k;
}

// ---------------------------------------------------------------------------
// value helpers
// ---------------------------------------------------------------------------
template <class T>
__device__ inline bool fg(T v) {
  return v != T(0);
}

// cc3d/cc3d_continuous.hpp:79-88: unsigned types compare the absolute
// difference without underflow; float types use std::abs(a - b) <= delta
// evaluated in T.
template <class T>
__device__ inline bool match(T a, T b, T delta) {
  if constexpr (T(-1) > T(0)) {
    const T d = a > b ? T(a - b) : T(b - a);
    return d <= delta;
  } else {
    const T d = a - b;
    return (d < T(0) ? -d : d) <= delta;
  }
}

template <class T>
__device__ __host__ inline T from_bits(uint64_t b) {
  T v;
  __builtin_memcpy(&v, &b, sizeof(T));
  return v;
}

// ---------------------------------------------------------------------------
// union-find
// ---------------------------------------------------------------------------
// LDS: parents are 16-bit node ids (a tile box has fewer than 2^16 nodes);
// the root of a tree is its minimum node id and every other node's parent is
// a smaller id.
//
// LDS has no 16-bit atomic min, so a union lowers a root's parent with a
// compare-and-swap of the aligned 32-bit word holding it (pmin).  Path
// halving (lfind_h) stores the grandparent with a plain 16-bit store: it only
// ever writes a node that is no longer a root, and always an ancestor of it,
// so trees stay trees and no link is lost (a union that loses a race retries
// from the parent it displaced).
__device__ inline uint32_t lload(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline uint32_t pload(const uint16_t* par, uint32_t a) {
  return __hip_atomic_load(par + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void pstore(uint16_t* par, uint32_t a, uint32_t v) {
  __hip_atomic_store(par + a, (uint16_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// par[a] = min(par[a], v); returns the previous par[a]
__device__ inline uint32_t pmin(uint16_t* par, uint32_t a, uint32_t v) {
  uint32_t* w = reinterpret_cast<uint32_t*>(par) + (a >> 1);
  const uint32_t sh = (a & 1u) << 4;
  uint32_t old = lload(w);
  while (true) {
    const uint32_t cur = (old >> sh) & 0xFFFFu;
    if (cur <= v) return cur;
    const uint32_t prev = atomicCAS(w, old, (old & ~(0xFFFFu << sh)) | (v << sh));
    if (prev == old) return cur;
    old = prev;
  }
}
__device__ inline uint32_t lfind_h(uint16_t* par, uint32_t a) {
  uint32_t p = pload(par, a);
  while (p != a) {
    const uint32_t gp = pload(par, p);
    if (gp == p) return p;
    pstore(par, a, gp);
    a = gp;
    p = pload(par, a);
  }
  return a;
}

// Lock-free union: link the larger root under the smaller and retry from the
// displaced parent when another lane got there first.  Returns a node of the
// merged set (a good starting point for the next union of a).
__device__ inline uint32_t lunion(uint16_t* par, uint32_t a, uint32_t b) {
  while (true) {
    a = lfind_h(par, a);
    b = lfind_h(par, b);
    if (a == b) return a;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = pmin(par, a, b);
    if (old == a) return b;
    a = old;
  }
}

// Global union-find over local roots (entries in tile order, RL[i] = key of
// entry i, read-only while unions run): a root is linked under the root with
// the smaller (key, entry) pair, so every tree is rooted at its component's
// smallest key.  (Two entries can share a key: a halo voxel that is a local
// root of a neighbouring tile and of its own tile.)  A link is a
// compare-and-swap on the root's own entry, so a concurrent link is never
// overwritten; (key, entry) only decreases along a path, so no cycle forms.
// Finds compress: the first three nodes of the walked path are pointed
// straight at the root found.  They are non-roots below it, a non-root never
// becomes a root again, and that root stays their ancestor whatever links
// follow, so a plain store is safe (a store racing another compression of the
// same node leaves one of two ancestors).  Compared with path halving this
// shortens the chains that concurrent links build in components spanning
// many tiles (measured: k_requests 0.67 -> 0.43 ms on a 512^3 binary volume
// at p = 0.5, 26-connected).
__device__ inline uint32_t kfind(uint32_t* GP, uint32_t a) {
  uint32_t n0 = a, n1 = a, n2 = a;  // path nodes 0..2 (only those below the root are stored)
  int depth = 0;
  uint32_t r = a, p = __hip_atomic_load(GP + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (p != r) {  // r is path node `depth`, p its parent
    depth++;
    r = p;
    if (depth == 1) n1 = r;
    else if (depth == 2) n2 = r;
    p = __hip_atomic_load(GP + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // node i (i < depth) has parent node i + 1; store where that is not r
  if (depth >= 2) __hip_atomic_store(GP + n0, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (depth >= 3) __hip_atomic_store(GP + n1, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (depth >= 4) __hip_atomic_store(GP + n2, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return r;
}
// Root marks: the local kernel sets the bit of every local root's key, and a
// link clears the key of the root it buries unless the root it joins has the
// same key.  Once the unions are done the bits left are the global roots'
// keys: an entry whose key is not its component's minimum ends below a root
// with a smaller key, so some link on its path took its key to a smaller one
// and cleared it; a root with the minimum key is only ever linked under an
// entry with that same key (entries sharing a key are one component), which
// clears nothing.
__device__ inline void kunion(uint32_t* GP, const uint32_t* RL, uint32_t* bits, uint32_t a, uint32_t b) {
  while (true) {
    a = kfind(GP, a);
    b = kfind(GP, b);
    if (a == b) return;
    uint64_t ka = ((uint64_t)RL[a] << 32) | a, kb = ((uint64_t)RL[b] << 32) | b;
    if (ka < kb) {
      const uint32_t t = a;
      a = b;
      b = t;
      const uint64_t tk = ka;
      ka = kb;
      kb = tk;
    }
    if (atomicCAS(GP + a, a, b) == a) {  // a was still a root: linked under b
      const uint32_t key = (uint32_t)(ka >> 32);
      if (key != (uint32_t)(kb >> 32)) atomicAnd(bits + (key >> 5), ~(1u << (key & 31)));
      return;
    }
  }
}
// Key of the global root of entry a, once every union is made (plain loads:
// the forest no longer changes).  Most entries are roots, so the entry's own
// key is loaded together with its parent.
__device__ inline uint32_t kroot_key(const uint32_t* GP, const uint32_t* RL, uint32_t a) {
  uint32_t p = GP[a];
  const uint32_t k = RL[a];
  if (p == a) return k;
  for (a = p, p = GP[a]; p != a; p = GP[a]) a = p;
  return RL[a];
}

// Global union-find over keys.  Loads bypass the non-coherent L1 so retries see
// other workgroups' links.
__device__ inline uint32_t gload(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Path halving with atomicMin to the grandparent: parents only ever decrease,
// so a node already pointed straight at its root (the minimum of its set) can
// never be moved back to an intermediate ancestor by a concurrent find on a
// longer path.
__device__ inline uint32_t gfind(uint32_t* GP, uint32_t a) {
  uint32_t p = gload(GP + a);
  while (p != a) {
    const uint32_t gp = gload(GP + p);
    if (gp == p) return p;
    atomicMin(GP + a, gp);
    a = gp;
    p = gload(GP + a);
  }
  return a;
}
// Find once every union is made (after the last union kernel): the forest now
// only gets compressed, so every value a node's parent takes is one of its
// ancestors and roots stay fixed.  A stale cached parent still leads to the
// root, so plain loads (which may hit in L2) replace the coherent ones.
__device__ inline uint32_t gfind_settled(uint32_t* GP, uint32_t a) {
  uint32_t p = GP[a];
  while (p != a) {
    const uint32_t gp = GP[p];
    if (gp == p) return p;
    atomicMin(GP + a, gp);
    a = gp;
    p = GP[a];
  }
  return a;
}
__device__ inline void gunion(uint32_t* GP, uint32_t a, uint32_t b) {
  while (true) {
    a = gfind(GP, a);
    b = gfind(GP, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = atomicMin(GP + a, b);
    if (old == a) return;
    a = old;
  }
}

// ---------------------------------------------------------------------------
// backward neighbourhoods, in the order the covering rule visits them
// ---------------------------------------------------------------------------
struct Off {
  int dx, dy, dz;
};
// 3D: E K M B H D F J L A C G I  (first 3: 6-conn, first 9: 18-conn, all: 26)
__host__ __device__ constexpr Off off3(int i) {
  switch (i) {
    case 0: return {0, 0, -1};
    case 1: return {0, -1, 0};
    case 2: return {-1, 0, 0};
    case 3: return {0, -1, -1};
    case 4: return {0, 1, -1};
    case 5: return {-1, 0, -1};
    case 6: return {1, 0, -1};
    case 7: return {-1, -1, 0};
    case 8: return {1, -1, 0};
    case 9: return {-1, -1, -1};
    case 10: return {1, -1, -1};
    case 11: return {-1, 1, -1};
    default: return {1, 1, -1};
  }
}
// 2D: K M J L  (first 2: 4-conn, all: 8-conn)
__host__ __device__ constexpr Off off2(int i) {
  switch (i) {
    case 0: return {0, -1, 0};
    case 1: return {-1, 0, 0};
    case 2: return {-1, -1, 0};
    default: return {1, -1, 0};
  }
}

template <int NB>
struct Nbhd {
  static constexpr bool IS3D = (NB == 26 || NB == 18 || NB == 6);
  static constexpr int N = NB == 26 ? 13 : NB == 18 ? 9 : NB == 6 ? 3 : NB == 8 ? 4 : 2;
  __host__ __device__ static constexpr Off off(int i) { return IS3D ? off3(i) : off2(i); }
  // neighbours j adjacent (under NB) to neighbour i
  __host__ __device__ static constexpr uint32_t adj(int i) {
    uint32_t m = 0;
    for (int j = 0; j < N; j++) {
      if (j == i) continue;
      const Off a = off(i), b = off(j);
      const int ddx = a.dx - b.dx, ddy = a.dy - b.dy, ddz = a.dz - b.dz;
      const int ax = ddx < 0 ? -ddx : ddx, ay = ddy < 0 ? -ddy : ddy, az = ddz < 0 ? -ddz : ddz;
      const int linf = ax > ay ? (ax > az ? ax : az) : (ay > az ? ay : az);
      const int l1 = ax + ay + az;
      bool ok = linf <= 1;
      if (NB == 18) ok = ok && l1 <= 2;
      if (NB == 6 || NB == 4) ok = ok && l1 <= 1;
      if (ok) m |= 1u << j;
    }
    return m;
  }
};


// Bit sets over the positions of the full backward neighbourhood (the 13 of
// 26-connectivity in 3D, the 4 of 8-connectivity in 2D); the smaller
// neighbourhoods are prefixes of these tables.
template <bool IS3D>
struct NbSets {
  static constexpr int N = IS3D ? 13 : 4;
  static constexpr uint32_t ALL = (1u << N) - 1u;
  __host__ __device__ static constexpr Off off(int i) { return IS3D ? off3(i) : off2(i); }
  __host__ __device__ static constexpr uint32_t sel(int axis, int val) {
    uint32_t m = 0;
    for (int i = 0; i < N; i++) {
      const Off o = off(i);
      const int c = axis == 0 ? o.dx : axis == 1 ? o.dy : o.dz;
      if (c == val) m |= 1u << i;
    }
    return m;
  }
  // positions after position j in raster order (z, then y, then x)
  __host__ __device__ static constexpr uint32_t later(int j) {
    uint32_t m = 0;
    const Off a = off(j);
    for (int i = 0; i < N; i++) {
      const Off b = off(i);
      const bool gt = b.dz != a.dz ? b.dz > a.dz : (b.dy != a.dy ? b.dy > a.dy : b.dx > a.dx);
      if (gt) m |= 1u << i;
    }
    return m;
  }
  static constexpr uint32_t XLO = sel(0, -1), XHI = sel(0, 1), X0 = sel(0, 0);
  static constexpr uint32_t YLO = sel(1, -1), YHI = sel(1, 1), Y0 = sel(1, 0);
  static constexpr uint32_t ZLO = sel(2, -1), Z0 = sel(2, 0);
};

// Where a tile voxel's backward neighbourhood leaves the tile.  H = positions
// in the halo; SX0/SX2 (SY0/SY2, SZ0) = positions in the neighbouring tile
// below/above along x (y, z).  All zero for a voxel whose neighbourhood lies
// inside the tile.
struct Border {
  uint32_t H, SX0, SX2, SY0, SY2, SZ0;
};

template <bool IS3D>
__device__ inline Border make_border(bool xlo, bool xhi, bool ylo, bool yhi, bool zlo) {
  using S = NbSets<IS3D>;
  Border b;
  b.SX0 = xlo ? S::XLO : 0u;
  b.SX2 = xhi ? S::XHI : 0u;
  b.SY0 = ylo ? S::YLO : 0u;
  b.SY2 = yhi ? S::YHI : 0u;
  b.SZ0 = zlo ? S::ZLO : 0u;
  b.H = b.SX0 | b.SX2 | b.SY0 | b.SY2 | b.SZ0;
  return b;
}

// Positions i such that the pair (j, i) is unioned inside this tile's
// structure: its later voxel lies in the tile, or both voxels are halo voxels
// of the same neighbouring tile.
template <bool IS3D>
__device__ inline uint32_t proc_mask(int j, const Border& b) {
  using S = NbSets<IS3D>;
  if (!b.H) return S::ALL;
  const uint32_t lat = S::later(j);
  uint32_t m = lat & ~b.H;
  if (!((b.H >> j) & 1u)) {
    m |= S::ALL & ~lat;
  } else {
    const uint32_t sx = ((b.SX0 >> j) & 1u) ? b.SX0 : ((b.SX2 >> j) & 1u) ? b.SX2 : (S::ALL & ~(b.SX0 | b.SX2));
    const uint32_t sy = ((b.SY0 >> j) & 1u) ? b.SY0 : ((b.SY2 >> j) & 1u) ? b.SY2 : (S::ALL & ~(b.SY0 | b.SY2));
    const uint32_t sz = ((b.SZ0 >> j) & 1u) ? b.SZ0 : (S::ALL & ~b.SZ0);
    m |= sx & sy & sz & b.H;
  }
  return m;
}

// Neighbours that a linked neighbour j may cover (connectivity NB): adjacent
// to j, with the pair (j, i) unioned inside this tile's structure.
template <int NB>
__device__ inline uint32_t cover_from(int j, const Border& b) {
  return Nbhd<NB>::adj(j) & proc_mask<Nbhd<NB>::IS3D>(j, b);
}

// Border class of a tile voxel: x (inside / first / last column of the tile)
// + 3 * y (inside / first / last row) + 9 * z (inside / first plane).  2D
// tiles use classes 0..5 (their last row borders nothing backward).
__host__ __device__ constexpr int border_class(bool xlo, bool xhi, bool ylo, bool yhi, bool zlo) {
  return (xlo ? 1 : xhi ? 2 : 0) + 3 * (ylo ? 1 : yhi ? 2 : 0) + 9 * (zlo ? 1 : 0);
}

template <bool IS3D>
__device__ inline Border class_border(int cls) {
  const int xc = cls % 3, yc = (cls / 3) % 3;
  return make_border<IS3D>(xc == 1, xc == 2, yc == 1, yc == 2, cls >= 9);
}

// Union tables of the transitive connectivities (26, 18 and 2D 8).  Entry
// (class, lk) lists the backward positions that a voxel of that border class
// whose linked positions are lk unions with: the first linked position, then
// growth over cover_from marks every linked position adjacent (through a pair
// unioned in the tile's structure) to a connected one, and the next linked
// position that growth cannot reach is the next union.  Class 0 (no border)
// also serves halo voxels, whose lk holds only eligible positions.  Rows of
// 2^N entries; the three tables share one device array of UTAB_TOTAL entries.
template <int NB>
struct UTab {
  static constexpr int N = Nbhd<NB>::N;
  static constexpr int ROWS = Nbhd<NB>::IS3D ? 18 : 6;
  static constexpr int OFF = NB == 26 ? 0 : NB == 18 ? (18 << 13) : (18 << 13) + (18 << 9);
};
constexpr int UTAB_TOTAL = (18 << 13) + (18 << 9) + (6 << 4);

}  // namespace cc3d_amd
