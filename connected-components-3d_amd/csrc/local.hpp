// Local labelling kernel: one workgroup labels one tile in LDS.
//
// Replaces the first raster pass of the reference
// (cc3d/cc3d.hpp:344-1421, cc3d/cc3d_binary.hpp:31-1263,
// cc3d/cc3d_continuous.hpp:90-392).  Instead of a sequential decision tree the
// tile's voxels are unioned in parallel in an LDS union-find; the per-voxel
// link sets are equivalent to the reference's (same transitive closure):
//
//   multilabel  v[q] == v[p]                 (T's ==; v[p] != 0)
//   binary      v[q] != 0
//   continuous  v[q] != 0 && match(v[p], v[q], delta), plus the voxel below
//               when exactly equal under 26-connectivity
//               (cc3d/cc3d_continuous.hpp:147-150)
//   Q_BIN4      binary 2D-4: row 0 links on value equality; later rows link
//               up always and left unless up and up-left are both set
//               (cc3d/cc3d_binary.hpp:822-935)
//   Q_CONT8     continuous 2D-8: only the pixel above when equal or when both
//               lie within delta of the global min/max, else all matches
//               (cc3d/cc3d_continuous.hpp:335-384)
//
// Work layout.  Thread t owns one column of PER_THREAD voxels along the
// column axis (z in 3D, y in 2D); the column neighbour is backward neighbour 0
// (E in 3D, K in 2D).  A voxel linked to its column predecessor inside the
// tile continues a run: its parent is set to the run's first voxel with a
// plain store before any union, so most voxels of large components never
// touch the union-find.  The voxels that still need neighbour unions (run
// starts, and run members whose column link does not cover the rest of the
// neighbourhood) are compacted into an LDS queue and processed densely.
//
// Covering: a linked backward neighbour need not be unioned when it is
// adjacent to another neighbour already connected to p and linked to it,
// because that pair's own union (made when the later of the two was
// processed) connects them.  For multilabel/binary links are transitive, so
// adjacency suffices; for continuous links it takes an exactly equal value or
// a direct match with neighbour 0.  6/4-connectivity uses the phantom voxels
// D, B, J (2D: J) of the 2x2(x2) cube.  On a tile border a pair may only be
// relied on when it is unioned inside this tile's structure (bit masks in
// common.hpp), so tiles never lean on each other's cross-tile links.
//
// Outputs per tile:
//   RL[root0 + r]    global key of local root r (root0 = tile_root0)
//   TR[tile]         number of local roots R
//   J8/J16           per voxel (thread-major: thread t's column is entries
//                    8t..8t+7): 0 (background) or 1 + index of its local root;
//                    8-bit when R <= 255, else 16-bit
//   GP[root0 + r]    initialised to itself for every local root
//   bits             bit (key) set for every local root's key (kunion clears
//                    those of the roots it buries)
//   REQ/TQ           cross-tile union requests: (root index r, halo index)
//                    for each halo "piece" (halo voxels connected through links
//                    that stay inside one neighbouring tile) touching a local
//                    component
//   tcnt[tile]       epl transitions, provisional-label count, first/last
//                    foreground row (cc3d/cc3d.hpp:287-315), folded by
//                    k_tile_counters (one store per tile: thousands of
//                    workgroups atomically updating one address serialise)
#pragma once

#include "common.hpp"

namespace cc3d_amd {

struct LocalParams {
  const void* in;
  int64_t sx, sy, sz, osx;
  int64_t ntx, nty, ntz, ntiles;
  uint64_t delta_bits;
  const uint64_t* minmax;  // Q_CONT8: global {min, max} bit patterns
  uint32_t* GP;
  uint32_t* RL;
  uint32_t* bits;  // root bitmask over keys, cleared beforehand
  uint32_t* TR;
  uint8_t* J8;
  uint16_t* J16;
  uint32_t* REQ;
  uint32_t* TQ;
  TileCnt* tcnt;
  const uint16_t* utab;  // union tables (UTab), UTAB_TOTAL entries
};

// Offsets of backward position i in the staged box and in the tile's node
// numbering (raster order).
template <class G>
__host__ __device__ constexpr int box_off(int i) {
  const Off o = NbSets<G::IS3D>::off(i);
  return o.dx + G::BX * (o.dy + G::BY * o.dz);
}
template <class G>
__host__ __device__ constexpr int node_off(int i) {
  const Off o = NbSets<G::IS3D>::off(i);
  return o.dx + G::TX * (o.dy + G::TY * o.dz);
}

// Offsets of the backward positions along one axis, packed as (d + 1) in 2-bit
// fields, for run-time position indices.
template <bool IS3D>
__host__ __device__ constexpr uint32_t packed_offsets(int axis) {
  uint32_t m = 0;
  for (int i = 0; i < NbSets<IS3D>::N; i++) {
    const Off o = NbSets<IS3D>::off(i);
    const int d = axis == 0 ? o.dx : axis == 1 ? o.dy : o.dz;
    m |= (uint32_t)(d + 1) << (2 * i);
  }
  return m;
}

// Continuous 2D-8 shortcut (cc3d/cc3d_continuous.hpp:341-349): both values
// within delta of the global minimum and maximum.
template <class T>
__device__ inline bool q3_span(T cur, T vb, T delta, T gmin, T gmax) {
  const T lo = (vb < cur) ? vb : cur;
  const T hi = (cur < vb) ? vb : cur;
  const T d1 = lo - gmin, d2 = gmax - hi;
  return d1 <= delta && d2 <= delta;
}

// Link between a foreground voxel and its column predecessor.
template <class T, int MODE, int NB, int QK>
__device__ inline bool column_link(T cur, T prev, T delta, T gmin, T gmax) {
  if constexpr (QK == Q_BIN4) {
    return fg(prev);  // rows after the first link up whenever the pixel above is set
  } else if constexpr (QK == Q_CONT8) {
    return prev == cur || (fg(prev) && (q3_span(cur, prev, delta, gmin, gmax) || match(cur, prev, delta)));
  } else if constexpr (MODE == M_CONT) {
    return fg(prev) && (match(cur, prev, delta) || (NB == 26 && prev == cur));
  } else if constexpr (MODE == M_MULTI) {
    return prev == cur;
  } else {
    return fg(prev);
  }
}

// Node of backward position i (a run-time index) of the voxel at box
// (bx, by, bz) with node pn; offsets are unpacked from 2-bit fields.
template <int QK, class G>
__device__ inline uint32_t node_at(uint32_t pn, int bx, int by, int bz, int i) {
  constexpr bool IS3D = G::IS3D;
  constexpr uint32_t PX = packed_offsets<IS3D>(0), PY = packed_offsets<IS3D>(1), PZ = packed_offsets<IS3D>(2);
  const int dx = (int)((PX >> (2 * i)) & 3u) - 1;
  const int dy = (int)((PY >> (2 * i)) & 3u) - 1;
  const int dz = (int)((PZ >> (2 * i)) & 3u) - 1;
  if (QK != Q_BLOCK8) return pn + dx + G::BX * (dy + G::BY * dz);
  return G::template nid<QK>(bx + dx, by + dy, bz + dz);
}

// Link mask of a foreground voxel (box index b, value cur) under a
// transitive connectivity (26/18/8): bit i when backward position i is in
// okm and linked.  A run member (col) has bit 0 set without reading it.
template <class T, int MODE, int NB, class G>
__device__ inline uint32_t link_mask(const T* sv, int b, T cur, uint32_t okm, bool col) {
  uint32_t lk = col ? 1u : 0u;
#pragma unroll
  for (int i = 0; i < Nbhd<NB>::N; i++) {
    if (col && i == 0) continue;
    if (!((okm >> i) & 1u)) continue;
    const T q = sv[b + box_off<G>(i)];
    const bool l = MODE == M_MULTI ? (q == cur) : fg(q);
    lk |= l ? (1u << i) : 0u;
  }
  return lk;
}

// Raw buffer loads (gfx9 buffer resource, 32-bit byte offsets): an offset
// at or above the resource's record count reads 0 without touching memory,
// which is background for every voxel type.
constexpr uint32_t OOB = 0x80000000u;
constexpr int BUF_DW3 = 0x00020000;
template <class T>
__device__ inline T bload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 1) {
    return __builtin_bit_cast(T, (uint8_t)__builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0));
  } else if constexpr (sizeof(T) == 2) {
    return __builtin_bit_cast(T, (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0));
  } else if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  }
}

// 6-connected voxel p (box index b): the links to M (position 2) and K (1)
// and whether each still needs a union here, given whether p is linked to E
// (0).  A phantom path x - c - y (both pairs unioned in this tile) makes the
// x - y union redundant; positions E K M B D J = 0 1 2 3 5 7.  Bits: 0 union
// M, 1 union K, 2 M linked, 3 K linked.
template <class T, int MODE, class G, bool HALO>
__device__ inline uint32_t six_links(const T* sv, int b, T cur, const Border& bd, uint32_t elig, bool lE) {
  auto lk = [&](int i) -> bool {
    if (HALO && !((elig >> i) & 1u)) return false;
    const T q = sv[b + box_off<G>(i)];
    return MODE == M_MULTI ? (q == cur) : fg(q);
  };
  auto path = [&](int x, int c, int y) -> bool {
    if (HALO) return true;
    return ((proc_mask<true>(x, bd) >> c) & 1u) && ((proc_mask<true>(c, bd) >> y) & 1u);
  };
  const bool lM = lk(2);
  const bool uM = lM && !(lE && lk(5) && path(0, 5, 2));
  const bool lK = lk(1);
  const bool uK = lK && !((lE && lk(3) && path(0, 3, 1)) || (lM && lk(7) && path(2, 7, 1)));
  return (uM ? 1u : 0u) | (uK ? 2u : 0u) | (lM ? 4u : 0u) | (lK ? 8u : 0u);
}

// Unions for one foreground voxel p (box index b, node pn, box coordinates
// bx, by, bz).  HALO: a halo voxel, restricted to the positions in `elig`
// (halo voxels of the same neighbouring tile).  COL: p continues a run, so
// its link to neighbour 0 is already in place.  Returns 0 exactly when p has
// no linked backward neighbour (the reference gives p a new provisional label).
template <class T, int MODE, int NB, int QK, class G, bool HALO>
__device__ inline int link_voxel(const T* sv, uint16_t* par, int b, uint32_t pn, int bx, int by, int bz,
                                 T cur, const Border& bd, const uint16_t* ut, uint32_t elig, bool col, T delta,
                                 T gmin, T gmax, bool row0) {
  using NH = Nbhd<NB>;
  constexpr bool IS3D = G::IS3D;
  auto val = [&](int i) -> T { return sv[b + box_off<G>(i)]; };
  auto nd = [&](int i) -> uint32_t {
    const Off o = NbSets<IS3D>::off(i);
    if (QK != Q_BLOCK8) return pn + box_off<G>(i);  // node ids are box indices
    return G::template nid<QK>(bx + o.dx, by + o.dy, bz + o.dz);
  };
  auto cover = [&](int i) -> uint32_t { return HALO ? NH::adj(i) : cover_from<NB>(i, bd); };
  // node of position i for a run-time i (offsets unpacked from 2-bit fields)
  auto nd_dyn = [&](int i) -> uint32_t { return node_at<QK, G>(pn, bx, by, bz, i); };
  auto ok = [&](int i) -> bool { return !HALO || ((elig >> i) & 1u); };
  uint32_t a = pn;

  if constexpr (QK == Q_BIN4) {
    // positions K M J L = 0 1 2 3; cc3d/cc3d_binary.hpp:822-935
    if (row0) {
      if (val(1) == cur) {
        lunion(par, a, nd(1));
        return 1;
      }
      return 0;
    }
    const bool C = col || fg(val(0));
    const bool B = fg(val(1));
    const bool D = fg(val(2));
    int n = 0;
    if (C) {
      if (!col) a = lunion(par, a, nd(0));
      n++;
    }
    if (B && !(C && D)) {
      lunion(par, a, nd(1));
      n++;
    }
    return n;
  } else if constexpr (QK == Q_CONT8) {
    // cc3d/cc3d_continuous.hpp:335-384
    int n = 0;
    if (col) {
      n = 1;  // a match-only link to K: the other matches follow
    } else {
      const T vb = val(0);
      if (vb == cur || (fg(vb) && q3_span(cur, vb, delta, gmin, gmax))) {
        lunion(par, a, nd(0));
        return 1;
      }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (i == 0 && col) continue;
      const T q = val(i);
      if (fg(q) && match(cur, q, delta)) {
        a = lunion(par, a, nd(i));
        n++;
      }
    }
    return n;
  } else if constexpr (MODE != M_CONT && (NB == 6 || NB == 4)) {
    auto lk = [&](int i) -> bool {
      if (!ok(i)) return false;
      const T q = val(i);
      return MODE == M_MULTI ? (q == cur) : fg(q);
    };
    // a phantom path x - c - y connects x and y when both pairs are unioned here
    auto path = [&](int x, int c, int y) -> bool {
      if (HALO) return true;
      return ((proc_mask<IS3D>(x, bd) >> c) & 1u) && ((proc_mask<IS3D>(c, bd) >> y) & 1u);
    };
    if constexpr (NB == 6) {
      const bool lE = col || lk(0);
      if (lE && !col) a = lunion(par, a, nd(0));
      const uint32_t s6 = six_links<T, MODE, G, HALO>(sv, b, cur, bd, elig, lE);
      if (s6 & 1u) a = lunion(par, a, nd(2));
      if (s6 & 2u) lunion(par, a, nd(1));
      return (int)lE + (int)((s6 >> 2) & 1u) + (int)((s6 >> 3) & 1u);
    } else {
      // positions K M J = 0 1 2
      const bool lK = col || lk(0);
      if (lK && !col) a = lunion(par, a, nd(0));
      const bool lM = lk(1);
      if (lM && !(lK && lk(2) && path(0, 2, 1))) lunion(par, a, nd(1));
      return (int)lK + (int)lM;
    }
  } else if constexpr (MODE != M_CONT) {
    // 26 / 18 / 8: transitive links.  Read every neighbour into a link mask;
    // the union table row of p's border class (ut) lists the positions to
    // union with (UTab in common.hpp).  A run member's link to neighbour 0 is
    // in place already.
    constexpr uint32_t ALLN = (1u << NH::N) - 1u;
    const uint32_t lk = link_mask<T, MODE, NB, G>(sv, b, cur, HALO ? (elig & ALLN) : ALLN, col);
    uint32_t u = (uint32_t)ut[lk] & ~(col ? 1u : 0u);
    while (u) {
      const int i = __builtin_ctz(u);
      u &= u - 1u;
      a = lunion(par, a, nd_dyn(i));
    }
    return lk ? 1 : 0;
  } else {
    // continuous: exact-equality covering plus direct matches with neighbour
    // 0.  Covered neighbours are still read: an exactly equal one extends the
    // cover.
    uint32_t cov = 0;
    int n = 0;
    bool hub = false;
    uint32_t hubcov = 0;
    T hv = T(0);
    if (col) {
      hv = val(0);
      hub = true;
      hubcov = cover(0);
      n = 1;
      cov = 1u | (hv == cur ? hubcov : 0u);
    }
#pragma unroll
    for (int i = 0; i < NH::N; i++) {
      if (!ok(i) || (col && i == 0)) continue;
      const T q = val(i);
      bool l = fg(q) && match(cur, q, delta);
      if (NB == 26 && i == 0 && q == cur) l = true;
      if (!l) continue;
      n++;
      const bool covered = ((cov >> i) & 1u) || (i > 0 && hub && ((hubcov >> i) & 1u) && match(hv, q, delta));
      if (!covered) a = lunion(par, a, nd(i));
      const uint32_t cf = cover(i);
      if (i == 0) {
        hub = true;
        hv = q;
        hubcov = cf;
      }
      if (q == cur) cov |= cf;
    }
    return n;
  }
}

// Positions a halo voxel at box (bx, by, bz) may link to: backward neighbours
// inside the box that are halo voxels of the same neighbouring tile.
template <class G>
__device__ inline uint32_t halo_elig(int bx, int by, int bz) {
  using S = NbSets<G::IS3D>;
  const bool xmid = bx >= 1 && bx <= G::TX;
  const bool ymid = by >= 1 && by <= G::TY;
  const uint32_t ex = xmid ? (S::X0 | (bx >= 2 ? S::XLO : 0u) | (bx <= G::TX - 1 ? S::XHI : 0u)) : S::X0;
  const uint32_t ey = ymid ? (S::Y0 | (by >= 2 ? S::YLO : 0u) | (by <= G::TY - 1 ? S::YHI : 0u)) : S::Y0;
  uint32_t ez = S::ALL;
  if (G::IS3D) ez = bz == 0 ? S::Z0 : (S::Z0 | (bz >= 2 ? S::ZLO : 0u));
  return ex & ey & ez;
}

// Is a foreground run member fully linked by its column link alone?
template <class T, int MODE, int NB, int QK>
__device__ inline bool column_done(T cur, T prev, bool col_covers_all, T delta, T gmin, T gmax) {
  if constexpr (QK == Q_CONT8) return prev == cur || q3_span(cur, prev, delta, gmin, gmax);
  else if constexpr (QK == Q_BIN4) return false;
  else if constexpr (NB == 6 || NB == 4) return false;  // face links: M and K still checked
  else if constexpr (MODE == M_CONT) return col_covers_all && prev == cur;
  else return col_covers_all;
}

// Work-queue capacity.  A whole tile where LDS allows it: for 1- and
// 2-byte voxels, and where most foreground voxels are queued (FULLQ:
// 6/4/18-connectivity, continuous links).  Otherwise the queue gets 42 % of a tile, which keeps 8
// workgroups per CU for 4-byte voxels and 5 for 8-byte ones: the tiles of
// the 512^3 uint32 Voronoi workload queue 621 voxels on average and at most
// 847 of 2048.  Voxels past the capacity are unioned by their own thread
// after the queue.
template <class T, class G, bool FULLQ>
__host__ __device__ constexpr int queue_cap() {
  return (sizeof(T) <= 2 || FULLQ) ? G::TV : (G::TV * 107 / 256) & ~7;
}

// Waves per SIMD the local kernel is compiled for, which caps its registers:
// what the LDS of one tile admits.  3D tiles take 13.7 KB (1-byte voxels),
// 16.7 KB (2), 20.4 KB (4; 22.1 KB with a whole-tile queue) and 32.7 KB (8;
// 35.1 KB): 8 waves per SIMD (the limit) for 1- and 2-byte voxels and for
// 4-byte ones with the short queue, 7 for 4-byte voxels with a whole-tile
// queue, 5 and 4 for 8-byte voxels.
template <class T, class G, bool FULLQ>
__host__ __device__ constexpr int local_waves() {
  return sizeof(T) <= 2 ? 8 : sizeof(T) == 4 ? (FULLQ ? 7 : 8) : (FULLQ ? 4 : 5);
}

// Whole-tile queue where most foreground voxels are queued: under 4 and
// binary 6 every foreground voxel, under 18 every foreground run member (a
// column link leaves two in-plane corners uncovered), and for continuous
// links every run member whose value is not exactly equal to its
// predecessor's.  Multi-label 6-connected tiles queue their run starts and
// the run members that still need a face union, about as many as under 26.
template <int MODE, int NB>
__host__ __device__ constexpr bool full_queue() {
  return (NB == 6 && MODE == M_BINARY) || NB == 4 || NB == 18 || MODE == M_CONT;
}

// LDS of one tile (declared once per kernel, so kernels that dispatch to
// several connectivities at run time allocate it once).
template <class T, class G, bool FULLQ>
struct LocalLds {
  static constexpr int NW = G::NT / 64;
  static constexpr int RW = (G::NIDS_MAX + 31) / 32;  // words of the root bitmask
  static constexpr int QC = queue_cap<T, G, FULLQ>();
  T sv[G::BOXN];
  alignas(4) uint16_t par[(G::NIDS_MAX + 1) & ~1];  // pmin works on aligned 32-bit words
  // storage of steps that are never live together, separated by barriers
  union {
    struct {  // steps 1-3
      uint16_t hqueue[G::IS3D ? G::HALO : 1];  // box indices of halo run starts
      // per ring column (3D): bit k set unless its voxel k is linked to voxel
      // k - 1 (bit 0 always set)
      uint8_t rbreak[G::IS3D ? G::RING : 1];
    } h;
    struct {
      uint16_t queue[QC];  // steps 3-4: voxel position | run member << 15
      uint32_t rbits[RW];  // steps 5-8: local roots that own tile voxels
    } w;
    struct {               // steps 6-9, over the queue only
      uint32_t rprefix[RW];  // exclusive popcount prefix of rbits
      uint32_t red_epl[NW], red_prov[NW];
      int32_t red_first[NW], red_last[NW];
      uint32_t nroots;
    } late;
  };
  uint32_t qcount, nreq, hcount;
  static_assert(QC % 2 == 0 && RW * 4 + NW * 16 + 4 <= QC * 2, "the late storage stays clear of rbits");
};

// Labels tile `tile` of the volume described by p.  Per-tile outputs go to
// index `slot` of RL/J/REQ/TR/TQ/tcnt; keys are offset by `keybase` (both
// are `tile` and 0 for a single volume; a batch concatenates volumes).
//
// Union-find nodes are box positions numbered in global key order (tile and
// halo voxels alike), so every LDS tree is rooted at its member with the
// smallest key.  A local component's root can then be a halo voxel: its key
// stands for the component in the global forest, and the requests of its
// halo pieces join it to the neighbouring tiles' forests.  The first voxel of
// a column links to its predecessor (the halo voxel below in 3D, the voxel
// above in 2D) with a plain store, like the rest of the column.
template <class T, int MODE, int NB, int QK, class G, bool FULLQ>
__device__ inline void local_tile(const LocalParams& p, int64_t tile, int64_t slot, uint32_t keybase,
                                  LocalLds<T, G, FULLQ>& L) {
  constexpr int NT = G::NT;  // threads of the workgroup, one column each
  constexpr int NW = NT / 64;
  constexpr int PT = G::PER_THREAD;
  constexpr int HPT = (G::HALO + NT - 1) / NT;
  constexpr bool PIECES = !(QK == Q_BIN4 || QK == Q_CONT8);
  constexpr bool IS3D = G::IS3D;
  constexpr int CSTEP = IS3D ? G::BX * G::BY : G::BX;  // box step along the column
  constexpr int RW = LocalLds<T, G, FULLQ>::RW;
  static_assert(HPT <= 32 && (!G::IS3D || (G::TZ <= 32 && G::BY <= 32)), "halo piece flags fit one word");
  constexpr int WPL = (RW + 63) / 64;  // root bitmask words per lane of the scanning wave
  static_assert(WPL <= 4, "one wave scans the root bitmask");
  // 3D transitive 26/18: halo staged as runs (see step 1)
  constexpr bool HRUNS = PIECES && IS3D && MODE != M_CONT && (NB == 26 || NB == 18);
  // 26-connected border run members consult the ring columns' run breaks (step 3)
  constexpr bool RBREAK = HRUNS && NB == 26;
  static_assert(!RBREAK || G::TZ <= 8, "one byte of run breaks per ring column");
  constexpr uint32_t NONE = 0xFFFFFFFFu;

  T* sv = L.sv;
  uint16_t* par = L.par;
  // union table of this connectivity, class 0 (transitive connectivities only)
  constexpr bool UT = MODE != M_CONT && (NB == 26 || NB == 18 || NB == 8) && QK != Q_CONT8;
  const uint16_t* ut0 = UT ? p.utab + UTab<UT ? NB : 26>::OFF : nullptr;
  uint16_t* queue = L.w.queue;
  constexpr uint32_t QC = LocalLds<T, G, FULLQ>::QC;
  uint16_t* hqueue = L.h.hqueue;
  uint32_t& qcount = L.qcount;
  uint32_t& nreq = L.nreq;
  uint32_t& hcount = L.hcount;

  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int64_t tz = tile / (p.ntx * p.nty);
  const int64_t trem = tile - tz * p.ntx * p.nty;
  const int64_t ty = trem / p.ntx, tx = trem - ty * p.ntx;
  const int64_t x0 = tx * G::TX, y0 = ty * G::TY, z0 = tz * G::TZ;
  const int64_t sx = p.sx, sy = p.sy, sz = p.sz;
  const T* in = static_cast<const T*>(p.in);
  const T delta = from_bits<T>(p.delta_bits);
  T gmin = T(0), gmax = T(0);
  if constexpr (QK == Q_CONT8) {
    gmin = from_bits<T>(p.minmax[0]);
    gmax = from_bits<T>(p.minmax[1]);
  }

  if (t == 0) {
    qcount = 0;
    nreq = 0;
    hcount = 0;
  }
  __syncthreads();

  // this thread's column
  int lx, ly0, lz0;
  G::voxel(t, 0, lx, ly0, lz0);
  const int64_t gx = x0 + lx;
  const int b0 = G::box(lx + 1, ly0 + 1, IS3D ? 1 : 0);
  auto vnode = [&](int k) -> uint32_t {  // node of column voxel k
    if (QK != Q_BLOCK8) return (uint32_t)(b0 + k * CSTEP);
    int vx, vy, vz;
    G::voxel(t, k, vx, vy, vz);
    return (uint32_t)G::template nid<QK>(vx + 1, vy + 1, IS3D ? vz + 1 : 0);
  };

  // 1. stage the column (registers + LDS) and the halo; voxels outside the
  //    volume read as background.  Column runs get their parent here.  All
  //    loads (column, its predecessor, halo) are issued before any value is
  //    used, so one wait covers them.
  // HRUNS bottom plane: columns along y cut into NBS segments of BSEG voxels
  constexpr int NBS = (G::BY + 9) / 10;
  constexpr int BSEG = (G::BY + NBS - 1) / NBS;
  constexpr int HLEN = G::TZ > BSEG ? G::TZ : BSEG;
  constexpr int HV = HRUNS ? HLEN : HPT;
  T col[PT];
  T hv[HV];
  T pv;
  uint32_t runs = 0;  // bit k: voxel k continues the run of its predecessor
  // 3D halo as columns (HRUNS): ring position r (threads 0..RING-1) along z;
  // bottom plane column bx, segment g (thread RING + bx + BX * g) along y
  // from y = g * BSEG.  Otherwise thread t stages halo voxels t, t + NT, ...
  constexpr int RING = G::RING;
  const bool hcol = HRUNS && t < RING + G::BX * NBS;
  const bool ring = t < RING;
  int cbx = 0, cby = 0, hby0 = 0;  // hby0: first y of a bottom segment
  if (HRUNS) {
    if (ring) {
      if (t < G::BX) {
        cbx = t;
        cby = 0;
      } else if (t < 2 * G::BX) {
        cbx = t - G::BX;
        cby = G::BY - 1;
      } else {
        cbx = ((t - 2 * G::BX) & 1) ? G::BX - 1 : 0;
        cby = ((t - 2 * G::BX) >> 1) + 1;
      }
    } else {
      cbx = (t - RING) % G::BX;
      cby = 0;
      if constexpr (NBS > 1) hby0 = ((t - RING) / G::BX) * BSEG;
    }
  }
  const int hlen = ring ? G::TZ : (G::BY - hby0 < BSEG ? G::BY - hby0 : BSEG);
  T hpv = T(0);  // bottom segment after the first: the voxel before it
#pragma unroll
  for (int k = 0; k < HV; k++) hv[k] = T(0);
  // The box (tile plus halo) spans fewer than 2^31 bytes of the volume for
  // any realistic plane size: then every voxel is one raw buffer load at a
  // 32-bit byte offset from box position 0, and voxels outside the volume
  // read 0 (offset OOB, or past the volume's end, which the record count
  // covers).  Otherwise 64-bit addresses with explicit bounds.
  const bool buf_ok =
    (uint64_t)(G::BX + sx * G::BY + (IS3D ? sx * sy * G::BZ : 0)) * sizeof(T) < (uint64_t)OOB;
  if (buf_ok) {
    const int64_t base_el = (x0 - 1) + sx * ((y0 - 1) + (IS3D ? sy * (z0 - 1) : 0));
    const uint64_t rem = (uint64_t)((IS3D ? sx * sy * sz : sx * sy) - base_el) * sizeof(T);
    const int nrec = (int)(rem < OOB ? rem : OOB - 1u);
    const char* bp = reinterpret_cast<const char*>(in) + base_el * (int64_t)sizeof(T);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(bp), 0, nrec, BUF_DW3);
    constexpr uint32_t SZ = sizeof(T);
    const uint32_t RB = (uint32_t)sx * SZ, PB = IS3D ? (uint32_t)(sx * sy) * SZ : 0u;
    const uint32_t CST = IS3D ? PB : RB;                                          // column step
    const uint32_t c0 = (uint32_t)(lx + 1) * SZ + (uint32_t)(ly0 + 1) * RB + PB;  // column voxel 0
    const uint32_t coff = (gx < sx && y0 + ly0 < sy) ? c0 : OOB;
#pragma unroll
    for (int k = 0; k < PT; k++) col[k] = bload<T>(rs, coff + (uint32_t)k * CST);
    // predecessor of voxel 0: (x, y, z0 - 1) in 3D, (x, y - 1) in 2D
    const bool pok = gx < sx && (IS3D ? (y0 + ly0 < sy && z0 > 0) : (y0 + ly0 > 0));
    pv = bload<T>(rs, pok ? c0 - CST : OOB);
    if constexpr (HRUNS) {
      if (hcol) {
        // ring column: box (cbx, cby, 1 + k); bottom segment: box (cbx, hby0 + k, 0)
        const int64_t hx = x0 + cbx - 1, hy = y0 + cby - 1, hyb = y0 - 1 + hby0;
        const bool fixok = hx >= 0 && hx < sx && (ring ? (hy >= 0 && hy < sy) : z0 > 0);
        const uint32_t h0 = (uint32_t)cbx * SZ + (ring ? (uint32_t)cby * RB + PB : (uint32_t)hby0 * RB);
        const uint32_t hst = ring ? PB : RB;
#pragma unroll
        for (int k = 0; k < HV; k++) {
          const bool along = ring ? (k < G::TZ) : (k < hlen && hyb + k >= 0 && hyb + k < sy);
          hv[k] = bload<T>(rs, (fixok && along) ? h0 + (uint32_t)k * hst : OOB);
        }
        if constexpr (NBS > 1) {
          const bool pbok = !ring && hby0 > 0 && fixok && hyb - 1 >= 0 && hyb - 1 < sy;
          hpv = bload<T>(rs, pbok ? h0 - RB : OOB);
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < HV; k++) {
        const bool live = t + NT * k < G::HALO;
        int bx, by, bz;
        G::halo_coords(live ? t + NT * k : 0, bx, by, bz);
        const int64_t hx = x0 + bx - 1, hy = y0 + by - 1, hz = IS3D ? z0 + bz - 1 : 0;
        const bool ok = live && hx >= 0 && hx < sx && hy >= 0 && hy < sy && hz >= 0;
        hv[k] = bload<T>(rs, ok ? (uint32_t)bx * SZ + (uint32_t)by * RB + (uint32_t)bz * PB : OOB);
      }
    }
  } else {
    // All global loads are issued before any value is used: out-of-volume
    // voxels load a valid address and are replaced by background afterwards.
    bool inb[PT];
    // one base address per column, stepped by a stride that is the same for
    // every thread; the bounds along the column are per tile
    const int64_t cstride = IS3D ? sx * sy : sx;                   // voxel k -> k + 1
    const int64_t ccount = IS3D ? sz - z0 : sy - (y0 + ly0);       // voxels left in the volume
    const bool colok = gx < sx && y0 + ly0 < sy;  // the column's first voxel is in the volume
    const T* cbase = in + (colok ? gx + sx * ((y0 + ly0) + sy * z0) : 0);
#pragma unroll
    for (int k = 0; k < PT; k++) {
      inb[k] = colok && k < ccount;
      col[k] = *(inb[k] ? cbase + k * cstride : in);
    }
    const int64_t py = IS3D ? y0 + ly0 : y0 + ly0 - 1, pz = IS3D ? z0 - 1 : 0;
    const bool pok = gx < sx && py >= 0 && py < sy && pz >= 0 && pz < sz;
    pv = in[pok ? gx + sx * (py + sy * pz) : 0];
    bool hok[HV];
#pragma unroll
    for (int k = 0; k < HV; k++) hok[k] = false;
    if constexpr (HRUNS) {
      if (hcol) {  // waves holding no halo column skip the address work
        const int64_t hx = x0 + cbx - 1;
        const int64_t hy0 = ring ? y0 + cby - 1 : y0 - 1 + hby0, hz0 = ring ? z0 : z0 - 1;
        const bool fixok = hx >= 0 && hx < sx && (ring ? (hy0 >= 0 && hy0 < sy) : (hz0 >= 0 && hz0 < sz));
        const int64_t hstride = ring ? sx * sy : sx;
        // the base may lie one row before the volume (y0 = 0); only addresses
        // of voxels inside the volume are dereferenced
        const T* hbase = in + (fixok ? hx + sx * (hy0 + sy * hz0) : 0);
#pragma unroll
        for (int k = 0; k < HV; k++) {
          const bool along = ring ? (k < G::TZ && z0 + k < sz) : (k < hlen && hy0 + k >= 0 && hy0 + k < sy);
          hok[k] = fixok && along;
          hv[k] = *(hok[k] ? hbase + k * hstride : in);
        }
        if constexpr (NBS > 1) {
          const bool pbok = !ring && hby0 > 0 && fixok && hy0 - 1 >= 0 && hy0 - 1 < sy;
          hpv = *(pbok ? hbase - sx : in);
          hpv = pbok ? hpv : T(0);
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < HV; k++) {
        int bx, by, bz;
        G::halo_coords(t + NT * k < G::HALO ? t + NT * k : 0, bx, by, bz);
        const int64_t hx = x0 + bx - 1, hy = y0 + by - 1, hz = IS3D ? z0 + bz - 1 : 0;
        const bool live = t + NT * k < G::HALO;
        hok[k] = live && hx >= 0 && hx < sx && hy >= 0 && hy < sy && hz >= 0 && hz < sz;
        hv[k] = in[hok[k] ? hx + sx * (hy + sy * hz) : 0];
      }
    }
#pragma unroll
    for (int k = 0; k < PT; k++) col[k] = inb[k] ? col[k] : T(0);
    pv = pok ? pv : T(0);
#pragma unroll
    for (int k = 0; k < HV; k++) hv[k] = hok[k] ? hv[k] : T(0);
  }
  {
    uint32_t start = QK != Q_BLOCK8 ? (uint32_t)(b0 - CSTEP)
                                    : (uint32_t)G::template nid<QK>(lx + 1, ly0, 0);  // predecessor node
#pragma unroll
    for (int k = 0; k < PT; k++) {
      const T v = col[k];
      sv[b0 + k * CSTEP] = v;
      const uint32_t nk = vnode(k);
      if (fg(v) && column_link<T, MODE, NB, QK>(v, k > 0 ? col[k - 1] : pv, delta, gmin, gmax)) {
        runs |= 1u << k;
      } else {
        start = nk;
      }
      par[nk] = start;
    }
  }
  if constexpr (HRUNS) {
    // A halo voxel linked to its predecessor inside the same neighbouring
    // tile continues a run; its link then covers every other position of its
    // piece, so only run starts are queued for step 2.
    uint32_t hneed = 0;  // bit k: halo voxel k of this column is a run start
    if (hcol) {
      const int cpos = ring ? 0 : 1;  // E or K
      // Does the link to the predecessor cover every eligible position?  For
      // linkable voxels the eligible set depends on the column only (ring,
      // z >= 2), or on whether y is the last tile row (bottom plane).
      const uint32_t cmask = Nbhd<NB>::adj(cpos) | (1u << cpos);
      auto covers = [&](uint32_t el) { return (cmask & el) == (el & ((1u << Nbhd<NB>::N) - 1u)); };
      const bool cov_a = covers(ring ? halo_elig<G>(cbx, cby, 2) : halo_elig<G>(cbx, 2, 0));
      const bool cov_b = covers(halo_elig<G>(cbx, G::TY, 0));
      // a bottom segment after the first continues the run of the voxel
      // before it (staged by the previous segment's thread)
      T prev = (NBS > 1 && !ring) ? hpv : T(0);
      uint32_t start = (NBS > 1 && !ring) ? (uint32_t)G::box(cbx, hby0 > 0 ? hby0 - 1 : 0, 0) : 0u;
      uint32_t brk = 0;  // ring column: bit k = voxel k is not linked to voxel k - 1
#pragma unroll
      for (int k = 0; k < HLEN; k++) {
        if (k >= hlen) break;
        const int bx = cbx, by = ring ? cby : hby0 + k, bz = ring ? k + 1 : 0;
        const int b = G::box(bx, by, bz);
        const T v = hv[k];
        sv[b] = v;
        // predecessor in the same neighbouring tile: along z for the ring,
        // rows 1..TY of the bottom plane (the tile below) along y
        const bool linkable = ring ? k >= 1 : (by >= 2 && by <= G::TY);
        const bool linked = fg(v) && linkable && column_link<T, MODE, NB, QK>(v, prev, delta, gmin, gmax);
        const bool member = linked && (ring || by < G::TY ? cov_a : cov_b);
        if (!member) start = (uint32_t)b;
        par[b] = start;
        if (fg(v) && !member) hneed |= 1u << k;
        if (!linked) brk |= 1u << k;
        prev = v;
      }
      if (RBREAK && ring) L.h.rbreak[t] = (uint8_t)brk;
    }
    // queue the run starts: a wave-level scan and one LDS atomic per wave
    const uint32_t hc = (uint32_t)__popc(hneed);
    const uint32_t hincl = wave_incl_add(hc);
    const uint32_t htot = wave_lane63(hincl);
    if (htot) {
      uint32_t hbase = 0;
      if (lane == 63) hbase = atomicAdd(&hcount, htot);
      hbase = wave_lane63(hbase) + hincl - hc;
#pragma unroll
      for (int k = 0; k < HLEN; k++)
        if ((hneed >> k) & 1u)
          hqueue[hbase++] = (uint16_t)(ring ? G::box(cbx, cby, k + 1) : G::box(cbx, hby0 + k, 0));  // box index
    }
  } else {
#pragma unroll
    for (int k = 0; k < HV; k++) {
      const int h = t + NT * k;
      if (h < G::HALO) {
        int bx, by, bz;
        G::halo_coords(h, bx, by, bz);
        sv[G::box(bx, by, bz)] = hv[k];
        const int n = G::template nid<QK>(bx, by, bz);
        par[n] = n;
      }
    }
  }
  __syncthreads();

  // 2. halo pieces: links between halo voxels of the same neighbouring tile.
  //    Under HRUNS the link mask and union-table load of the thread's first
  //    halo run start are issued here and its unions made after step 3, which
  //    does not depend on them, so the load's latency overlaps step 3.
  uint32_t h_nh = 0, h_b = 0, h_u = 0;
  int h_bx = 0, h_by = 0, h_bz = 0;
  if constexpr (HRUNS) {
    h_nh = hcount;
    if ((uint32_t)t < h_nh) {
      G::template nid_box<Q_NONE>(hqueue[t], h_bx, h_by, h_bz);
      h_b = (uint32_t)G::box(h_bx, h_by, h_bz);
      constexpr uint32_t ALLN = (1u << Nbhd<NB>::N) - 1u;
      const uint32_t lk =
        link_mask<T, MODE, NB, G>(sv, (int)h_b, sv[h_b], halo_elig<G>(h_bx, h_by, h_bz) & ALLN, false);
      h_u = ut0[lk];  // class 0: halo pairs of one neighbouring tile are all unioned here
    }
  } else if constexpr (PIECES) {
    for (int h = t; h < G::HALO; h += NT) {
      int bx, by, bz;
      G::halo_coords(h, bx, by, bz);
      const int b = G::box(bx, by, bz);
      const T cur = sv[b];
      if (!fg(cur)) continue;
      link_voxel<T, MODE, NB, QK, G, true>(sv, par, b, (uint32_t)G::template nid<QK>(bx, by, bz), bx, by, bz, cur,
                                           Border{}, ut0, halo_elig<G>(bx, by, bz), false, delta, gmin, gmax, false);
    }
  }

  // 3. counters over the column and the work queue: voxels still needing
  //    neighbour unions (run starts, and run members whose column link does
  //    not cover their neighbourhood)
  // per-thread counts fit 32 bits; rows are tile-local (ly + TY * lz), whose
  // order is the order of the global rows y + sy * z
  uint32_t my_epl = 0, my_prov = 0;
  int my_first = 0x7FFFFFFF, my_last = -1;
  uint32_t need = 0;  // bit k: voxel k is queued
  uint32_t qs = 0;    // this thread's first queue slot
  {
    const bool xlo = lx == 0, xhi = lx == G::TX - 1;
    // the column link of voxel 0 leaves the tile (3D: the z-1 halo plane; 2D:
    // the halo row above for the first column group)
    const Border bd0 = make_border<IS3D>(xlo, xhi, IS3D ? ly0 == 0 : ly0 == 0, IS3D && ly0 == G::TY - 1, IS3D);
    const Border bdc = make_border<IS3D>(xlo, xhi, IS3D ? ly0 == 0 : false, IS3D && ly0 == G::TY - 1, false);
    constexpr uint32_t ALLN = (1u << Nbhd<NB>::N) - 1u;
    const bool covers_all0 = (cover_from<NB>(0, bd0) | 1u) == ALLN;
    const bool covers_all = (cover_from<NB>(0, bdc) | 1u) == ALLN;
    // 26-connected run member p at k >= 1 of a border column: the positions
    // its column link leaves uncovered are halo voxels q of ring columns.  If
    // q is linked to the voxel q' below it in its ring column (so q ~ q'),
    // then q' is linked to p's predecessor E at the same offset (values are
    // equal), E ~ q' holds by the same argument one voxel down (or by E's
    // own unions), and p ~ E: the union p - q is redundant.  So p needs
    // unions only where some uncovered q sits at a break of its ring column
    // (bit k of the column's mask for q at p's z, bit k - 1 one plane below;
    // bit 0 always set, as ring voxel 0 is not joined to the plane below).
    uint32_t rcand = ~0u;  // bit k: a border run member at k may need unions
    if constexpr (RBREAK) {
      if (!covers_all) {
        constexpr uint32_t PX = packed_offsets<true>(0), PY = packed_offsets<true>(1), PZ = packed_offsets<true>(2);
        uint32_t un = ALLN & ~(cover_from<NB>(0, bdc) | 1u);
        rcand = 0;
        while (un) {
          const int i = __builtin_ctz(un);
          un &= un - 1u;
          const int cbx = lx + (int)((PX >> (2 * i)) & 3u);  // box x: lx + 1 + dx
          const int cby = ly0 + (int)((PY >> (2 * i)) & 3u);
          const bool below = ((PZ >> (2 * i)) & 3u) == 0u;  // dz = -1
          const int r = cby == 0 ? cbx : cby == G::BY - 1 ? G::BX + cbx : 2 * G::BX + 2 * (cby - 1) + (cbx != 0);
          const uint32_t m = L.h.rbreak[r];
          rcand |= below ? (m << 1) : m;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < PT; k++) {
      const T cur = col[k];
      if (!fg(cur)) continue;
      int vx, vy, vz;
      G::voxel(t, k, vx, vy, vz);
      if (cur != sv[b0 + k * CSTEP - 1]) my_epl++;
      const int row = vy + G::TY * vz;
      my_first = row < my_first ? row : my_first;
      my_last = row > my_last ? row : my_last;
      bool q = !((runs >> k) & 1u) ||
               !column_done<T, MODE, NB, QK>(cur, k > 0 ? col[k - 1] : pv, k > 0 ? covers_all : covers_all0,
                                             delta, gmin, gmax);
      if constexpr (RBREAK) {
        if (q && k > 0 && ((runs >> k) & 1u)) q = ((rcand >> k) & 1u) != 0u;
      }
      if constexpr (NB == 6 && MODE != M_CONT) {
        // a 6-connected run member is queued only for a union it still needs
        if (q && ((runs >> k) & 1u)) {
          const Border bd = make_border<true>(vx == 0, vx == G::TX - 1, vy == 0, vy == G::TY - 1, vz == 0);
          q = (six_links<T, MODE, G, false>(sv, b0 + k * CSTEP, cur, bd, 0u, true) & 3u) != 0u;
        }
      }
      if (q) need |= 1u << k;
    }
    // wave-level exclusive scan of the counts, one LDS atomic per wave
    const uint32_t cnt = (uint32_t)__popc(need);
    const uint32_t incl = wave_incl_add(cnt);
    uint32_t base = 0;
    if (lane == 63) base = atomicAdd(&qcount, incl);
    qs = wave_lane63(base) + incl - cnt;
  }
  if constexpr (HRUNS) {
    // unions of the halo run start prepared in step 2, then any further ones
    if ((uint32_t)t < h_nh) {
      uint32_t a = h_b, u = h_u;
      while (u) {
        const int i = __builtin_ctz(u);
        u &= u - 1u;
        a = lunion(par, a, node_at<QK, G>(h_b, h_bx, h_by, h_bz, i));
      }
    }
#pragma unroll 1
    for (uint32_t q = t + NT; q < h_nh; q += NT) {
      int bx, by, bz;
      G::template nid_box<Q_NONE>(hqueue[q], bx, by, bz);
      const int b = G::box(bx, by, bz);
      link_voxel<T, MODE, NB, QK, G, true>(sv, par, b, (uint32_t)b, bx, by, bz, sv[b], Border{}, ut0,
                                           halo_elig<G>(bx, by, bz), false, delta, gmin, gmax, false);
    }
  }
  __syncthreads();

  // The queue and the root bitmask share LDS with hqueue, whose last reads
  // are above.  Voxels past the queue's capacity stay with their thread.
  uint32_t over = 0;  // bit k: voxel k did not fit the queue
#pragma unroll
  for (int k = 0; k < PT; k++) {
    if (!((need >> k) & 1u)) continue;
    int vx, vy, vz;
    G::voxel(t, k, vx, vy, vz);
    if (qs < QC) queue[qs] = (uint16_t)(G::pos(vx, vy, vz) | (((runs >> k) & 1u) << 15));
    else over |= 1u << k;
    qs++;
  }
  for (int w = t; w < RW; w += NT) L.w.rbits[w] = 0;

  // halo pieces are the roots left by step 2 (without pieces every
  // foreground halo voxel stands for itself).  Entry i of this thread: halo
  // column voxel i under HRUNS, else halo voxel t + NT * i.
  auto hnode = [&](int i) -> uint32_t {
    if (HRUNS) return (uint32_t)G::box(cbx, ring ? cby : hby0 + i, ring ? i + 1 : 0);
    int bx, by, bz;
    G::halo_coords(t + NT * i, bx, by, bz);
    return (uint32_t)G::template nid<QK>(bx, by, bz);
  };
  auto hindex = [&](int i) -> uint32_t {
    if (HRUNS) return (uint32_t)(ring ? G::BX * G::BY + i * RING + t : cbx + G::BX * (hby0 + i));
    return (uint32_t)(t + NT * i);
  };
  uint32_t hpiece = 0;
#pragma unroll
  for (int i = 0; i < HV; i++) {
    if (fg(hv[i])) {  // hv is background outside the halo and the volume
      const uint32_t n = hnode(i);
      if (!PIECES || par[n] == n) hpiece |= 1u << i;
    }
  }
  const uint32_t nQ = qcount < QC ? qcount : QC;
  __syncthreads();

  // 4. queued voxels
  if constexpr (UT) {
    // Transitive connectivities: link mask, then the union table.
    struct Item {
      uint32_t pn, lk, u, cm;
      int bx, by, bz;
    };
    auto prep_at = [&](int ps, bool colq) -> Item {
      const int vx = ps & (G::TX - 1), vy = (ps / G::TX) & (G::TY - 1), vz = ps / (G::TX * G::TY);
      Item it;
      it.bx = vx + 1;
      it.by = vy + 1;
      it.bz = IS3D ? vz + 1 : 0;
      const int b = G::box(it.bx, it.by, it.bz);
      it.pn = (uint32_t)G::template nid<QK>(it.bx, it.by, it.bz);
      it.lk = link_mask<T, MODE, NB, G>(sv, b, sv[b], (1u << Nbhd<NB>::N) - 1u, colq);
      const int cls = border_class(vx == 0, vx == G::TX - 1, vy == 0, IS3D && vy == G::TY - 1, IS3D && vz == 0);
      it.u = ut0[(cls << Nbhd<NB>::N) | it.lk];
      it.cm = colq ? ~1u : ~0u;  // a run member's link to neighbour 0 is in place
      return it;
    };
    auto prep = [&](uint32_t q) -> Item {
      const uint32_t e = queue[q];
      return prep_at((int)(e & 0x7FFFu), (e >> 15) != 0u);
    };
    auto unions = [&](const Item& it) {
      uint32_t a = it.pn, u = it.u & it.cm;
      while (u) {
        const int i = __builtin_ctz(u);
        u &= u - 1u;
        a = lunion(par, a, node_at<QK, G>(it.pn, it.bx, it.by, it.bz, i));
      }
      if (QK != Q_BLOCK8 && it.lk == 0) my_prov++;  // block labels are counted separately
    };
    // Two entries per round: the second one's table load is issued first and
    // the first one's right after, so both latencies overlap.
#pragma unroll 1
    for (uint32_t q = t; q < nQ; q += 2 * NT) {
      const uint32_t q2 = q + NT;
      Item i2;  // left unset past the end
      if (q2 < nQ) i2 = prep(q2);
      const Item i1 = prep(q);
      unions(i1);
      if (q2 < nQ) unions(i2);
    }
#pragma unroll 1
    while (over) {
      const int k = __builtin_ctz(over);
      over &= over - 1u;
      int vx, vy, vz;
      G::voxel(t, k, vx, vy, vz);
      unions(prep_at(G::pos(vx, vy, vz), ((runs >> k) & 1u) != 0u));
    }
  } else {
    auto one = [&](int ps, bool colq) {
      const int vx = ps & (G::TX - 1), vy = (ps / G::TX) & (G::TY - 1), vz = ps / (G::TX * G::TY);
      const int bx = vx + 1, by = vy + 1, bz = IS3D ? vz + 1 : 0;
      const int b = G::box(bx, by, bz);
      const T cur = sv[b];
      const bool xlo = vx == 0, xhi = vx == G::TX - 1, ylo = vy == 0, yhi = IS3D && vy == G::TY - 1;
      const bool zlo = IS3D && vz == 0;
      const Border bd = make_border<IS3D>(xlo, xhi, ylo, yhi, zlo);
      const uint16_t* ut = ut0 + (border_class(xlo, xhi, ylo, yhi, zlo) << Nbhd<NB>::N);
      const uint32_t pn = (uint32_t)G::template nid<QK>(bx, by, bz);
      const int nl = link_voxel<T, MODE, NB, QK, G, false>(sv, par, b, pn, bx, by, bz, cur, bd, ut, 0u, colq, delta,
                                                            gmin, gmax, y0 + vy == 0);
      if (QK != Q_BLOCK8 && nl == 0) my_prov++;  // block labels are counted separately
    };
#pragma unroll 1
    for (uint32_t q = t; q < nQ; q += NT) {
      const uint32_t e = queue[q];
      one((int)(e & 0x7FFFu), (e >> 15) != 0u);
    }
#pragma unroll 1
    while (over) {
      const int k = __builtin_ctz(over);
      over &= over - 1u;
      int vx, vy, vz;
      G::voxel(t, k, vx, vy, vz);
      one(G::pos(vx, vy, vz), ((runs >> k) & 1u) != 0u);
    }
  }
  __syncthreads();

  // 5. resolve roots and mark the roots that own tile voxels.  The column's
  //    run starts and the thread's halo pieces walk to their roots in
  //    lockstep, one parent hop per round for all of them, so their LDS
  //    latencies overlap; run members share their run start's root.  Roots
  //    are fixed points, so a round may hop every slot; unused slots sit on
  //    the node of box position (0, 0, 0), the smallest node in use, which is
  //    a root (a parent never exceeds its node).  No lane writes par here, so
  //    plain loads suffice; a path is shorter than the node count, which
  //    bounds the rounds.
  constexpr uint32_t ROOT0 = (uint32_t)G::template nid<QK>(0, 0, 0);
  uint32_t fa[PT], fh[HV];
#pragma unroll
  for (int k = 0; k < PT; k++)
    fa[k] = (fg(col[k]) && !(k > 0 && ((runs >> k) & 1u))) ? vnode(k) : ROOT0;
#pragma unroll
  for (int i = 0; i < HV; i++) fh[i] = ((hpiece >> i) & 1u) ? hnode(i) : ROOT0;
  const bool wave_halo = __any(hpiece != 0u);  // waves without halo pieces skip those slots
  bool more = true;
#pragma unroll 1
  for (int round = 0; more && round < G::template NIDS<QK>; round++) {
    more = false;
    // all loads of the round are issued before the first compare
    uint32_t pa[PT];
#pragma unroll
    for (int k = 0; k < PT; k++) pa[k] = par[fa[k]];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < PT; k++) {
      more |= pa[k] != fa[k];
      fa[k] = pa[k];
    }
    if (wave_halo) {
      uint32_t ph[HV];
#pragma unroll
      for (int i = 0; i < HV; i++) ph[i] = par[fh[i]];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < HV; i++) {
        more |= ph[i] != fh[i];
        fh[i] = ph[i];
      }
    }
  }
  uint32_t root[PT];
#pragma unroll
  for (int k = 0; k < PT; k++) {
    if (!fg(col[k])) root[k] = NONE;
    else if (k > 0 && ((runs >> k) & 1u)) root[k] = root[k - 1];
    else root[k] = fa[k];
  }
  uint32_t hroot[HV];
#pragma unroll
  for (int i = 0; i < HV; i++) hroot[i] = ((hpiece >> i) & 1u) ? fh[i] : NONE;
#pragma unroll
  for (int k = 0; k < PT; k++) {
    const uint32_t r = root[k];
    // most columns of a component share its root: skip the atomic once set
    if (r != NONE && (k == 0 || r != root[k - 1]) && !((lload(&L.w.rbits[r >> 5]) >> (r & 31)) & 1u))
      atomicOr(&L.w.rbits[r >> 5], 1u << (r & 31));
  }
  __syncthreads();

  // 6. rank the marked roots in node (= key) order: one wave scans the bitmask
  if (wv == 0) {
    const int w0 = WPL * lane;
    uint32_t c[WPL];
    uint32_t own = 0;
#pragma unroll
    for (int i = 0; i < WPL; i++) {
      c[i] = w0 + i < RW ? (uint32_t)__popc(L.w.rbits[w0 + i]) : 0u;
      own += c[i];
    }
    const uint32_t incl = wave_incl_add(own);
    uint32_t ex = incl - own;
#pragma unroll
    for (int i = 0; i < WPL; i++) {
      if (w0 + i < RW) L.late.rprefix[w0 + i] = ex;
      ex += c[i];
    }
    if (lane == 63) L.late.nroots = incl;
  }
  __syncthreads();
  const uint32_t R = L.late.nroots;
  auto rank = [&](uint32_t r) -> uint32_t {
    return L.late.rprefix[r >> 5] + (uint32_t)__popc(L.w.rbits[r >> 5] & ((1u << (r & 31)) - 1u));
  };
  const uint32_t root0 = tile_root0<G>(sx, sy, sz, tx, ty, tz, keybase);
  uint32_t* RLt = p.RL + root0;
  // Keys are below 2^32, so they can be formed in 32-bit arithmetic: the
  // wrap-around of the intermediate terms cancels.  Box position (bx, by, bz)
  // is voxel (x0 + bx - 1, y0 + by - 1, z0 + bz - 1).
  const uint32_t kbox0 = (uint32_t)(x0 - 1) + (uint32_t)sx * ((uint32_t)(y0 - 1) + (uint32_t)sy * (uint32_t)(z0 - 1));
  for (int w = t; w < RW; w += NT) {
    uint32_t bits = L.w.rbits[w];
    uint32_t rk = L.late.rprefix[w];
    while (bits) {
      const int bpos = __builtin_ctz(bits);
      bits &= bits - 1u;
      int bx, by, bz;
      G::template nid_box<QK>(32 * w + bpos, bx, by, bz);
      uint32_t key;
      if (QK == Q_BLOCK8)
        key = gkey<QK>(x0 + bx - 1, y0 + by - 1, 0, sx, sy, p.osx) + keybase;
      else if (IS3D)
        key = kbox0 + (uint32_t)bx + (uint32_t)sx * ((uint32_t)by + (uint32_t)sy * (uint32_t)bz) + keybase;
      else
        key = kbox0 + (uint32_t)sx * (uint32_t)sy + (uint32_t)bx + (uint32_t)sx * (uint32_t)by + keybase;
      p.GP[root0 + rk] = root0 + rk;
      RLt[rk++] = key;
      atomicOr(p.bits + (key >> 5), 1u << (key & 31));
    }
  }

  // 7. per-voxel root index, thread-major: one 8- or 16-byte store per thread
  static_assert(PT == 8 || PT == 4, "J stores take 4 or 8 voxels per thread");
  uint32_t jv[PT];
#pragma unroll
  for (int k = 0; k < PT; k++)
    jv[k] = root[k] == NONE ? 0u : (k > 0 && root[k] == root[k - 1]) ? jv[k > 0 ? k - 1 : 0] : rank(root[k]) + 1u;
  uint32_t jw[PT / 2];  // 16-bit entries, two per word
#pragma unroll
  for (int i = 0; i < PT / 2; i++) jw[i] = jv[2 * i] | (jv[2 * i + 1] << 16);
  if (R <= 255) {
    uint32_t w8[PT / 4];  // 8-bit entries, four per word
#pragma unroll
    for (int i = 0; i < PT / 4; i++) w8[i] = (jw[2 * i] & 0xFFu) | ((jw[2 * i] >> 8) & 0xFF00u) |
                                             ((jw[2 * i + 1] & 0xFFu) << 16) | ((jw[2 * i + 1] >> 16) << 24);
    if constexpr (PT == 8) reinterpret_cast<uint2*>(p.J8 + slot * G::TV)[t] = make_uint2(w8[0], w8[PT / 4 - 1]);
    else reinterpret_cast<uint32_t*>(p.J8 + slot * G::TV)[t] = w8[0];
  } else {
    if constexpr (PT == 8)
      reinterpret_cast<uint4*>(p.J16 + slot * G::TV)[t] = make_uint4(jw[0], jw[1], jw[PT / 2 - 2], jw[PT / 2 - 1]);
    else reinterpret_cast<uint2*>(p.J16 + slot * G::TV)[t] = make_uint2(jw[0], jw[PT / 2 - 1]);
  }

  // 8. cross-tile requests: one per halo piece whose tree owns tile voxels,
  //    as (root node, halo index)
  //    slots from a wave-level scan and one LDS atomic per wave
  uint32_t rq = 0;  // bit i: halo entry i makes a request
#pragma unroll
  for (int i = 0; i < HV; i++) {
    const uint32_t r = hroot[i];
    if (r != NONE && ((L.w.rbits[r >> 5] >> (r & 31)) & 1u)) rq |= 1u << i;
  }
  if (__any(rq != 0u)) {
    const uint32_t rc = (uint32_t)__popc(rq);
    const uint32_t rincl = wave_incl_add(rc);
    uint32_t rbase = 0;
    if (lane == 63) rbase = atomicAdd(&nreq, rincl);
    uint32_t q = wave_lane63(rbase) + rincl - rc;
    uint32_t* REQt = p.REQ + slot * G::HALO;
#pragma unroll
    for (int i = 0; i < HV; i++)
      if ((rq >> i) & 1u) REQt[q++] = (rank(hroot[i]) << 16) | hindex(i);
  }

  // 9. counters
  my_epl = wave_lane63(wave_incl_add(my_epl));
  my_prov = wave_lane63(wave_incl_add(my_prov));
  my_first = (int)wave_lane63((uint32_t)wave_incl_min(my_first));
  my_last = (int)wave_lane63((uint32_t)wave_incl_max(my_last));
  if (lane == 0) {
    L.late.red_epl[wv] = my_epl;
    L.late.red_prov[wv] = my_prov;
    L.late.red_first[wv] = my_first;
    L.late.red_last[wv] = my_last;
  }
  __syncthreads();
  if (t == 0) {
    unsigned long long e = 0, pv2 = 0;
    long long f = 0x7FFFFFFFFFFFFFFFll, l = -1;
    int fl = 0x7FFFFFFF, ll = -1;
    for (int w = 0; w < NW; w++) {
      e += L.late.red_epl[w];
      pv2 += L.late.red_prov[w];
      fl = (int)L.late.red_first[w] < fl ? (int)L.late.red_first[w] : fl;
      ll = (int)L.late.red_last[w] > ll ? (int)L.late.red_last[w] : ll;
    }
    if (ll >= 0) {  // tile-local rows to global rows
      f = (y0 + fl % G::TY) + sy * (IS3D ? z0 + fl / G::TY : 0);
      l = (y0 + ll % G::TY) + sy * (IS3D ? z0 + ll / G::TY : 0);
    }
    p.TR[slot] = R;
    p.TQ[slot] = nreq;
    TileCnt c;
    c.epl = (uint32_t)e;
    c.prov = (uint32_t)pv2;
    c.first_row = f;
    c.last_row = l;
    p.tcnt[slot] = c;
  }
}

template <class T, int MODE, int NB, int QK, class G>
__global__ __launch_bounds__(G::NT, (local_waves<T, G, full_queue<MODE, NB>()>())) void k_local(LocalParams p) {
  __shared__ LocalLds<T, G, full_queue<MODE, NB>()> L;
  const int64_t tile = xcd_tile(blockIdx.x, p.ntiles);
  local_tile<T, MODE, NB, QK, G>(p, tile, tile, 0u, L);
}

// The volumes of a batch that share a connectivity, in one launch: their
// tiles are contiguous in the batch, [gtile0, gtile0 + gtiles); each
// workgroup reads its volume's table entry from vix.  p carries the shared per-tile
// arrays; the geometry fields come from the table.  Queue capacity and
// occupancy are those of the single-volume kernel of the same connectivity.
template <class T, int MODE, int NB, class G, int QK = Q_NONE>
__global__ __launch_bounds__(G::NT, (local_waves<T, G, full_queue<MODE, NB>()>())) void k_local_batch(
  LocalParams p, const BatchVol* vols, const uint32_t* vix, int64_t gtile0, int64_t gtiles) {
  __shared__ LocalLds<T, G, full_queue<MODE, NB>()> L;
  const int64_t slot = gtile0 + xcd_tile(blockIdx.x, gtiles);
  const BatchVol& v = vols[vix[slot]];
  LocalParams q = p;
  q.in = v.in;
  q.sx = v.sx;
  q.sy = v.sy;
  q.sz = v.sz;
  q.osx = v.osx;
  q.ntx = v.ntx;
  q.nty = v.nty;
  q.ntz = v.ntz;
  q.ntiles = v.ntiles;
  q.minmax = v.minmax;
  local_tile<T, MODE, NB, QK, G, full_queue<MODE, NB>()>(q, slot - v.tile0, slot, v.key0, L);
}

// Launchers (one translation unit per mode keeps compile times short).
// tall: 3D tiles of Geo3T instead of Geo3.
hipError_t launch_local_multi(int dt, int nb, int qk, bool tall, const LocalParams& p, hipStream_t s);
hipError_t launch_local_binary(int dt, int nb, int qk, bool tall, const LocalParams& p, hipStream_t s);
hipError_t launch_local_cont(int dt, int nb, int qk, bool tall, const LocalParams& p, hipStream_t s);
// A batch launches once per connectivity: tiles [gtile0, gtile0 + gtiles) all
// belong to volumes of connectivity nb.
hipError_t launch_local_batch_multi(int dt, int nb, bool tall, const LocalParams& p, const BatchVol* vols,
                                    const uint32_t* vix, int64_t gtile0, int64_t gtiles, hipStream_t s);
hipError_t launch_local_batch_binary(int dt, int nb, bool tall, const LocalParams& p, const BatchVol* vols,
                                     const uint32_t* vix, int64_t gtile0, int64_t gtiles, hipStream_t s);
hipError_t launch_local_batch_cont(int dt, int nb, bool tall, const LocalParams& p, const BatchVol* vols,
                                   const uint32_t* vix, int64_t gtile0, int64_t gtiles, hipStream_t s);
// Batch volumes on the 2D tiles (single planes whose numbering or links need
// the 2D kernel): binary 2D-4 (Q_BIN4) and 2D-8 (Q_BLOCK8), continuous 2D-8
// (Q_CONT8).
hipError_t launch_local_batch_plane_binary(int dt, int nb, const LocalParams& p, const BatchVol* vols,
                                           const uint32_t* vix, int64_t gtile0, int64_t gtiles, hipStream_t s);
hipError_t launch_local_batch_plane_cont(int dt, int nb, const LocalParams& p, const BatchVol* vols,
                                         const uint32_t* vix, int64_t gtile0, int64_t gtiles, hipStream_t s);

template <int MODE>
hipError_t launch_local_batch_plane_mode(int dt, int nb, const LocalParams& p, const BatchVol* vols,
                                         const uint32_t* vix, int64_t gtile0, int64_t gtiles, hipStream_t s) {
  const dim3 grid((unsigned)gtiles), block(Geo2::NT);
  constexpr int Q8 = MODE == M_BINARY ? Q_BLOCK8 : MODE == M_CONT ? Q_CONT8 : Q_NONE;
  constexpr int Q4 = MODE == M_BINARY ? Q_BIN4 : Q_NONE;
#define CC3D_P(T)                                                                                                 \
  do {                                                                                                            \
    if (nb == 8 && Q8 != Q_NONE)                                                                                  \
      hipLaunchKernelGGL((k_local_batch<T, MODE, 8, Geo2, Q8>), grid, block, 0, s, p, vols, vix, gtile0, gtiles); \
    else if (nb == 4 && Q4 != Q_NONE)                                                                             \
      hipLaunchKernelGGL((k_local_batch<T, MODE, 4, Geo2, Q4>), grid, block, 0, s, p, vols, vix, gtile0, gtiles); \
    else                                                                                                          \
      return hipErrorInvalidValue;                                                                                \
  } while (0)
  switch (dt) {
    case DT_U8: CC3D_P(uint8_t); break;
    case DT_U16: CC3D_P(uint16_t); break;
    case DT_U32: CC3D_P(uint32_t); break;
    case DT_U64: CC3D_P(uint64_t); break;
    case DT_F32: CC3D_P(float); break;
    case DT_F64: CC3D_P(double); break;
    default: return hipErrorInvalidValue;
  }
#undef CC3D_P
  return hipGetLastError();
}

template <int MODE, class G>
hipError_t launch_local_batch_geo(int dt, int nb, const LocalParams& p, const BatchVol* vols, const uint32_t* vix,
                                  int64_t gtile0, int64_t gtiles, hipStream_t s) {
  const dim3 grid((unsigned)gtiles), block(G::NT);
#define CC3D_B(T, NB) \
  hipLaunchKernelGGL((k_local_batch<T, MODE, NB, G>), grid, block, 0, s, p, vols, vix, gtile0, gtiles)
#define CC3D_BNB(T)                          \
  switch (nb) {                              \
    case 26: CC3D_B(T, 26); break;           \
    case 18: CC3D_B(T, 18); break;           \
    case 6: CC3D_B(T, 6); break;             \
    default: return hipErrorInvalidValue;    \
  }
  switch (dt) {
    case DT_U8: CC3D_BNB(uint8_t); break;
    case DT_U16: CC3D_BNB(uint16_t); break;
    case DT_U32: CC3D_BNB(uint32_t); break;
    case DT_U64: CC3D_BNB(uint64_t); break;
    case DT_F32: CC3D_BNB(float); break;
    case DT_F64: CC3D_BNB(double); break;
    default: return hipErrorInvalidValue;
  }
#undef CC3D_BNB
#undef CC3D_B
  return hipGetLastError();
}
template <int MODE>
hipError_t launch_local_batch_mode(int dt, int nb, bool tall, const LocalParams& p, const BatchVol* vols,
                                   const uint32_t* vix, int64_t gtile0, int64_t gtiles, hipStream_t s) {
  return tall ? launch_local_batch_geo<MODE, Geo3T>(dt, nb, p, vols, vix, gtile0, gtiles, s)
              : launch_local_batch_geo<MODE, Geo3>(dt, nb, p, vols, vix, gtile0, gtiles, s);
}

template <int MODE, class G3>
hipError_t launch_local_geo(int dt, int nb, int qk, const LocalParams& p, hipStream_t s) {
  const dim3 grid((unsigned)p.ntiles);
#define CC3D_L(T, NB, QK, G) \
  hipLaunchKernelGGL((k_local<T, MODE, NB, QK, G>), grid, dim3(G::NT), 0, s, p)
#define CC3D_NB(T)                                                               \
  switch (nb) {                                                                  \
    case 26: CC3D_L(T, 26, Q_NONE, G3); break;                                   \
    case 18: CC3D_L(T, 18, Q_NONE, G3); break;                                   \
    case 6: CC3D_L(T, 6, Q_NONE, G3); break;                                     \
    case 8:                                                                      \
      if (MODE == M_BINARY && qk == Q_BLOCK8) CC3D_L(T, 8, (MODE == M_BINARY ? Q_BLOCK8 : Q_NONE), Geo2); \
      else if (MODE == M_CONT && qk == Q_CONT8) CC3D_L(T, 8, (MODE == M_CONT ? Q_CONT8 : Q_NONE), Geo2); \
      else CC3D_L(T, 8, Q_NONE, Geo2);                                           \
      break;                                                                     \
    case 4:                                                                      \
      if (MODE == M_BINARY && qk == Q_BIN4) CC3D_L(T, 4, (MODE == M_BINARY ? Q_BIN4 : Q_NONE), Geo2); \
      else CC3D_L(T, 4, Q_NONE, Geo2);                                           \
      break;                                                                     \
    default: return hipErrorInvalidValue;                                        \
  }
  switch (dt) {
    case DT_U8: CC3D_NB(uint8_t); break;
    case DT_U16: CC3D_NB(uint16_t); break;
    case DT_U32: CC3D_NB(uint32_t); break;
    case DT_U64: CC3D_NB(uint64_t); break;
    case DT_F32: CC3D_NB(float); break;
    case DT_F64: CC3D_NB(double); break;
    default: return hipErrorInvalidValue;
  }
#undef CC3D_NB
#undef CC3D_L
  return hipGetLastError();
}
template <int MODE>
hipError_t launch_local_mode(int dt, int nb, int qk, bool tall, const LocalParams& p, hipStream_t s) {
  return tall ? launch_local_geo<MODE, Geo3T>(dt, nb, qk, p, s) : launch_local_geo<MODE, Geo3>(dt, nb, qk, p, s);
}

}  // namespace cc3d_amd
