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
// Covering: a linked backward neighbour need not be unioned when it is
// adjacent to another neighbour already connected to p and linked to it,
// because that pair's own union (made when the later of the two was
// processed) connects them.  For multilabel/binary links are transitive, so
// adjacency suffices; for continuous links it takes an exactly equal value or
// a direct match with the first neighbour.  6/4-connectivity uses the
// phantom voxels J, D, B (2D: J) of the 2x2(x2) cube.
//
// Outputs per tile:
//   RL[tile*TV + r]  global key of local root r (roots in ballot order)
//   TR[tile]         number of local roots R
//   J8/J16           per voxel: 0 (background) or 1 + index of its local root;
//                    8-bit when R <= 255, else 16-bit
//   GP[key]          initialised to key for every local root
//   REQ/TQ           cross-tile union requests: (root position, halo index)
//                    for each halo "piece" (halo voxels connected through links
//                    that stay inside one neighbouring tile) touching a local
//                    component
//   counters         epl transitions, provisional-label count, first/last
//                    foreground row (cc3d/cc3d.hpp:287-315)
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
  uint32_t* TR;
  uint8_t* J8;
  uint16_t* J16;
  uint32_t* REQ;
  uint32_t* TQ;
  Counters* ctr;
};

template <class G>
__device__ inline int tile_slot(int bx, int by, int bz) {
  const int sxl = bx == 0 ? 0 : (bx <= G::TX ? 1 : 2);
  const int syl = by == 0 ? 0 : (by <= G::TY ? 1 : 2);
  const int szl = G::IS3D ? (bz == 0 ? 0 : 1) : 0;
  return sxl + 3 * (syl + 3 * szl);
}

template <class G>
__device__ inline bool in_box(int bx, int by, int bz) {
  return bx >= 0 && bx < G::BX && by >= 0 && by < G::BY && bz >= 0 && bz < G::BZ;
}

// Raster order of two box positions (z, then y, then x).
__device__ inline bool box_later(int ax, int ay, int az, int bx, int by, int bz) {
  if (az != bz) return az > bz;
  if (ay != by) return ay > by;
  return ax > bx;
}

// Is the adjacent pair (a, b) unioned inside this tile's structure: by the
// tile phase (its later voxel is a tile voxel) or by the halo phase (both are
// halo voxels of the same neighbouring tile)?
template <class G>
__device__ inline bool pair_here(int ax, int ay, int az, int bx, int by, int bz) {
  const bool ta = G::is_tile(ax, ay, az), tb = G::is_tile(bx, by, bz);
  if (box_later(ax, ay, az, bx, by, bz) ? ta : tb) return true;
  return !ta && !tb && tile_slot<G>(ax, ay, az) == tile_slot<G>(bx, by, bz);
}

// All backward neighbours of this tile voxel lie inside the tile.
template <class G, int NB>
__device__ inline bool interior_voxel(int bx, int by, int bz) {
  if constexpr (NB == 26 || NB == 18)
    return bx >= 2 && bx <= G::TX - 1 && by >= 2 && by <= G::TY - 1 && bz >= 2;
  else if constexpr (NB == 6)
    return bx >= 2 && by >= 2 && bz >= 2;
  else if constexpr (NB == 8)
    return bx >= 2 && bx <= G::TX - 1 && by >= 2;
  else
    return bx >= 2 && by >= 2;
}


// Unions for one foreground voxel at box (bx,by,bz).  HALO: restrict to halo
// neighbours in the same neighbouring tile.  Returns the number of linked
// neighbours found (0 <=> the reference gives this voxel a new label).
template <class T, int MODE, int NB, int QK, class G, bool HALO>
__device__ inline int link_voxel(const T* sv, uint32_t* par, int bx, int by, int bz, T cur,
                                 T delta, T gmin, T gmax, int64_t gy, int slot) {
  using NH = Nbhd<NB>;
  const uint32_t pn = G::template node_box<QK>(bx, by, bz);
  auto ok = [&](int qx, int qy, int qz) -> bool {
    if (!HALO) return true;
    return in_box<G>(qx, qy, qz) && !G::is_tile(qx, qy, qz) && tile_slot<G>(qx, qy, qz) == slot;
  };
  auto val = [&](int qx, int qy, int qz) -> T { return sv[G::box(qx, qy, qz)]; };
  auto nd = [&](int qx, int qy, int qz) -> uint32_t {
    return G::template node_box<QK>(qx, qy, qz);
  };

  if constexpr (QK == Q_BIN4) {
    // cc3d/cc3d_binary.hpp:822-935
    if (gy == 0) {
      if (val(bx - 1, by, bz) == cur) {
        lunion(par, pn, nd(bx - 1, by, bz));
        return 1;
      }
      return 0;
    }
    const bool C = fg(val(bx, by - 1, bz));
    const bool B = fg(val(bx - 1, by, bz));
    const bool D = fg(val(bx - 1, by - 1, bz));
    int n = 0;
    if (C) {
      lunion(par, pn, nd(bx, by - 1, bz));
      n++;
    }
    if (B && !(C && D)) {
      lunion(par, pn, nd(bx - 1, by, bz));
      n++;
    }
    return n;
  } else if constexpr (QK == Q_CONT8) {
    // cc3d/cc3d_continuous.hpp:335-384
    const T vb = val(bx, by - 1, bz);
    if (vb == cur) {
      lunion(par, pn, nd(bx, by - 1, bz));
      return 1;
    }
    if (fg(vb)) {
      const T lo = (vb < cur) ? vb : cur;
      const T hi = (cur < vb) ? vb : cur;
      const T d1 = lo - gmin, d2 = gmax - hi;
      if (d1 <= delta && d2 <= delta) {
        lunion(par, pn, nd(bx, by - 1, bz));
        return 1;
      }
    }
    int n = 0;
    const int qxs[4] = {bx, bx - 1, bx + 1, bx - 1};
    const int qys[4] = {by - 1, by - 1, by - 1, by};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const T q = val(qxs[i], qys[i], bz);
      if (fg(q) && match(cur, q, delta)) {
        lunion(par, pn, nd(qxs[i], qys[i], bz));
        n++;
      }
    }
    return n;
  } else if constexpr (MODE != M_CONT && (NB == 6 || NB == 4)) {
    // face connectivity: M, K (and E) with phantoms J, D, B
    auto lk = [&](int qx, int qy, int qz) -> bool {
      if (!ok(qx, qy, qz)) return false;
      const T q = val(qx, qy, qz);
      return MODE == M_MULTI ? (q == cur) : fg(q);
    };
    // a phantom path a-c-b connects a and b when both of its pairs are
    // processed in this tile's structure
    auto path = [&](int ax, int ay, int az, int cx, int cy, int cz, int qx, int qy, int qz) -> bool {
      return HALO || (pair_here<G>(ax, ay, az, cx, cy, cz) && pair_here<G>(cx, cy, cz, qx, qy, qz));
    };
    const bool lM = lk(bx - 1, by, bz);
    const bool lK = lk(bx, by - 1, bz);
    int n = 0;
    if (lM) {
      lunion(par, pn, nd(bx - 1, by, bz));
      n++;
    }
    if (lK) {
      if (!(lM && lk(bx - 1, by - 1, bz) && path(bx - 1, by, bz, bx - 1, by - 1, bz, bx, by - 1, bz)))
        lunion(par, pn, nd(bx, by - 1, bz));
      n++;
    }
    if constexpr (NB == 6) {
      if (lk(bx, by, bz - 1)) {
        const bool viaD = lM && lk(bx - 1, by, bz - 1) && path(bx - 1, by, bz, bx - 1, by, bz - 1, bx, by, bz - 1);
        const bool viaB = lK && lk(bx, by - 1, bz - 1) && path(bx, by - 1, bz, bx, by - 1, bz - 1, bx, by, bz - 1);
        if (!(viaD || viaB)) lunion(par, pn, nd(bx, by, bz - 1));
        n++;
      }
    }
    return n;
  } else {
    // Covering masks.  Interior voxels (and the halo phase, where every pair is
    // two halo voxels of one neighbouring tile) use plain adjacency.  On a tile
    // border a neighbour j may cover i only when the pair (j, i) is itself
    // unioned in this tile's structure: the later voxel of the pair lies in the
    // tile, or both are halo voxels of the same neighbouring tile.  Otherwise
    // the tiles' local forests could lean on each other's cross-tile links.
    const bool interior = HALO || interior_voxel<G, NB>(bx, by, bz);
    auto cover_from = [&](int j) -> uint32_t {
      uint32_t m = NH::adj(j);
      if (interior) return m;
      const Off oj = NH::off(j);
      uint32_t r = 0;
#pragma unroll
      for (int i = 0; i < NH::N; i++) {
        if (!((m >> i) & 1u)) continue;
        const Off oi = NH::off(i);
        if (pair_here<G>(bx + oj.dx, by + oj.dy, bz + oj.dz, bx + oi.dx, by + oi.dy, bz + oi.dz)) r |= 1u << i;
      }
      return r;
    };
    if constexpr (MODE != M_CONT) {
      // 26 / 18 / 8: transitive links
      uint32_t cov = 0;
      int n = 0;
#pragma unroll
      for (int i = 0; i < NH::N; i++) {
        if ((cov >> i) & 1u) continue;
        const Off o = NH::off(i);
        const int qx = bx + o.dx, qy = by + o.dy, qz = bz + o.dz;
        if (!ok(qx, qy, qz)) continue;
        const T q = val(qx, qy, qz);
        const bool l = MODE == M_MULTI ? (q == cur) : fg(q);
        if (!l) continue;
        n++;
        lunion(par, pn, nd(qx, qy, qz));
        cov |= cover_from(i);
      }
      return n;
    } else {
      // continuous: exact-equality covering plus direct matches with neighbour 0
      uint32_t cov = 0;
      int n = 0;
      bool hub = false;
      uint32_t hubcov = 0;
      T hv = T(0);
#pragma unroll
      for (int i = 0; i < NH::N; i++) {
        if ((cov >> i) & 1u) continue;
        const Off o = NH::off(i);
        const int qx = bx + o.dx, qy = by + o.dy, qz = bz + o.dz;
        if (!ok(qx, qy, qz)) continue;
        const T q = val(qx, qy, qz);
        bool l = fg(q) && match(cur, q, delta);
        if (NB == 26 && i == 0 && q == cur) l = true;
        if (!l) continue;
        n++;
        bool covered = false;
        if (i > 0 && hub && ((hubcov >> i) & 1u) && match(hv, q, delta)) covered = true;
        if (!covered) lunion(par, pn, nd(qx, qy, qz));
        const uint32_t cf = cover_from(i);
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
}

template <class T, int MODE, int NB, int QK, class G>
__global__ __launch_bounds__(WG) void k_local(LocalParams p) {
  constexpr int NW = WG / 64;
  constexpr int PT = G::PER_THREAD;
  constexpr int HPT = (G::HALO + WG - 1) / WG;
  constexpr bool PIECES = !(QK == Q_BIN4 || QK == Q_CONT8);

  __shared__ T sv[G::BOXN];
  __shared__ uint32_t par[G::NN];
  __shared__ uint8_t hpiece[G::HALO];
  __shared__ uint32_t wbase[PT * NW + 1];
  __shared__ uint32_t nreq;
  __shared__ unsigned long long red_epl[NW], red_prov[NW];
  __shared__ long long red_first[NW], red_last[NW];

  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int64_t tile = xcd_tile(blockIdx.x, p.ntiles);
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

  // 1. stage tile + halo; voxels outside the volume read as background
  for (int b = t; b < G::BOXN; b += WG) {
    const int bz = b / (G::BX * G::BY);
    const int r = b - bz * (G::BX * G::BY);
    const int by = r / G::BX, bx = r - by * G::BX;
    const int64_t gx = x0 + bx - 1, gy = y0 + by - 1, gz = G::IS3D ? z0 + bz - 1 : 0;
    T v = T(0);
    if (gx >= 0 && gx < sx && gy >= 0 && gy < sy && gz >= 0 && gz < sz) v = in[gx + sx * (gy + sy * gz)];
    sv[b] = v;
  }
  for (int n = t; n < G::NN; n += WG) par[n] = n;
  if (t == 0) nreq = 0;
  __syncthreads();

  // 2. halo pieces: links between halo voxels of the same neighbouring tile
  if constexpr (PIECES) {
    for (int h = t; h < G::HALO; h += WG) {
      int bx, by, bz;
      G::halo_coords(h, bx, by, bz);
      const T cur = sv[G::box(bx, by, bz)];
      if (!fg(cur)) continue;
      link_voxel<T, MODE, NB, QK, G, true>(sv, par, bx, by, bz, cur, delta, gmin, gmax, 0,
                                           tile_slot<G>(bx, by, bz));
    }
    __syncthreads();
    for (int h = t; h < G::HALO; h += WG) hpiece[h] = (par[G::TV + h] == (uint32_t)(G::TV + h));
    __syncthreads();
  }

  // 3. tile voxels (not unrolled: keeps the kernel inside the instruction cache)
  unsigned long long my_epl = 0, my_prov = 0;
  long long my_first = 0x7FFFFFFFFFFFFFFFll, my_last = -1;
#pragma unroll 1
  for (int k = 0; k < PT; k++) {
    const int ps = t + WG * k;
    const int lx = ps & (G::TX - 1), ly = (ps / G::TX) & (G::TY - 1), lz = ps / (G::TX * G::TY);
    const int bx = lx + 1, by = ly + 1, bz = G::IS3D ? lz + 1 : 0;
    const int b = G::box(bx, by, bz);
    const T cur = sv[b];
    if (!fg(cur)) continue;
    const int64_t gy = y0 + ly, gz = z0 + lz;
    if (cur != sv[b - 1]) my_epl++;
    const long long row = gy + sy * gz;
    my_first = row < my_first ? row : my_first;
    my_last = row > my_last ? row : my_last;
    const int nl = link_voxel<T, MODE, NB, QK, G, false>(sv, par, bx, by, bz, cur, delta, gmin,
                                                           gmax, gy, 0);
    if (QK != Q_BLOCK8 && nl == 0) my_prov++;  // block labels are counted separately
  }
  __syncthreads();

  // 4. resolve roots
  uint32_t root[PT];
#pragma unroll
  for (int k = 0; k < PT; k++) {
    const int ps = t + WG * k;
    const int lx = ps & (G::TX - 1), ly = (ps / G::TX) & (G::TY - 1), lz = ps / (G::TX * G::TY);
    const T cur = sv[G::box(lx + 1, ly + 1, G::IS3D ? lz + 1 : 0)];
    root[k] = fg(cur) ? lfind(par, G::template node<QK>(lx, ly, lz)) : 0xFFFFFFFFu;
  }
  uint32_t hroot[HPT];
#pragma unroll
  for (int i = 0; i < HPT; i++) {
    const int h = t + WG * i;
    hroot[i] = 0xFFFFFFFFu;
    // without pieces every linked halo voxel is its own request
    if (h < G::HALO && (!PIECES || hpiece[h])) {
      int bx, by, bz;
      G::halo_coords(h, bx, by, bz);
      if (fg(sv[G::box(bx, by, bz)])) hroot[i] = lfind(par, G::TV + h);
    }
  }
  // 5. rank local roots (ballot order: slice k, then wave, then lane)
  bool isroot[PT];
  uint64_t bal[PT];
#pragma unroll
  for (int k = 0; k < PT; k++) {
    const int ps = t + WG * k;
    const int lx = ps & (G::TX - 1), ly = (ps / G::TX) & (G::TY - 1), lz = ps / (G::TX * G::TY);
    isroot[k] = root[k] == (uint32_t)G::template node<QK>(lx, ly, lz);
    bal[k] = __ballot(isroot[k]);
    if (lane == 0) wbase[k * NW + wv] = __popcll(bal[k]);
  }
  __syncthreads();
  if (t == 0) {
    uint32_t s = 0;
    for (int i = 0; i < PT * NW; i++) {
      const uint32_t c = wbase[i];
      wbase[i] = s;
      s += c;
    }
    wbase[PT * NW] = s;
  }
  __syncthreads();
  const uint32_t R = wbase[PT * NW];
  const uint64_t ltmask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t* RLt = p.RL + tile * G::TV;
#pragma unroll
  for (int k = 0; k < PT; k++) {
    if (isroot[k]) {
      const uint32_t rk = wbase[k * NW + wv] + __popcll(bal[k] & ltmask);
      par[root[k]] = rk;
      const int ps = t + WG * k;
      const int lx = ps & (G::TX - 1), ly = (ps / G::TX) & (G::TY - 1), lz = ps / (G::TX * G::TY);
      const uint32_t key = gkey<QK>(x0 + lx, y0 + ly, z0 + lz, sx, sy, p.osx);
      RLt[rk] = key;
      p.GP[key] = key;
    }
  }
  __syncthreads();

  // 6. per-voxel root index
  const bool narrow = R <= 255;
#pragma unroll
  for (int k = 0; k < PT; k++) {
    const int ps = t + WG * k;
    const uint32_t j = root[k] == 0xFFFFFFFFu ? 0u : par[root[k]] + 1u;
    if (narrow) p.J8[tile * G::TV + ps] = (uint8_t)j;
    else p.J16[tile * G::TV + ps] = (uint16_t)j;
  }

  // 7. cross-tile requests: one per halo piece attached to a local component
  uint32_t* REQt = p.REQ + tile * G::HALO;
#pragma unroll
  for (int i = 0; i < HPT; i++) {
    const int h = t + WG * i;
    const uint32_t r = hroot[i];
    if (r < (uint32_t)G::TV) {
      int lx, ly, lz;
      G::template node_to_local<QK>(r, lx, ly, lz);
      const uint32_t slot = atomicAdd(&nreq, 1u);
      REQt[slot] = ((uint32_t)G::pos(lx, ly, lz) << 16) | (uint32_t)h;
    }
  }

  // 8. counters
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    my_epl += __shfl_xor(my_epl, o);
    my_prov += __shfl_xor(my_prov, o);
    const long long f = __shfl_xor(my_first, o), l = __shfl_xor(my_last, o);
    my_first = f < my_first ? f : my_first;
    my_last = l > my_last ? l : my_last;
  }
  if (lane == 0) {
    red_epl[wv] = my_epl;
    red_prov[wv] = my_prov;
    red_first[wv] = my_first;
    red_last[wv] = my_last;
  }
  __syncthreads();
  if (t == 0) {
    unsigned long long e = 0, pv = 0;
    long long f = 0x7FFFFFFFFFFFFFFFll, l = -1;
    for (int w = 0; w < NW; w++) {
      e += red_epl[w];
      pv += red_prov[w];
      f = red_first[w] < f ? red_first[w] : f;
      l = red_last[w] > l ? red_last[w] : l;
    }
    p.TR[tile] = R;
    p.TQ[tile] = nreq;
    if (e) atomicAdd(&p.ctr->epl, e);
    if (pv) atomicAdd(&p.ctr->prov, pv);
    if (l >= 0) {
      atomicMin(&p.ctr->first_row, f);
      atomicMax(&p.ctr->last_row, l);
    }
  }
}

// Launchers (one translation unit per mode keeps compile times short).
hipError_t launch_local_multi(int dt, int nb, int qk, const LocalParams& p, hipStream_t s);
hipError_t launch_local_binary(int dt, int nb, int qk, const LocalParams& p, hipStream_t s);
hipError_t launch_local_cont(int dt, int nb, int qk, const LocalParams& p, hipStream_t s);

template <int MODE>
hipError_t launch_local_mode(int dt, int nb, int qk, const LocalParams& p, hipStream_t s) {
  const dim3 grid((unsigned)p.ntiles), block(WG);
#define CC3D_L(T, NB, QK, G) \
  hipLaunchKernelGGL((k_local<T, MODE, NB, QK, G>), grid, block, 0, s, p)
#define CC3D_NB(T)                                                               \
  switch (nb) {                                                                  \
    case 26: CC3D_L(T, 26, Q_NONE, Geo3); break;                                 \
    case 18: CC3D_L(T, 18, Q_NONE, Geo3); break;                                 \
    case 6: CC3D_L(T, 6, Q_NONE, Geo3); break;                                   \
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

}  // namespace cc3d_amd
