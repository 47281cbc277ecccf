// Row kernel: local labelling of 3D multilabel and binary volumes
// (connectivity 26, 18, 6) with one voxel per lane.
//
// Replaces the first raster pass of the reference for these paths
// (cc3d/cc3d.hpp:344-1076, cc3d/cc3d_binary.hpp:31-761).  The tile is
// 64 x 8 x 8 voxels; a 512-thread workgroup gives one wave to each z-plane,
// and a lane to each x.  A row (64 voxels along x) is handled with 64-bit
// lane masks:
//
//   F      foreground                       (v != 0)
//   S      run starts: foreground voxels not equal (multilabel) / not linked
//          (binary) to the voxel before them in the row; lane 0 always starts
//          a run, so runs never leave the tile
//   E_d    links of the row to a backward row r' at x + d, d in {-1, 0, +1}
//          (equal values / both foreground), from DPP-shifted copies of r'
//          whose shifted-in lane is the x-halo voxel of r'
//
// Union-find nodes are the runs (identified by their first voxel) and the
// halo nodes: runs of the halo rows and the x-halo voxels.  Node ids are box
// positions (bx, by, bz) = (x + 1, y + 1, z + 1) in raster order, which is
// global key order, so every tree is rooted at its member with the smallest
// key.  A run makes one union with each run of r' it touches: at the first
// voxel of each maximal E_0 interval inside it (U0), and through the
// diagonals only where nothing else joins the two runs -- at its first voxel
// for d = -1 (Um) and at its last voxel for d = +1 (Up).  Unions are appended
// to a per-wave queue in LDS and made 64 at a time, one per lane.
//
// Halo pieces (halo nodes connected through links inside one neighbouring
// tile) are unioned first, and their roots recorded; after the tile's own
// unions each piece whose tree holds tile voxels makes one cross-tile request
// (root index, piece node), resolved by the global kernels.  Outputs (RL, TR,
// J8/J16, GP, REQ/TQ, tcnt) have the same meaning as those of the voxel
// kernel in local.hpp; J is row-major: entry x + 64 * (y + 8 * z).
#pragma once

#include "common.hpp"
#include "local.hpp"  // LocalParams

namespace cc3d_amd {

// Tile geometry of the row kernels.  Box positions: bx in [0, 65]
// (x = -1 .. 64), by in [0, TY + 1], bz in [0, TZ].  Halo index = box index.
template <int TY_, int TZ_>
struct RowGeo {
  static constexpr int TX = 64, TY = TY_, TZ = TZ_;
  static constexpr bool IS3D = true;
  static constexpr int NT = 64 * TZ;  // one wave per plane
  static constexpr int NW = TZ;
  static constexpr int TV = TX * TY * TZ;
  static constexpr int BX = TX + 2, BY = TY + 2, BZ = TZ + 1;
  static constexpr int BOXN = BX * BY * BZ;
  static constexpr int NROWS = BY * BZ;
  // box positions that can be halo nodes: plane bz = 0; rows by = 0 of
  // planes 1..TZ; rows by = TY + 1 of planes 1..TZ-1; the x-halo of the tile
  // rows.  Bounds the requests of a tile.
  static constexpr int HALO = BX * BY + BX * TZ + BX * (TZ - 1) + 2 * TY * TZ;
  static_assert(BOXN < 65536, "16-bit nodes");

  __host__ __device__ static constexpr int box(int bx, int by, int bz) { return bx + BX * (by + BY * bz); }
  __host__ __device__ static constexpr int rid(int by, int bz) { return by + BY * bz; }
  __device__ static inline void halo_coords(int h, int& bx, int& by, int& bz) {
    bz = h / (BX * BY);
    const int r = h - bz * (BX * BY);
    by = r / BX;
    bx = r - by * BX;
  }
  template <int QK>
  __device__ static inline void nid_box(int n, int& bx, int& by, int& bz) {
    halo_coords(n, bx, by, bz);
  }
  __device__ static inline int jidx(int lx, int ly, int lz) { return lx + TX * (ly + TY * lz); }
};
using GeoR = RowGeo<8, 8>;

template <class T>
struct RowLds {
  using G = GeoR;
  static constexpr int QCAP = 128;  // per-wave union queue: 64 pending + one append
  static constexpr int RW = (G::BOXN + 31) / 32;
  T v[G::BOXN];                                 // box values
  alignas(4) uint16_t par[(G::BOXN + 1) & ~1];  // parents (pmin works on aligned words)
  uint32_t rbits[RW];                           // roots of trees that hold tile voxels
  union {
    uint32_t q[G::NW][QCAP];  // union queues (phases 2 and 4)
    uint32_t rprefix[RW];     // exclusive popcount prefix of rbits (phases 6-7)
  } u;
  uint16_t plist[G::HALO];  // piece roots
  uint32_t npiece, nreq, nroots;
  uint32_t red_epl[G::NW], red_prov[G::NW];
  int32_t red_first[G::NW], red_last[G::NW];
};

// DPP lane shifts across the whole wave (wave_shr:1 / wave_shl:1): lane x
// receives src from lane x - 1 (x + 1); the lane with no source keeps `old`.
__device__ inline uint32_t shr1_u32(uint32_t old, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, 0x138, 0xf, 0xf, false);
}
__device__ inline uint32_t shl1_u32(uint32_t old, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, 0x130, 0xf, 0xf, false);
}
template <class T>
__device__ inline T lane_shift(T old, T src, bool right) {
  if constexpr (sizeof(T) <= 4) {
    uint32_t o = 0, s = 0;
    __builtin_memcpy(&o, &old, sizeof(T));
    __builtin_memcpy(&s, &src, sizeof(T));
    const uint32_t r = right ? shr1_u32(o, s) : shl1_u32(o, s);
    T out;
    __builtin_memcpy(&out, &r, sizeof(T));
    return out;
  } else {
    uint64_t o, s;
    __builtin_memcpy(&o, &old, 8);
    __builtin_memcpy(&s, &src, 8);
    const uint32_t lo = right ? shr1_u32((uint32_t)o, (uint32_t)s) : shl1_u32((uint32_t)o, (uint32_t)s);
    const uint32_t hi =
      right ? shr1_u32((uint32_t)(o >> 32), (uint32_t)(s >> 32)) : shl1_u32((uint32_t)(o >> 32), (uint32_t)(s >> 32));
    const uint64_t r = (uint64_t)lo | ((uint64_t)hi << 32);
    T out;
    __builtin_memcpy(&out, &r, 8);
    return out;
  }
}

__device__ inline uint64_t wave_ballot(bool p) { return (uint64_t)__ballot(p); }

// lfind_h / lunion (common.hpp) with every loop bounded by the node count of
// a box: parents only decrease, so a correct forest never needs more hops;
// the bound only guarantees that the kernel ends.
constexpr int RHOPS = 6000;
__device__ inline uint32_t rfind(uint16_t* par, uint32_t a) {
  uint32_t p = pload(par, a);
  for (int i = 0; p != a && i < RHOPS; i++) {
    const uint32_t gp = pload(par, p);
    if (gp == p) return p;
    pstore(par, a, gp);
    a = gp;
    p = pload(par, a);
  }
  return a;
}
// Root of a, reading only (bounded like rfind).
__device__ inline uint32_t rfind_ro(const uint16_t* par, uint32_t a) {
  for (int i = 0; i < RHOPS; i++) {
    const uint32_t p = pload(par, a);
    if (p == a) break;
    a = p;
  }
  return a;
}
__device__ inline void runion(uint16_t* par, uint32_t a, uint32_t b) {
  for (int i = 0; i < RHOPS; i++) {
    a = rfind(par, a);
    b = rfind(par, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = pmin(par, a, b);
    if (old == a) return;
    a = old;
  }
}

// The n queued unions (a << 16 | b) of a wave, one per lane.  Called from a
// few places only, out of line: one copy of the union loops keeps the
// kernel's code small.
__device__ __noinline__ void run_queue(uint16_t* par, const uint32_t* q, uint32_t n, int lane) {
  for (uint32_t i = lane; i < n; i += 64) {
    const uint32_t e = q[i];
    runion(par, e >> 16, e & 0xFFFFu);
  }
}


// Lanes below x (1-based rank of x among the set bits of m, minus one).
__device__ inline uint32_t mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Position of the highest set bit of m at or below lane x (m & lemask != 0).
__device__ inline uint32_t run_start(uint64_t m, uint64_t lemask) {
  const uint64_t b = m & lemask;
  const uint32_t hi = (uint32_t)(b >> 32);
  return hi ? 63u - (uint32_t)__builtin_clz(hi) : 31u - (uint32_t)__builtin_clz((uint32_t)b | 1u);
}

// Buffer loads of box positions (raw buffer loads at 32-bit byte offsets
// from box position (0, 0, 0); an offset at or above the record count reads
// 0), or 64-bit addresses when the box spans 2 GiB or more.
template <class T>
struct BoxLoader {
  __amdgpu_buffer_rsrc_t rs;
  const T* in;
  int64_t x0, y0, z0, sx, sy, sz;
  bool buf;
  uint32_t RB, PB;

  __device__ inline void init(const T* in_, int64_t x0_, int64_t y0_, int64_t z0_, int64_t sx_, int64_t sy_,
                              int64_t sz_, int BXn, int BYn, int BZn) {
    in = in_;
    x0 = x0_;
    y0 = y0_;
    z0 = z0_;
    sx = sx_;
    sy = sy_;
    sz = sz_;
    buf = (uint64_t)(BXn + sx * BYn + sx * sy * BZn) * sizeof(T) < (uint64_t)0x80000000u;
    RB = (uint32_t)(sx * (int64_t)sizeof(T));
    PB = (uint32_t)(sx * sy * (int64_t)sizeof(T));
    if (buf) {
      const int64_t base_el = (x0 - 1) + sx * ((y0 - 1) + sy * (z0 - 1));
      const uint64_t rem = (uint64_t)(sx * sy * sz - base_el) * sizeof(T);
      const int nrec = (int)(rem < 0x80000000ull ? rem : 0x7FFFFFFFull);
      const char* bp = reinterpret_cast<const char*>(in) + base_el * (int64_t)sizeof(T);
      rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(bp), 0, nrec, 0x00020000);
    }
  }
  // value of box position (bx, by, bz); 0 outside the volume
  __device__ inline T load(int bx, int by, int bz) const {
    const int64_t gx = x0 + bx - 1, gy = y0 + by - 1, gz = z0 + bz - 1;
    const bool ok = gx >= 0 && gx < sx && gy >= 0 && gy < sy && gz >= 0 && gz < sz;
    if (buf) {
      const uint32_t off = ok ? (uint32_t)bx * (uint32_t)sizeof(T) + (uint32_t)by * RB + (uint32_t)bz * PB : 0x80000000u;
      if constexpr (sizeof(T) == 1) {
        return __builtin_bit_cast(T, (uint8_t)__builtin_amdgcn_raw_buffer_load_b8(rs, off, 0, 0));
      } else if constexpr (sizeof(T) == 2) {
        return __builtin_bit_cast(T, (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(rs, off, 0, 0));
      } else if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
      } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
      }
    }
    const T v = in[ok ? gx + sx * (gy + sy * gz) : 0];
    return ok ? v : T(0);
  }
};

// Link masks of row `cur` (foreground F) to a backward row r' (values w,
// x-halo values hl / hr, foreground Fw): bit x of E0 / Em / Ep when cur[x]
// links to r'[x] / r'[x - 1] / r'[x + 1].
template <class T, int MODE>
__device__ inline void row_links(T cur, uint64_t F, T w, T hl, T hr, uint64_t Fw, bool diag, uint64_t& E0,
                                 uint64_t& Em, uint64_t& Ep) {
  if constexpr (MODE == M_MULTI) {
    E0 = wave_ballot(cur == w) & F;
    if (diag) {
      Em = wave_ballot(cur == lane_shift<T>(hl, w, true)) & F;
      Ep = wave_ballot(cur == lane_shift<T>(hr, w, false)) & F;
    } else {
      Em = Ep = 0;
    }
  } else {
    E0 = F & Fw;
    if (diag) {
      Em = F & ((Fw << 1) | (fg(hl) ? 1ull : 0ull));
      Ep = F & ((Fw >> 1) | (fg(hr) ? (1ull << 63) : 0ull));
    } else {
      Em = Ep = 0;
    }
  }
}

template <class T, int MODE, int NB>
__device__ inline void rows_tile(const LocalParams& p, const T* in, int64_t sx, int64_t sy, int64_t sz, int64_t ntx,
                                 int64_t nty, int64_t tile, int64_t slot, uint32_t keybase, RowLds<T>& L) {
  using G = GeoR;
  constexpr int BX = G::BX;
  constexpr int QCAP = RowLds<T>::QCAP;
  constexpr int RW = RowLds<T>::RW;
  constexpr bool DIAG = NB != 6;  // x +- 1 neighbours in the rows the connectivity reaches
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t tz = tile / (ntx * nty);
  const int64_t trem = tile - tz * ntx * nty;
  const int64_t ty = trem / ntx, tx = trem - ty * ntx;
  const int64_t x0 = tx * G::TX, y0 = ty * G::TY, z0 = tz * G::TZ;
  uint16_t* par = L.par;
  uint32_t* q = L.u.q[wv];
  const uint64_t lemask = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes <= x

  for (int i = t; i < RW; i += G::NT) L.rbits[i] = 0u;
  if (t == 0) {
    L.npiece = 0;
    L.nreq = 0;
  }

  BoxLoader<T> ld;
  ld.init(in, x0, y0, z0, sx, sy, sz, G::BX, G::BY, G::BZ);

  // 1. staging.  Wave wv loads the rows by = 0..8 of its plane bz = wv + 1
  //    (kept in registers), row by = 9 of that plane (wv < 7) and rows wv
  //    and wv + 8 of the halo plane bz = 0: every box row but (9, 8), which
  //    no voxel of the tile reaches.  Each row: values to LDS, run starts,
  //    parents of run starts, the run start of every lane.
  const int bzo = wv + 1;
  auto stage_row = [&](T v, int by, int bz) {
    const int r = G::rid(by, bz);
    L.v[r * BX + 1 + lane] = v;
    const uint64_t F = wave_ballot(fg(v));
    uint64_t S;
    if constexpr (MODE == M_MULTI) {
      const uint64_t eq = wave_ballot(v == lane_shift<T>(v, v, true));
      S = F & ~(eq & ~1ull);
    } else {
      S = F & ~(F << 1);
    }
    if ((S >> lane) & 1ull) par[r * BX + 1 + lane] = (uint16_t)(r * BX + 1 + lane);
  };
  {
    T own[G::TY + 1];  // rows by = 0..8 of plane bzo
    T extra[3];
    const int eby0 = wv, eby1 = wv + 8;
#pragma unroll
    for (int k = 0; k <= G::TY; k++) own[k] = ld.load(1 + lane, k, bzo);
    extra[0] = wv < G::TZ - 1 ? ld.load(1 + lane, G::TY + 1, bzo) : T(0);
    extra[1] = ld.load(1 + lane, eby0, 0);
    extra[2] = eby1 <= G::TY + 1 ? ld.load(1 + lane, eby1, 0) : T(0);
    // x-halo of these rows: lane 2k + side, k over the rows above in order
    const int nrow = (G::TY + 1) + (wv < G::TZ - 1 ? 1 : 0) + 1 + (eby1 <= G::TY + 1 ? 1 : 0);
    int hrow = lane >> 1;
    int hby, hbz;
    if (hrow <= G::TY) {
      hby = hrow;
      hbz = bzo;
    } else if (hrow == G::TY + 1 && wv < G::TZ - 1) {
      hby = G::TY + 1;
      hbz = bzo;
    } else {
      const int e = hrow - (G::TY + 1) - (wv < G::TZ - 1 ? 1 : 0);
      hby = e == 0 ? eby0 : eby1;
      hbz = 0;
    }
    const bool hlive = hrow < nrow;
    const int hbx = (lane & 1) ? BX - 1 : 0;
    const T hv = hlive ? ld.load(hbx, hby, hbz) : T(0);
#pragma unroll
    for (int k = 0; k <= G::TY; k++) stage_row(own[k], k, bzo);
    if (wv < G::TZ - 1) stage_row(extra[0], G::TY + 1, bzo);
    stage_row(extra[1], eby0, 0);
    if (eby1 <= G::TY + 1) stage_row(extra[2], eby1, 0);
    if (hlive) {
      const int n = G::box(hbx, hby, hbz);
      L.v[n] = hv;
      if (fg(hv)) par[n] = (uint16_t)n;
    }
  }
  __syncthreads();

  // union queue of this wave
  uint32_t qn = 0;
  auto flush = [&]() {
    run_queue(par, q, qn, lane);
    qn = 0;
  };
  auto append = [&](uint64_t U, uint32_t a, uint32_t b) {
    if (U == 0) return;
    if ((U >> lane) & 1ull) q[qn + mbcnt64(U)] = (a << 16) | b;
    qn += (uint32_t)__builtin_popcountll(U);
    if (qn >= 64) flush();
  };
  const uint64_t ltmask = lemask >> 1;            // lanes below x
  const uint64_t le1mask = (lemask << 1) | 1ull;  // lanes up to x + 1
  auto row_vals = [&](int r) -> T { return L.v[r * BX + 1 + lane]; };
  auto row_F = [&](T v) -> uint64_t { return wave_ballot(fg(v)); };
  auto row_S = [&](T v, uint64_t F) -> uint64_t {
    if constexpr (MODE == M_MULTI) {
      const uint64_t eq = wave_ballot(v == lane_shift<T>(v, v, true));
      return F & ~(eq & ~1ull);
    } else {
      return F & ~(F << 1);
    }
  };
  // Unions of row r (values cur, F, S) with backward row rp.  Links in
  // sk0 / skm / skp (offsets 0, -1, +1) are implied by others and make no
  // union; E0 / Em / Ep return the row's links to rp.
  // a: node of the run of each lane of row r.  Partner nodes come from the
  // run starts of rp, recomputed from its values.
  auto contact = [&](uint32_t a, T cur, uint64_t F, uint64_t S, uint64_t Send, int rp, bool diag, bool halo_only,
                     uint64_t& any, uint64_t sk0, uint64_t skm, uint64_t skp, uint64_t& E0, uint64_t& Em,
                     uint64_t& Ep) {
    const T w = row_vals(rp);
    const T hl = L.v[rp * BX], hr = L.v[rp * BX + BX - 1];
    const uint64_t Fw = row_F(w);
    row_links<T, MODE>(cur, F, w, hl, hr, Fw, diag, E0, Em, Ep);
    if (halo_only) {  // the x-halo voxels of r' lie in other tiles
      Em &= ~1ull;
      Ep &= ~(1ull << 63);
    }
    any |= E0 | Em | Ep;
    const uint64_t E0u = E0 & ~sk0;
    const uint64_t U0 = E0u & ~((E0u << 1) & ~S);
    const uint64_t Um = Em & ~skm & ~E0 & S;
    const uint64_t Up = Ep & ~skp & ~E0 & Send;
    if ((U0 | Um | Up) == 0) return;
    const uint64_t Sw = row_S(w, Fw);
    const uint32_t base = (uint32_t)(rp * BX + 1);
    if (U0) append(U0, a, base + run_start(Sw, lemask));
    if (Um) append(Um, a, lane == 0 ? base - 1u : base + run_start(Sw, ltmask));
    if (Up) append(Up, a, lane == 63 ? base + 64u : base + run_start(Sw, le1mask));
  };

  // 2. halo pieces: links between halo nodes of one neighbouring tile.
  //    Row pairs: plane 0 rows 2..8 with the row before (the tile below);
  //    rows by = 0 of planes 2..8 and rows by = 9 of planes 2..7 with the
  //    same row of the plane before (the tiles beside in y).
  for (int pi = wv; pi < 20; pi += G::NW) {
    int r, rp;
    if (pi < 7) {
      r = G::rid(pi + 2, 0);
      rp = G::rid(pi + 1, 0);
    } else if (pi < 14) {
      r = G::rid(0, pi - 5);
      rp = G::rid(0, pi - 6);
    } else {
      r = G::rid(G::TY + 1, pi - 12);
      rp = G::rid(G::TY + 1, pi - 13);
    }
    const T cur = row_vals(r);
    const uint64_t F = row_F(cur);
    const uint64_t S = row_S(cur, F);
    const uint64_t Send = F & ~((F & ~S) >> 1);
    uint64_t any = 0, e0, em, ep;
    const uint32_t a = (uint32_t)(r * BX + 1) + run_start(S, lemask);
    contact(a, cur, F, S, Send, rp, DIAG, true, any, 0, 0, 0, e0, em, ep);
  }
  flush();
  // x-halo columns of the tile rows (the tiles beside in x): each voxel with
  // the voxels of its column before it, as the connectivity reaches them
  if (t < 128) {
    const int bxh = t < 64 ? 0 : BX - 1;
    const int k = t & 63, by = 1 + (k & 7), bz = 1 + (k >> 3);
    const uint32_t n = (uint32_t)G::box(bxh, by, bz);
    const T cur = L.v[n];
    if (fg(cur)) {
      auto link = [&](int by2, int bz2) {
        const uint32_t m = (uint32_t)G::box(bxh, by2, bz2);
        const T o = L.v[m];
        const bool l = MODE == M_MULTI ? (o == cur) : fg(o);
        if (l) runion(par, n, m);
      };
      if (by >= 2) link(by - 1, bz);
      if (bz >= 2) {
        link(by, bz - 1);
        if (NB != 6) {
          if (by >= 2) link(by - 1, bz - 1);
          if (by <= G::TY - 1) link(by + 1, bz - 1);
        }
      }
    }
  }
  __syncthreads();

  // 3. piece roots: halo nodes that are roots now
  {
    auto push = [&](bool is_root, uint32_t n) {
      const uint64_t m = wave_ballot(is_root);
      if (m == 0) return;
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(&L.npiece, (uint32_t)__builtin_popcountll(m));
      base = __shfl(base, 0);
      if (is_root) L.plist[base + mbcnt64(m)] = (uint16_t)n;
    };
    // halo rows: plane 0 (10 rows), by = 0 of planes 1..8, by = 9 of planes 1..7
    for (int hr = wv; hr < 25; hr += G::NW) {
      const int r = hr < 10 ? G::rid(hr, 0) : hr < 18 ? G::rid(0, hr - 9) : G::rid(G::TY + 1, hr - 17);
      const T v = row_vals(r);
      const uint64_t F = row_F(v);
      const uint64_t S = row_S(v, F);
      const uint32_t n = (uint32_t)(r * BX + 1 + lane);
      push(((S >> lane) & 1ull) && pload(par, n) == n, n);
    }
    // x-halo voxels of the 89 staged rows (rows 0..88), a wave at a time
    for (int i0 = wv * 64; i0 < 2 * 89; i0 += G::NT) {
      const int i = i0 + lane;
      const bool live = i < 2 * 89;
      const int r = live ? i >> 1 : 0;
      const uint32_t n = (uint32_t)(r * BX + ((i & 1) ? BX - 1 : 0));
      push(live && fg(L.v[n]) && pload(par, n) == n, n);
    }
  }
  __syncthreads();

  // 4. the tile's rows: unions with the backward rows; counters
  uint32_t my_epl = 0, my_prov = 0;
  int my_first = 0x7FFFFFFF, my_last = -1;
#pragma unroll 1
  for (int by = 1; by <= G::TY; by++) {
    const int r = G::rid(by, bzo);
    const T cur = row_vals(r);
    const uint64_t F = row_F(cur);
    const T hlc = L.v[r * BX];
    const uint64_t eqv = wave_ballot(cur == lane_shift<T>(hlc, cur, true));  // bit 0: vs the x-halo voxel
    uint64_t S, link0;
    if constexpr (MODE == M_MULTI) {
      S = F & ~(eqv & ~1ull);
      link0 = eqv & F;
    } else {
      S = F & ~(F << 1);
      link0 = F & ((F << 1) | (fg(hlc) ? 1ull : 0ull));
    }
    const uint64_t Send = F & ~((F & ~S) >> 1);
    my_epl += (uint32_t)__builtin_popcountll(F & ~eqv);
    if (F) {
      const int lr = (by - 1) + G::TY * (bzo - 1);
      my_first = lr < my_first ? lr : my_first;
      my_last = lr > my_last ? lr : my_last;
    }
    uint64_t any = link0;
    if (link0 & 1ull) append(1ull, (uint32_t)(r * BX + 1), (uint32_t)(r * BX));  // the x-halo voxel before the row
    // Same plane, row before (B); plane before: rows by (C), by - 1 (D) and
    // by + 1 (E).  A link p - q of voxel p = (x, y, z) to q = (x + d, y - 1,
    // z - 1) in D is implied when p links to (x + d, y - 1, z) in B or to
    // (x + d, y, z - 1) in C: that voxel is adjacent to both, and its links
    // to q (a C link of row B, a B link of row C) are always unioned by the
    // tile that owns the later voxel.  Likewise for E through C.
    uint64_t b0, bm, bp, c0, cm, cp, t0, tm, tp;
    const uint32_t a = (uint32_t)(r * BX + 1) + run_start(S, lemask);
    contact(a, cur, F, S, Send, G::rid(by - 1, bzo), DIAG, false, any, 0, 0, 0, b0, bm, bp);
    contact(a, cur, F, S, Send, G::rid(by, bzo - 1), DIAG, false, any, 0, 0, 0, c0, cm, cp);
    if (NB != 6) {
      contact(a, cur, F, S, Send, G::rid(by - 1, bzo - 1), NB == 26, false, any, b0 | c0, bm | cm, bp | cp, t0, tm,
              tp);
      contact(a, cur, F, S, Send, G::rid(by + 1, bzo - 1), NB == 26, false, any, c0, cm, cp, t0, tm, tp);
    }
    my_prov += (uint32_t)__builtin_popcountll(F & ~any);
  }
  flush();
  __syncthreads();

  // 5. roots of the tile's run starts, 64 at a time from a per-wave queue;
  //    mark the roots and store each run start's root as its parent.  Every
  //    walk here and in step 7 only reads, so the only stores are these: a
  //    run start's parent is its root from here on.
  {
    uint32_t fnum = 0;
    auto fflush = [&]() {
      for (uint32_t i = lane; i < fnum; i += 64) {
        const uint32_t n = q[i];
        const uint32_t root = rfind_ro(par, n);
        if (root != n) pstore(par, n, root);
        if (!((lload(&L.rbits[root >> 5]) >> (root & 31)) & 1u)) atomicOr(&L.rbits[root >> 5], 1u << (root & 31));
      }
      fnum = 0;
    };
#pragma unroll 1
    for (int by = 1; by <= G::TY; by++) {
      const int r = G::rid(by, bzo);
      const T cur = row_vals(r);
      const uint64_t S = row_S(cur, row_F(cur));
      if (S == 0) continue;
      if ((S >> lane) & 1ull) q[fnum + mbcnt64(S)] = (uint32_t)(r * BX + 1 + lane);
      fnum += (uint32_t)__builtin_popcountll(S);
      if (fnum >= 64) fflush();
    }
    fflush();
  }
  __syncthreads();

  // 6. rank the marked roots in node (= key) order
  if (wv == 0) {
    uint32_t c[3];
#pragma unroll
    for (int k = 0; k < 3; k++) c[k] = 3 * lane + k < RW ? (uint32_t)__popc(L.rbits[3 * lane + k]) : 0u;
    const uint32_t tot = c[0] + c[1] + c[2];
    uint32_t incl = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    uint32_t ex = incl - tot;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if (3 * lane + k < RW) L.u.rprefix[3 * lane + k] = ex;
      ex += c[k];
    }
    if (lane == 63) L.nroots = incl;
  }
  __syncthreads();
  static_assert(3 * 64 >= RW, "one wave scans the root bitmask");
  const uint32_t R = L.nroots;
  auto rank = [&](uint32_t r) -> uint32_t {
    return L.u.rprefix[r >> 5] + (uint32_t)__popc(L.rbits[r >> 5] & ((1u << (r & 31)) - 1u));
  };

  // 7. outputs.  Keys in 32-bit arithmetic (below 2^32; intermediate wraps
  //    cancel): box position (bx, by, bz) is voxel (x0 + bx - 1, ...).
  const uint32_t kbox0 =
    (uint32_t)(x0 - 1) + (uint32_t)sx * ((uint32_t)(y0 - 1) + (uint32_t)sy * (uint32_t)(z0 - 1)) + keybase;
  const uint32_t root0 = tile_root0<G>(sx, sy, sz, tx, ty, tz, keybase);
  uint32_t* RLt = p.RL + root0;
  if (t < RW) {
    uint32_t bits = L.rbits[t];
    uint32_t rk = L.u.rprefix[t];
    while (bits) {
      const int b = __builtin_ctz(bits);
      bits &= bits - 1u;
      int bx, by, bz;
      G::halo_coords(32 * t + b, bx, by, bz);
      const uint32_t key = kbox0 + (uint32_t)bx + (uint32_t)sx * ((uint32_t)by + (uint32_t)sy * (uint32_t)bz);
      p.GP[root0 + rk] = root0 + rk;
      RLt[rk++] = key;
      atomicOr(p.bits + (key >> 5), 1u << (key & 31));
    }
  }
#pragma unroll 2
  for (int by = 1; by <= G::TY; by++) {
    const int r = G::rid(by, bzo);
    const T cur = row_vals(r);
    const uint64_t F = row_F(cur);
    const uint64_t S = row_S(cur, F);
    uint32_t j = 0;
    if (fg(cur)) j = rank(pload(par, (uint32_t)(r * BX + 1) + run_start(S, lemask))) + 1u;  // its run's root
    const int64_t e = slot * G::TV + G::jidx(lane, by - 1, bzo - 1);
    if (R <= 255) p.J8[e] = (uint8_t)j;
    else p.J16[e] = (uint16_t)j;
  }
  for (uint32_t i = t; i < L.npiece; i += G::NT) {
    const uint32_t h = L.plist[i];
    const uint32_t root = rfind_ro(par, h);
    if ((L.rbits[root >> 5] >> (root & 31)) & 1u) {
      const uint32_t k = atomicAdd(&L.nreq, 1u);
      p.REQ[slot * G::HALO + k] = (rank(root) << 16) | h;
    }
  }
  // counters
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int f = __shfl_xor(my_first, o), l = __shfl_xor(my_last, o);
    my_first = f < my_first ? f : my_first;
    my_last = l > my_last ? l : my_last;
  }
  if (lane == 0) {
    L.red_epl[wv] = my_epl;
    L.red_prov[wv] = my_prov;
    L.red_first[wv] = my_first;
    L.red_last[wv] = my_last;
  }
  __syncthreads();
  if (t == 0) {
    uint32_t e = 0, pv = 0;
    int fl = 0x7FFFFFFF, ll = -1;
    for (int w = 0; w < G::NW; w++) {
      e += L.red_epl[w];
      pv += L.red_prov[w];
      fl = L.red_first[w] < fl ? L.red_first[w] : fl;
      ll = L.red_last[w] > ll ? L.red_last[w] : ll;
    }
    long long f = 0x7FFFFFFFFFFFFFFFll, l = -1;
    if (ll >= 0) {
      f = (y0 + fl % G::TY) + sy * (z0 + fl / G::TY);
      l = (y0 + ll % G::TY) + sy * (z0 + ll / G::TY);
    }
    p.TR[slot] = R;
    p.TQ[slot] = L.nreq;
    TileCnt c;
    c.epl = e;
    c.prov = pv;
    c.first_row = f;
    c.last_row = l;
    p.tcnt[slot] = c;
  }
}

// Waves per SIMD the row kernel is compiled for: what the LDS of one tile
// admits (3 workgroups of 8 waves per CU up to 4-byte voxels, 2 for 8-byte).
template <class T>
__host__ __device__ constexpr int rows_waves() {
  return sizeof(T) <= 2 ? 8 : sizeof(T) == 4 ? 6 : 4;
}

template <class T, int MODE, int NB>
__global__ __launch_bounds__(GeoR::NT, (rows_waves<T>())) void k_rows_local(LocalParams p) {
  __shared__ RowLds<T> L;
  const int64_t tile = xcd_tile(blockIdx.x, p.ntiles);
  rows_tile<T, MODE, NB>(p, static_cast<const T*>(p.in), p.sx, p.sy, p.sz, p.ntx, p.nty, tile, tile, 0u, L);
}

// Batch: tiles [gtile0, gtile0 + gtiles) of volumes of connectivity NB; each
// workgroup reads its volume's table entry.
template <class T, int MODE, int NB>
__global__ __launch_bounds__(GeoR::NT, (rows_waves<T>())) void k_rows_local_batch(LocalParams p, const BatchVol* vols,
                                                                                 const uint32_t* vix, int64_t gtile0,
                                                                                 int64_t gtiles) {
  __shared__ RowLds<T> L;
  const int64_t slot = gtile0 + xcd_tile(blockIdx.x, gtiles);
  const BatchVol& v = vols[vix[slot]];
  rows_tile<T, MODE, NB>(p, static_cast<const T*>(v.in), v.sx, v.sy, v.sz, v.ntx, v.nty, slot - v.tile0, slot, v.key0,
                         L);
}

hipError_t launch_rows_multi(int dt, int nb, const LocalParams& p, hipStream_t s);
hipError_t launch_rows_binary(int dt, int nb, const LocalParams& p, hipStream_t s);
hipError_t launch_rows_batch_multi(int dt, int nb, const LocalParams& p, const BatchVol* vols, const uint32_t* vix,
                                   int64_t gtile0, int64_t gtiles, hipStream_t s);
hipError_t launch_rows_batch_binary(int dt, int nb, const LocalParams& p, const BatchVol* vols, const uint32_t* vix,
                                    int64_t gtile0, int64_t gtiles, hipStream_t s);

template <int MODE>
hipError_t launch_rows_mode(int dt, int nb, const LocalParams& p, hipStream_t s) {
  const dim3 grid((unsigned)p.ntiles), block(GeoR::NT);
#define CC3D_R(T, NB) hipLaunchKernelGGL((k_rows_local<T, MODE, NB>), grid, block, 0, s, p)
#define CC3D_RNB(T)                       \
  switch (nb) {                           \
    case 26: CC3D_R(T, 26); break;        \
    case 18: CC3D_R(T, 18); break;        \
    case 6: CC3D_R(T, 6); break;          \
    default: return hipErrorInvalidValue; \
  }
  switch (dt) {
    case DT_U8: CC3D_RNB(uint8_t); break;
    case DT_U16: CC3D_RNB(uint16_t); break;
    case DT_U32: CC3D_RNB(uint32_t); break;
    case DT_U64: CC3D_RNB(uint64_t); break;
    case DT_F32: CC3D_RNB(float); break;
    case DT_F64: CC3D_RNB(double); break;
    default: return hipErrorInvalidValue;
  }
#undef CC3D_RNB
#undef CC3D_R
  return hipGetLastError();
}

template <int MODE>
hipError_t launch_rows_batch_mode(int dt, int nb, const LocalParams& p, const BatchVol* vols, const uint32_t* vix,
                                  int64_t gtile0, int64_t gtiles, hipStream_t s) {
  const dim3 grid((unsigned)gtiles), block(GeoR::NT);
#define CC3D_R(T, NB) \
  hipLaunchKernelGGL((k_rows_local_batch<T, MODE, NB>), grid, block, 0, s, p, vols, vix, gtile0, gtiles)
#define CC3D_RNB(T)                       \
  switch (nb) {                           \
    case 26: CC3D_R(T, 26); break;        \
    case 18: CC3D_R(T, 18); break;        \
    case 6: CC3D_R(T, 6); break;          \
    default: return hipErrorInvalidValue; \
  }
  switch (dt) {
    case DT_U8: CC3D_RNB(uint8_t); break;
    case DT_U16: CC3D_RNB(uint16_t); break;
    case DT_U32: CC3D_RNB(uint32_t); break;
    case DT_U64: CC3D_RNB(uint64_t); break;
    case DT_F32: CC3D_RNB(float); break;
    case DT_F64: CC3D_RNB(double); break;
    default: return hipErrorInvalidValue;
  }
#undef CC3D_RNB
#undef CC3D_R
  return hipGetLastError();
}

}  // namespace cc3d_amd
