// Voxel connectivity graph (cc3d/cc3d_graphs.hpp:31-236, bound at
// cc3d/fastcc3d.pyx:1021-1176): per voxel a bitfield with one bit per
// direction, set unless the neighbour in that direction exists and holds a
// different label (so directions out of the volume stay set).  The bit of
// each direction is the published output format (the docstring at
// cc3d/fastcc3d.pyx:1030-1064 and automated_test.py:878-950):
//   3D: faces +x -x +y -y +z -z (bits 0-5); in-plane edges (+1,+1,0)
//       (-1,+1,0) (+1,-1,0) (-1,-1,0) (6-9); edges (+1,0,+1) (-1,0,+1)
//       (0,+1,+1) (0,-1,+1) (+1,0,-1) (-1,0,-1) (0,+1,-1) (0,-1,-1) (10-17);
//       corners (+1,+1,+1) (-1,+1,+1) (+1,-1,+1) (-1,-1,+1) (+1,+1,-1)
//       (-1,+1,-1) (+1,-1,-1) (-1,-1,-1) (18-25).  18-connectivity keeps
//       bits 0-17, 6-connectivity bits 0-5 (uint8).
//   2D (4/8): +x -x +y -y (0-3), (+1,+1) (-1,+1) (+1,-1) (-1,-1) (4-7);
//       4-connectivity keeps bits 0-3.
// A tile and its halo are staged in LDS; the halo re-reads partly hit L2, so
// HBM sees the labels read 1.1 to 1.5 times and the graph written once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/cc3d_amd.h"
#include "common.hpp"

namespace cc3d_amd {

constexpr int8_t kDir3[26][3] = {
  {+1, 0, 0},   {-1, 0, 0},   {0, +1, 0},   {0, -1, 0},   {0, 0, +1},   {0, 0, -1},   {+1, +1, 0},
  {-1, +1, 0},  {+1, -1, 0},  {-1, -1, 0},  {+1, 0, +1},  {-1, 0, +1},  {0, +1, +1},  {0, -1, +1},
  {+1, 0, -1},  {-1, 0, -1},  {0, +1, -1},  {0, -1, -1},  {+1, +1, +1}, {-1, +1, +1}, {+1, -1, +1},
  {-1, -1, +1}, {+1, +1, -1}, {-1, +1, -1}, {+1, -1, -1}, {-1, -1, -1},
};
constexpr int8_t kDir2[8][2] = {
  {+1, 0}, {-1, 0}, {0, +1}, {0, -1}, {+1, +1}, {-1, +1}, {+1, -1}, {-1, -1},
};

// One workgroup per tile of VX x VY x VZ voxels (VZ = 1 in 2D), staged with
// a one-voxel halo in LDS by coalesced loads; thread (lx, ly) then builds the
// bitfields of its column of VZ voxels from LDS.  Whether a neighbour exists
// comes from the voxel's coordinates, not from the staged value.  Tiles are
// spread over the XCDs in contiguous ranges (xcd_tile) when the grid holds one
// workgroup per tile, else walked in a grid-stride loop.
constexpr int VX = 64, VY = 4;

template <bool D2>
struct VTile {
  static constexpr int VZ = D2 ? 1 : 8;
  static constexpr int BX = VX + 2, BY = VY + 2, BZ = D2 ? 1 : VZ + 2;
  static constexpr int BOX = BX * BY * BZ;
  static constexpr int Z0 = D2 ? 0 : 1;  // box z of the tile's first plane
  __host__ __device__ static constexpr int box(int x, int y, int z) { return x + BX * (y + BY * z); }
};

template <class T, class O, bool D2>
__global__ __launch_bounds__(WG) void k_vcg(const T* lab, int64_t sx, int64_t sy, int64_t sz, uint32_t keep,
                                            O* out) {
  using V = VTile<D2>;
  constexpr int ND = D2 ? 8 : 26;
  __shared__ T s[V::BOX];
  const int t = threadIdx.x, lx = t % VX, ly = t / VX;
  const int64_t ntx = (sx + VX - 1) / VX, nty = (sy + VY - 1) / VY, ntz = (sz + V::VZ - 1) / V::VZ;
  const int64_t ntiles = ntx * nty * ntz;
  const bool one = (int64_t)gridDim.x == ntiles;
  for (int64_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
    const int64_t tile = one ? xcd_tile(b, ntiles) : b;
    const int64_t tx = tile % ntx, tr = tile / ntx, ty = tr % nty, tz = tr / nty;
    const int64_t x0 = tx * VX, y0 = ty * VY, z0 = tz * V::VZ;
    __syncthreads();  // the previous tile's reads are done
    for (int h = t; h < V::BOX; h += WG) {
      const int hx = h % V::BX, hy = (h / V::BX) % V::BY, hz = h / (V::BX * V::BY);
      const int64_t gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - V::Z0;
      const bool in = gx >= 0 && gx < sx && gy >= 0 && gy < sy && gz >= 0 && gz < sz;
      s[h] = in ? lab[gx + sx * (gy + sy * gz)] : T(0);
    }
    __syncthreads();
    const int64_t gx = x0 + lx, gy = y0 + ly;
    if (gx >= sx || gy >= sy) continue;
    const bool xm = gx > 0, xp = gx < sx - 1, ym = gy > 0, yp = gy < sy - 1;
#pragma unroll
    for (int lz = 0; lz < V::VZ; lz++) {
      const int64_t gz = z0 + lz;
      if (gz >= sz) break;
      const bool zm = gz > 0, zp = gz < sz - 1;
      const int c0 = V::box(lx + 1, ly + 1, lz + V::Z0);
      const T c = s[c0];
      uint32_t g = D2 ? 0xFFu : 0x3FFFFFFu;
#pragma unroll
      for (int d = 0; d < ND; d++) {
        const int dx = D2 ? kDir2[d][0] : kDir3[d][0], dy = D2 ? kDir2[d][1] : kDir3[d][1];
        const int dz = D2 ? 0 : kDir3[d][2];
        const bool ex = (dx < 0 ? xm : dx > 0 ? xp : true) && (dy < 0 ? ym : dy > 0 ? yp : true) &&
                        (dz < 0 ? zm : dz > 0 ? zp : true);
        if (ex && s[c0 + dx + V::BX * (dy + V::BY * dz)] != c) g &= ~(1u << d);
      }
      out[gx + sx * (gy + sy * gz)] = (O)(g & keep);
    }
  }
}

static thread_local std::string g_err;

static int gfail(int code, const char* msg) {
  g_err = msg;
  return code;
}

template <class T>
static void launch_vcg(const void* labels, int64_t sx, int64_t sy, int64_t sz, int connectivity, void* out,
                       hipStream_t s) {
  const bool d2 = connectivity == 8 || connectivity == 4;
  const int64_t vz = d2 ? VTile<true>::VZ : VTile<false>::VZ;
  const int64_t ntiles = (sx + VX - 1) / VX * ((sy + VY - 1) / VY) * ((sz + vz - 1) / vz);
  const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>(ntiles, 1 << 20)));
  const T* in = static_cast<const T*>(labels);
  switch (connectivity) {
    case 26: hipLaunchKernelGGL((k_vcg<T, uint32_t, false>), grid, dim3(WG), 0, s, in, sx, sy, sz, 0x3FFFFFFu, (uint32_t*)out); break;
    case 18: hipLaunchKernelGGL((k_vcg<T, uint32_t, false>), grid, dim3(WG), 0, s, in, sx, sy, sz, 0x3FFFFu, (uint32_t*)out); break;
    case 6: hipLaunchKernelGGL((k_vcg<T, uint8_t, false>), grid, dim3(WG), 0, s, in, sx, sy, sz, 0x3Fu, (uint8_t*)out); break;
    case 8: hipLaunchKernelGGL((k_vcg<T, uint8_t, true>), grid, dim3(WG), 0, s, in, sx, sy, sz, 0xFFu, (uint8_t*)out); break;
    default: hipLaunchKernelGGL((k_vcg<T, uint8_t, true>), grid, dim3(WG), 0, s, in, sx, sy, sz, 0x0Fu, (uint8_t*)out); break;
  }
}

}  // namespace cc3d_amd

using namespace cc3d_amd;

extern "C" {

const char* cc3d_amd_graphs_last_error(void) { return g_err.c_str(); }

int cc3d_amd_voxel_connectivity_graph(const void* labels, int dtype, int64_t sx, int64_t sy, int64_t sz,
                                      int connectivity, void* out, void* stream) {
  if (!labels || !out || sx <= 0 || sy <= 0 || sz <= 0) return gfail(CC3D_AMD_ERR_INVALID, "bad argument");
  if (connectivity == 8 || connectivity == 4) {
    if (sz != 1) return gfail(CC3D_AMD_ERR_RUNTIME, "sz must be 1 for 2D connectivities.");
  } else if (connectivity != 26 && connectivity != 18 && connectivity != 6) {
    return gfail(CC3D_AMD_ERR_RUNTIME, "Only 4 and 8 2D and 6, 18, and 26 3D connectivities are supported.");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case DT_U8: launch_vcg<uint8_t>(labels, sx, sy, sz, connectivity, out, s); break;
    case DT_U16: launch_vcg<uint16_t>(labels, sx, sy, sz, connectivity, out, s); break;
    case DT_U32: launch_vcg<uint32_t>(labels, sx, sy, sz, connectivity, out, s); break;
    case DT_U64: launch_vcg<uint64_t>(labels, sx, sy, sz, connectivity, out, s); break;
    default: return gfail(CC3D_AMD_ERR_INVALID, "labels must be an unsigned view of an integer type");
  }
  if (hipGetLastError() != hipSuccess) return gfail(CC3D_AMD_ERR_HIP, "voxel_connectivity_graph launch failed");
  return CC3D_AMD_OK;
}

}  // extern "C"
