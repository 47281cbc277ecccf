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
// One thread per voxel; the neighbours of a row come through L1/L2, so HBM
// sees one read of the labels and one write of the graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/cc3d_amd.h"
#include "common.hpp"

namespace cc3d_amd {

struct Dir {
  int8_t dx, dy, dz;
};

__constant__ Dir g_dir3[26] = {
  {+1, 0, 0},   {-1, 0, 0},   {0, +1, 0},   {0, -1, 0},   {0, 0, +1},   {0, 0, -1},   {+1, +1, 0},
  {-1, +1, 0},  {+1, -1, 0},  {-1, -1, 0},  {+1, 0, +1},  {-1, 0, +1},  {0, +1, +1},  {0, -1, +1},
  {+1, 0, -1},  {-1, 0, -1},  {0, +1, -1},  {0, -1, -1},  {+1, +1, +1}, {-1, +1, +1}, {+1, -1, +1},
  {-1, -1, +1}, {+1, +1, -1}, {-1, +1, -1}, {+1, -1, -1}, {-1, -1, -1},
};
__constant__ Dir g_dir2[8] = {
  {+1, 0, 0}, {-1, 0, 0}, {0, +1, 0}, {0, -1, 0}, {+1, +1, 0}, {-1, +1, 0}, {+1, -1, 0}, {-1, -1, 0},
};

// One workgroup covers WG consecutive x of one row (y, z); rows are
// enumerated in a grid-stride loop over blocks.
template <class T, class O, bool D2>
__global__ __launch_bounds__(WG) void k_vcg(const T* lab, int64_t sx, int64_t sy, int64_t sz, uint32_t keep,
                                            O* out) {
  constexpr int ND = D2 ? 8 : 26;
  const int64_t nxb = (sx + WG - 1) / WG, nblocks = nxb * sy * sz;
  for (int64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
    const int64_t row = blk / nxb, x = (blk - row * nxb) * WG + threadIdx.x;
    if (x >= sx) continue;
    const int64_t y = row % sy, z = row / sy;
    const int64_t i = x + sx * row;
    const T c = lab[i];
    uint32_t g = D2 ? 0xFFu : 0x3FFFFFFu;
#pragma unroll
    for (int d = 0; d < ND; d++) {
      if (!((keep >> d) & 1u)) continue;
      const Dir o = D2 ? g_dir2[d] : g_dir3[d];
      const int64_t nx = x + o.dx, ny = y + o.dy, nz = z + o.dz;
      if (nx < 0 || nx >= sx || ny < 0 || ny >= sy || nz < 0 || nz >= sz) continue;
      if (lab[i + o.dx + sx * (o.dy + sy * (int64_t)o.dz)] != c) g &= ~(1u << d);
    }
    out[i] = (O)(g & keep);
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
  const int64_t nblocks = (sx + WG - 1) / WG * sy * sz;
  const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>(nblocks, 1 << 20)));
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
