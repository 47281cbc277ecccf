// color_connectivity_graph on the GPU (cc3d/cc3d_graphs.hpp:525-1106, bound
// at cc3d/fastcc3d.pyx:941-1018): connected components of a voxel
// connectivity graph.
//
// Every voxel is a node.  Voxel p is joined to a backward neighbour q when
// p's own bit for the direction p -> q is set and q lies in the volume (the
// reference takes the graph as undirected and reads only p's bits).  Backward
// neighbours: 3D 26 (uint32 graph) the four of plane z above and left and the
// nine of plane z - 1; 3D 6 -x, -y, -z; 2D 8 -x, (-1,-1), -y, (+1,-1); 2D 4
// -x, -y.  The bit of a direction is the voxel_connectivity_graph layout:
// the 3D one for uint32 graphs, in 2D uint8 graphs the diagonals are bits
// 4-7.  Components are numbered by their first voxel in raster order, the
// order of the reference's provisional labels.
//
// Kernels: one wave per 64-voxel row chunk.  k_ccg_init points every voxel at
// the start of its run along x (voxels joined by their -x bit).  k_ccg_link
// unions runs across the other directions on a global forest over voxel
// indices, one union per maximal interval of one direction's links that stays
// inside one run of each row.  Then roots set a bit in a bitmask, an
// exclusive popcount scan numbers them, and every voxel takes its root's
// number.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/cc3d_amd.h"
#include "common.hpp"

namespace cc3d_amd {

// Exclusive popcount prefix of a root bitmask (cc3d_amd.hip): prefix per
// word, total into ctr->N.
hipError_t scan_root_bits(const uint32_t* bits, int64_t nwords, uint32_t* bsum, uint32_t* prefix, Counters* ctr,
                          hipStream_t s);

struct CcgDir {
  int8_t dx, dy, dz, bit;
};
// Backward directions other than -x (which forms the runs), per variant:
// 0 3D 26, 1 3D 6, 2 2D 8 uint8, 3 2D 8 uint32, 4 2D 4.
__constant__ CcgDir kCcg[5][12] = {
  {{-1, -1, 0, 9}, {0, -1, 0, 3}, {+1, -1, 0, 8}, {-1, -1, -1, 25}, {0, -1, -1, 17}, {+1, -1, -1, 24},
   {-1, 0, -1, 15}, {0, 0, -1, 5}, {+1, 0, -1, 14}, {-1, +1, -1, 23}, {0, +1, -1, 16}, {+1, +1, -1, 22}},
  {{0, -1, 0, 3}, {0, 0, -1, 5}},
  {{-1, -1, 0, 7}, {0, -1, 0, 3}, {+1, -1, 0, 6}},
  {{-1, -1, 0, 9}, {0, -1, 0, 3}, {+1, -1, 0, 8}},
  {{0, -1, 0, 3}},
};
__device__ inline int ccg_ndirs(int var) { return var == 0 ? 12 : var == 1 ? 2 : var == 4 ? 1 : 3; }
constexpr uint32_t CCG_MX = 1u << 1;  // the -x bit in every layout

__device__ inline uint64_t ccg_ballot(bool p) { return (uint64_t)__ballot(p); }

// Run starts of a row chunk: voxels whose -x bit does not join them to the
// voxel before them inside the chunk (lane 0 always starts a run).
template <class VT>
__device__ inline uint64_t ccg_starts(VT v, bool valid, int lane) {
  const uint64_t cont = ccg_ballot(valid && lane > 0 && (v & CCG_MX));
  return ccg_ballot(valid) & ~cont;
}

template <class VT>
__global__ __launch_bounds__(WG) void k_ccg_init(const VT* vcg, int64_t sx, int64_t rows, uint32_t* GP) {
  const int lane = threadIdx.x & 63;
  const int64_t nch = (sx + 63) / 64;
  const int64_t wave = ((int64_t)blockIdx.x * WG + threadIdx.x) >> 6;
  if (wave >= rows * nch) return;
  const int64_t row = wave / nch, ch = wave - row * nch;
  const int64_t gx = ch * 64 + lane;
  const bool valid = gx < sx;
  const int64_t v = row * sx + gx;
  const VT c = valid ? vcg[v] : VT(0);
  const uint64_t S = ccg_starts<VT>(c, valid, lane);
  const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const uint64_t b = S & le;
  const uint32_t hi = (uint32_t)(b >> 32);
  const int rs = hi ? 63 - __builtin_clz(hi) : 31 - __builtin_clz((uint32_t)b | 1u);
  if (valid) GP[v] = (uint32_t)(row * sx + ch * 64 + rs);
}

template <class VT>
__global__ __launch_bounds__(WG) void k_ccg_link(const VT* vcg, int64_t sx, int64_t sy, int64_t sz, int var,
                                                 uint32_t* GP) {
  const int lane = threadIdx.x & 63;
  const int64_t nch = (sx + 63) / 64;
  const int64_t wave = ((int64_t)blockIdx.x * WG + threadIdx.x) >> 6;
  if (wave >= sy * sz * nch) return;
  const int64_t row = wave / nch, ch = wave - row * nch;
  const int64_t y = row % sy, z = row / sy;
  const int64_t gx = ch * 64 + lane;
  const bool valid = gx < sx;
  const int64_t v = row * sx + gx;
  const VT c = valid ? vcg[v] : VT(0);
  const uint64_t S = ccg_starts<VT>(c, valid, lane);
  // a run continuing from the chunk before
  if (lane == 0 && valid && gx > 0 && (c & CCG_MX)) gunion(GP, (uint32_t)v, (uint32_t)(v - 1));
  const int nd = ccg_ndirs(var);
  for (int k = 0; k < nd; k++) {
    const CcgDir d = kCcg[var][k];
    const int64_t qy = y + d.dy, qz = z + d.dz;
    if (qy < 0 || qy >= sy || qz < 0 || qz >= sz) continue;  // the whole row is out of the volume
    const int64_t qx = gx + d.dx;
    const bool link = valid && qx >= 0 && qx < sx && ((c >> d.bit) & 1);
    const uint64_t E = ccg_ballot(link);
    if (E == 0) continue;
    const int64_t qrow = qy + sy * qz;
    const VT w = valid ? vcg[qrow * sx + gx] : VT(0);
    const uint64_t Sq = ccg_starts<VT>(w, valid, lane);
    // run starts of the partner row at x + dx; unknown (other chunk) counts as a start
    const uint64_t Sd = d.dx == 0 ? Sq : d.dx < 0 ? ((Sq << 1) | 1ull) : ((Sq >> 1) | (1ull << 63));
    const uint64_t U = E & ~((E << 1) & ~S & ~Sd);
    if ((U >> lane) & 1ull) gunion(GP, (uint32_t)v, (uint32_t)(qrow * sx + qx));
  }
}

__global__ __launch_bounds__(WG) void k_ccg_roots(uint32_t* GP, int64_t V, uint32_t* bits) {
  for (int64_t v = (int64_t)blockIdx.x * WG + threadIdx.x; v < V; v += (int64_t)gridDim.x * WG) {
    const uint32_t r = gfind_settled(GP, (uint32_t)v);
    atomicMin(GP + v, r);  // the set's minimum: from here on every voxel's parent is its root
    if (r == (uint32_t)v) atomicOr(bits + (r >> 5), 1u << (r & 31));
  }
}

__global__ __launch_bounds__(WG) void k_ccg_final(const uint32_t* GP, int64_t V, const uint32_t* bits,
                                                  const uint32_t* prefix, uint32_t* out) {
  for (int64_t v = (int64_t)blockIdx.x * WG + threadIdx.x; v < V; v += (int64_t)gridDim.x * WG) {
    const uint32_t r = GP[v];
    out[v] = prefix[r >> 5] + (uint32_t)__popc(bits[r >> 5] & ((1u << (r & 31)) - 1u)) + 1u;
  }
}

__global__ void k_ccg_clear(uint32_t* w, int64_t n, Counters* c) {
  for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (int64_t)gridDim.x * WG) w[i] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) c->N = 0;
}

static thread_local std::string c_err;

static int cfail(int code, const std::string& msg) {
  c_err = msg;
  return code;
}

template <class VT>
static hipError_t ccg_launch(const VT* vcg, int64_t sx, int64_t sy, int64_t sz, int var, uint32_t* GP, hipStream_t s) {
  const int64_t waves = ((sx + 63) / 64) * sy * sz;
  const unsigned blocks = (unsigned)((waves + WG / 64 - 1) / (WG / 64));
  hipLaunchKernelGGL((k_ccg_init<VT>), dim3(blocks), dim3(WG), 0, s, vcg, sx, sy * sz, GP);
  hipLaunchKernelGGL((k_ccg_link<VT>), dim3(blocks), dim3(WG), 0, s, vcg, sx, sy, sz, var, GP);
  return hipGetLastError();
}

}  // namespace cc3d_amd

using namespace cc3d_amd;

extern "C" {

const char* cc3d_amd_color_last_error(void) { return c_err.c_str(); }

int cc3d_amd_color_connectivity_graph(const void* vcg, int vcg_size, int64_t sx, int64_t sy, int64_t sz,
                                      int connectivity, uint32_t* out, uint64_t* N, void* stream) {
  if (!vcg || !out || !N || sx <= 0 || sy <= 0 || sz <= 0 || (vcg_size != 1 && vcg_size != 4))
    return cfail(CC3D_AMD_ERR_INVALID, "cc3d_amd_color_connectivity_graph: bad argument");
  // the reference's dispatch (cc3d/cc3d_graphs.hpp:1077-1106)
  if (connectivity != 26 && connectivity != 6 && connectivity != 4 && connectivity != 8)
    return cfail(CC3D_AMD_ERR_RUNTIME, "Only 4, 8, 6 and 26 connectivities are supported.");
  if (sz > 1 && connectivity != 6 && connectivity != 26)
    return cfail(CC3D_AMD_ERR_RUNTIME, "Only 6 and 26 connectivity is supported in 3D.");
  int var;
  if (sz == 1) var = (connectivity == 6 || connectivity == 4) ? 4 : (vcg_size == 1 ? 2 : 3);
  else var = connectivity == 6 ? 1 : 0;
  if (var == 0 && vcg_size != 4)
    return cfail(CC3D_AMD_ERR_RUNTIME, "26-connectivity requires a 32-bit voxel graph for color_connectivity_graph_26.");
  const int64_t V = sx * sy * sz;
  if (V >= 0xFFFFFFFFll) return cfail(CC3D_AMD_ERR_INVALID, "volumes of 2^32-1 voxels or more are not supported");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t nwords = (V + 31) / 32;
  const int64_t nblk = (nwords + 4095) / 4096;
  auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
  const size_t oGP = 0, oB = oGP + al((size_t)V * 4), oP = oB + al((size_t)nwords * 4),
               oS = oP + al((size_t)nwords * 4), oC = oS + al((size_t)(nblk + 1) * 4), total = oC + al(sizeof(Counters));
  void* ws = nullptr;
  if (hipMallocAsync(&ws, total, s) != hipSuccess || !ws)
    return cfail(CC3D_AMD_ERR_NOMEM, "cc3d_amd_color_connectivity_graph: could not allocate workspace");
  char* w = static_cast<char*>(ws);
  uint32_t* GP = reinterpret_cast<uint32_t*>(w + oGP);
  uint32_t* bits = reinterpret_cast<uint32_t*>(w + oB);
  uint32_t* prefix = reinterpret_cast<uint32_t*>(w + oP);
  uint32_t* bsum = reinterpret_cast<uint32_t*>(w + oS);
  Counters* ctr = reinterpret_cast<Counters*>(w + oC);
  const unsigned gb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((V + WG - 1) / WG, 65536));
  hipLaunchKernelGGL(k_ccg_clear, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((nwords + WG - 1) / WG, 8192))),
                     dim3(WG), 0, s, bits, nwords, ctr);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    e = vcg_size == 1 ? ccg_launch<uint8_t>(static_cast<const uint8_t*>(vcg), sx, sy, sz, var, GP, s)
                      : ccg_launch<uint32_t>(static_cast<const uint32_t*>(vcg), sx, sy, sz, var, GP, s);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_ccg_roots, dim3(gb), dim3(WG), 0, s, GP, V, bits);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = scan_root_bits(bits, nwords, bsum, prefix, ctr, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_ccg_final, dim3(gb), dim3(WG), 0, s, GP, V, bits, prefix, out);
    e = hipGetLastError();
  }
  Counters h{};
  if (e == hipSuccess) e = hipMemcpyAsync(&h, ctr, sizeof(Counters), hipMemcpyDeviceToHost, s);
  const hipError_t ef = hipFreeAsync(ws, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = ef;
  if (e != hipSuccess) return cfail(CC3D_AMD_ERR_HIP, std::string("color_connectivity_graph: ") + hipGetErrorString(e));
  *N = h.N;
  return CC3D_AMD_OK;
}

}  // extern "C"
