// C-ABI shim over the reference cc3d headers, built only into oracle/_ref/.
//
// Test infrastructure: this file is the checker's binding to the unmodified
// reference implementation (cc3d/cc3d_continuous.hpp:394-455 and
// cc3d/cc3d.hpp:287-315), compiled in place from the reference tree by
// oracle/Makefile.  Nothing in the product links or loads it; tests, the
// smoke check and the bench's CPU-baseline leg may.
//
// Mirrors the Cython dispatch in cc3d/fastcc3d.pyx:471-613: the input is
// already viewed as one of u8/u16/u32/u64/f32/f64 and the output is one of
// u16/u32/u64.  Exceptions (std::runtime_error, e.g. the union-find
// "maximum length exception" of cc3d/cc3d.hpp:107-111) come back as status 1
// with the message copied into errbuf.

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <type_traits>

#include "cc3d_continuous.hpp"
#include "cc3d_graphs.hpp"

namespace {

enum : int { DT_U8 = 0, DT_U16 = 1, DT_U32 = 2, DT_U64 = 3, DT_F32 = 4, DT_F64 = 5 };

template <typename T, typename U>
void run(const void* in, int64_t sx, int64_t sy, int64_t sz, int64_t max_labels,
         int64_t connectivity, T delta, void* out, uint64_t* N, bool periodic,
         bool binary) {
  size_t n = 0;
  cc3d::connected_components3d<T, U>(
      const_cast<T*>(static_cast<const T*>(in)), sx, sy, sz, max_labels,
      connectivity, delta, static_cast<U*>(out), n, periodic, binary);
  *N = n;
}

template <typename T>
void run_out(const void* in, int64_t sx, int64_t sy, int64_t sz,
             int64_t max_labels, int64_t connectivity, T delta, void* out,
             int out_dtype, uint64_t* N, bool periodic, bool binary) {
  if (out_dtype == DT_U16)
    run<T, uint16_t>(in, sx, sy, sz, max_labels, connectivity, delta, out, N, periodic, binary);
  else if (out_dtype == DT_U32)
    run<T, uint32_t>(in, sx, sy, sz, max_labels, connectivity, delta, out, N, periodic, binary);
  else if (out_dtype == DT_U64)
    run<T, uint64_t>(in, sx, sy, sz, max_labels, connectivity, delta, out, N, periodic, binary);
  else
    throw std::invalid_argument("bad out dtype");
}

}  // namespace

template <class T>
static int64_t region_graph_of(const void* labels, int64_t sx, int64_t sy, int64_t sz, float wx, float wy, float wz,
                               int connectivity, int surface_area, uint64_t* lo, uint64_t* hi, float* val,
                               int64_t cap) {
  const auto res = cc3d::extract_region_graph<T>(static_cast<T*>(const_cast<void*>(labels)), sx, sy, sz, wx, wy, wz,
                                                 connectivity, surface_area != 0);
  if ((int64_t)res.size() > cap) return -1;
  int64_t m = 0;
  for (const auto& kv : res) {
    lo[m] = (uint64_t)(int64_t)kv.first.first;
    hi[m] = (uint64_t)(int64_t)kv.first.second;
    if (!std::is_signed<T>::value) {
      lo[m] = (uint64_t)kv.first.first;
      hi[m] = (uint64_t)kv.first.second;
    }
    val[m] = kv.second;
    m++;
  }
  return m;
}

template <class T>
static int64_t runs_of(const void* labels, uint64_t n, uint64_t* label, uint64_t* start, uint64_t* end, int64_t cap) {
  const auto res = cc3d::extract_runs<T>(static_cast<T*>(const_cast<void*>(labels)), n);
  int64_t m = 0;
  for (const auto& kv : res)
    for (const auto& r : kv.second) {
      if (m >= cap) return -1;
      label[m] = (uint64_t)kv.first;
      start[m] = r.first;
      end[m] = r.second;
      m++;
    }
  return m;
}

extern "C" {

// Returns 0 on success, 1 on a C++ exception (message in errbuf).
int ref_connected_components3d(const void* in, int in_dtype, int64_t sx,
                               int64_t sy, int64_t sz, int64_t max_labels,
                               int64_t connectivity, double delta_f,
                               uint64_t delta_u, void* out, int out_dtype,
                               uint64_t* N, int periodic, int binary,
                               char* errbuf, int errlen) {
  try {
    switch (in_dtype) {
      case DT_U8:
        run_out<uint8_t>(in, sx, sy, sz, max_labels, connectivity, (uint8_t)delta_u, out, out_dtype, N, periodic, binary);
        break;
      case DT_U16:
        run_out<uint16_t>(in, sx, sy, sz, max_labels, connectivity, (uint16_t)delta_u, out, out_dtype, N, periodic, binary);
        break;
      case DT_U32:
        run_out<uint32_t>(in, sx, sy, sz, max_labels, connectivity, (uint32_t)delta_u, out, out_dtype, N, periodic, binary);
        break;
      case DT_U64:
        run_out<uint64_t>(in, sx, sy, sz, max_labels, connectivity, (uint64_t)delta_u, out, out_dtype, N, periodic, binary);
        break;
      case DT_F32:
        run_out<float>(in, sx, sy, sz, max_labels, connectivity, (float)delta_f, out, out_dtype, N, periodic, binary);
        break;
      case DT_F64:
        run_out<double>(in, sx, sy, sz, max_labels, connectivity, delta_f, out, out_dtype, N, periodic, binary);
        break;
      default:
        throw std::invalid_argument("bad input dtype");
    }
  } catch (const std::exception& e) {
    if (errbuf && errlen > 0) {
      std::strncpy(errbuf, e.what(), errlen - 1);
      errbuf[errlen - 1] = 0;
    }
    return 1;
  }
  return 0;
}

// cc3d/cc3d.hpp:287-315 through cc3d/fastcc3d.pyx:168-242.
uint64_t ref_estimate_provisional_label_count(const void* in, int in_dtype,
                                              int64_t sx, int64_t voxels,
                                              int64_t* first_row,
                                              int64_t* last_row) {
  int64_t f = 0, l = 0;
  size_t r = 0;
  switch (in_dtype) {
    case DT_U8: r = cc3d::estimate_provisional_label_count<uint8_t>((uint8_t*)in, sx, voxels, f, l); break;
    case DT_U16: r = cc3d::estimate_provisional_label_count<uint16_t>((uint16_t*)in, sx, voxels, f, l); break;
    case DT_U32: r = cc3d::estimate_provisional_label_count<uint32_t>((uint32_t*)in, sx, voxels, f, l); break;
    case DT_U64: r = cc3d::estimate_provisional_label_count<uint64_t>((uint64_t*)in, sx, voxels, f, l); break;
    case DT_F32: r = cc3d::estimate_provisional_label_count<float>((float*)in, sx, voxels, f, l); break;
    case DT_F64: r = cc3d::estimate_provisional_label_count<double>((double*)in, sx, voxels, f, l); break;
    default: return 0;
  }
  *first_row = f;
  *last_row = l;
  return r;
}

// cc3d/cc3d_graphs.hpp:192-236 through cc3d/fastcc3d.pyx:1100-1166: the
// labels are one of u8/u16/u32/u64 (x fastest), out is zeroed uint8 (out_size
// 1) or uint32 (4) voxels.  Returns 0, or 1 on a C++ exception.
int ref_voxel_connectivity_graph(const void* in, int in_dtype, int64_t sx, int64_t sy, int64_t sz,
                                 int64_t connectivity, void* out, int out_size) {
  try {
    auto go = [&](auto* lab) {
      using T = std::remove_const_t<std::remove_pointer_t<decltype(lab)>>;
      T* l = const_cast<T*>(lab);
      if (out_size == 1)
        cc3d::extract_voxel_connectivity_graph<T, uint8_t>(l, sx, sy, sz, connectivity, static_cast<uint8_t*>(out));
      else
        cc3d::extract_voxel_connectivity_graph<T, uint32_t>(l, sx, sy, sz, connectivity, static_cast<uint32_t*>(out));
    };
    switch (in_dtype) {
      case DT_U8: go(static_cast<const uint8_t*>(in)); break;
      case DT_U16: go(static_cast<const uint16_t*>(in)); break;
      case DT_U32: go(static_cast<const uint32_t*>(in)); break;
      case DT_U64: go(static_cast<const uint64_t*>(in)); break;
      default: return 1;
    }
  } catch (const std::exception&) {
    return 1;
  }
  return 0;
}

// cc3d/cc3d_graphs.hpp:1077-1106 (color_connectivity_graph_N) through
// cc3d/fastcc3d.pyx:941-1018: vcg is uint8 (vcg_size 1) or uint32 (4), x
// fastest; out is zeroed uint32.  Returns 0, or 1 on a C++ exception
// (message in errbuf).
int ref_color_connectivity_graph(const void* vcg, int vcg_size, int64_t sx, int64_t sy, int64_t sz,
                                 int connectivity, uint32_t* out, uint64_t* N, char* errbuf, int errlen) {
  try {
    size_t n = 0;
    if (vcg_size == 1)
      cc3d::color_connectivity_graph_N<uint8_t, uint32_t>(static_cast<const uint8_t*>(vcg), sx, sy, sz, connectivity,
                                                          out, n);
    else
      cc3d::color_connectivity_graph_N<uint32_t, uint32_t>(static_cast<const uint32_t*>(vcg), sx, sy, sz,
                                                           connectivity, out, n);
    *N = n;
  } catch (const std::exception& e) {
    if (errbuf && errlen > 0) {
      std::strncpy(errbuf, e.what(), errlen - 1);
      errbuf[errlen - 1] = 0;
    }
    return 1;
  } catch (...) {  // the reference throws a pointer for 26 with a uint8 graph
    if (errbuf && errlen > 0) {
      std::strncpy(errbuf, "26-connectivity requires a 32-bit voxel graph for color_connectivity_graph_26.", errlen - 1);
      errbuf[errlen - 1] = 0;
    }
    return 1;
  }
  return 0;
}

// cc3d/cc3d_graphs.hpp:430-468 (extract_region_graph<T>) through
// cc3d/fastcc3d.pyx:1230-1252: labels x fastest, dtype 0 u8, 1 u16, 2 u32,
// 3 u64, 6 i8, 7 i16, 8 i32, 9 i64.  Pairs are written as 64-bit words
// (signed labels sign-extended) with their sums, at most cap of them.
// Returns the pair count, -1 when cap is too small, -2 on a C++ exception.
int64_t ref_contacts(const void* labels, int dtype, int64_t sx, int64_t sy, int64_t sz, int connectivity,
                     int surface_area, float wx, float wy, float wz, uint64_t* lo, uint64_t* hi, float* val,
                     int64_t cap, char* errbuf, int errlen) {
  try {
    switch (dtype) {
      case 0: return region_graph_of<uint8_t>(labels, sx, sy, sz, wx, wy, wz, connectivity, surface_area, lo, hi, val, cap);
      case 1: return region_graph_of<uint16_t>(labels, sx, sy, sz, wx, wy, wz, connectivity, surface_area, lo, hi, val, cap);
      case 2: return region_graph_of<uint32_t>(labels, sx, sy, sz, wx, wy, wz, connectivity, surface_area, lo, hi, val, cap);
      case 3: return region_graph_of<uint64_t>(labels, sx, sy, sz, wx, wy, wz, connectivity, surface_area, lo, hi, val, cap);
      case 6: return region_graph_of<int8_t>(labels, sx, sy, sz, wx, wy, wz, connectivity, surface_area, lo, hi, val, cap);
      case 7: return region_graph_of<int16_t>(labels, sx, sy, sz, wx, wy, wz, connectivity, surface_area, lo, hi, val, cap);
      case 8: return region_graph_of<int32_t>(labels, sx, sy, sz, wx, wy, wz, connectivity, surface_area, lo, hi, val, cap);
      default: return region_graph_of<int64_t>(labels, sx, sy, sz, wx, wy, wz, connectivity, surface_area, lo, hi, val, cap);
    }
  } catch (const std::exception& e) {
    if (errbuf && errlen > 0) {
      std::strncpy(errbuf, e.what(), errlen - 1);
      errbuf[errlen - 1] = 0;
    }
    return -2;
  }
}

// cc3d/cc3d_graphs.hpp:470-501 (extract_runs<T>) through
// cc3d/fastcc3d.pyx:1254-1277: the runs of each label in label order, each
// label's runs in position order, flattened as (label, start, end).
int64_t ref_runs(const void* labels, int dtype, uint64_t n, uint64_t* label, uint64_t* start, uint64_t* end,
                 int64_t cap) {
  switch (dtype) {
    case 0: return runs_of<uint8_t>(labels, n, label, start, end, cap);
    case 1: return runs_of<uint16_t>(labels, n, label, start, end, cap);
    case 2: return runs_of<uint32_t>(labels, n, label, start, end, cap);
    default: return runs_of<uint64_t>(labels, n, label, start, end, cap);
  }
}

}  // extern "C"
