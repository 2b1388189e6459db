// Shared helpers of the gfx950 sea-ice kernels: error reporting, domain
// geometry, numpy.roll-style wrapped indexing and the overlap ("halo") fill
// performed as mirror stores by the thread that computes the source cell.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>

#include "veris_hip.h"

namespace vh {

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

#define VH_TRY_HIP(expr)                                                        \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      ::vh::set_error("%s failed: %s", #expr, hipGetErrorString(e_));          \
      return VH_EHIP;                                                           \
    }                                                                           \
  } while (0)

#define VH_REQUIRE(cond, code, ...)                                             \
  do {                                                                          \
    if (!(cond)) {                                                              \
      ::vh::set_error(__VA_ARGS__);                                             \
      return (code);                                                            \
    }                                                                           \
  } while (0)

// Padded geometry of one domain.  i = axis 0 (x), j = axis 1 (y, contiguous).
struct Geom {
  int NX, NY;    // padded extents
  int nx, ny;    // interior extents
  int ol;        // overlap width (2)
  int fill_x;    // computed cells are mirrored into the x overlap
  int fill_y;    // computed cells are mirrored into the y overlap
};

// Validates a domain descriptor and derives the padded geometry.
int make_geom(const vh_domain* d, Geom* g);

inline size_t field_elems(const Geom& g) { return (size_t)g.NX * (size_t)g.NY; }
inline size_t elem_bytes(int dtype) { return dtype == VH_F64 ? 8 : 4; }
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

__host__ __device__ inline int wrap(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// numpy.sign for finite or infinite inputs (0 for +-0).
template <typename T>
__device__ inline T sgn(T x) {
  return (T)((x > (T)0) - (x < (T)0));
}

// True when cell (i, j) is computed by its own thread: cells inside an overlap
// that the fill overwrites take their value from a mirror store instead.
__device__ inline bool is_computed(const Geom& g, int i, int j) {
  if (g.fill_x && (i < g.ol || i >= g.nx + g.ol)) return false;
  if (g.fill_y && (j < g.ol || j >= g.ny + g.ol)) return false;
  return true;
}

// Stores the value of computed cell (i, j) at (i, j) and at every overlap cell
// that the periodic fill copies from it (x first, then y, as the fill does).
template <typename P, typename T>
__device__ inline void store_filled(const Geom& g, P dst, int i, int j, T val) {
  int ix[2] = {i, -1};
  int jy[2] = {j, -1};
  if (g.fill_x) {
    if (i < 2 * g.ol) ix[1] = i + g.nx;
    else if (i >= g.nx) ix[1] = i - g.nx;
  }
  if (g.fill_y) {
    if (j < 2 * g.ol) jy[1] = j + g.ny;
    else if (j >= g.ny) jy[1] = j - g.ny;
  }
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    if (ix[a] < 0) continue;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      if (jy[b] < 0) continue;
      dst[ix[a] * g.NY + jy[b]] = val;
    }
  }
}

// Replays the launch sequence that `record(stream)` issues as a hipGraph on
// `stream`, capturing it once per (device, key) on a private stream.  The
// key must name everything the recorded launches bake in (kernel arguments,
// grid sizes).  At most 256 graphs are kept per device; past that, new keys
// issue their launches directly on `stream`.  Returns VH_OK or VH_EHIP.
int graph_run(const std::string& key, const std::function<void(hipStream_t)>& record,
              hipStream_t stream);

// Whether graph replay is enabled (environment VERIS_EVP_GRAPH=0 disables it).
int graphs_enabled();

// Whether EVP subcycles under the default settings launch the substep
// instantiation with the optional branches compiled out (read at every call;
// environment VERIS_EVP_PLAIN=0 selects the general instantiation, which gives
// the same results bit for bit).
int evp_plain_kernels_enabled();

// 1-D launch helper: one thread per padded cell.
inline dim3 cell_grid(const Geom& g, int block) {
  size_t n = field_elems(g);
  return dim3((unsigned)((n + block - 1) / block));
}

}  // namespace vh
