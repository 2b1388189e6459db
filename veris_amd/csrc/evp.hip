// Single-domain EVP subcycle (veris/evp_solver.py:257-319): host side of the
// C ABI vh_evp_solver.  The substep kernel is evp_substep_kernel
// (evp_kernel.h); this file is built with -ffp-contract=off so the FP64
// instantiation evaluates the reference expressions in their own order.
//
// Launch plan per call:
//   1. one tiny kernel copies the argument block (all field pointers, the
//      ping-pong state pointers and the scalars) into the caller's workspace;
//   2. the n substep launches read it from there.  For n >= 8 they are
//      replayed from a hipGraph captured once per (device, geometry, tile, n,
//      workspace address), so moving inputs and outputs never forces a
//      re-capture and a 400-substep subcycle costs one graph launch.
#include <cstdio>
#include <mutex>

#include "evp_kernel.h"

namespace vh {
namespace {

int pick_tile(const Geom& g, int req) {
  if (req > 0 && req < kEvpNumTiles) return req;
  const long long cells = (long long)g.NX * g.NY;
  return cells < 400000 ? 5 : 3;
}

template <typename T, typename TD>
void launch_substeps(int cfg, const EvpArgBlock<T>* ab, int nsub, const Geom& g, bool so,
                     bool res, bool plain, hipStream_t st) {
  if (res) {  // one residual-writing launch per substep
    const int TI = evp_tile_extent_i(kEvpResidualCfg), TJ = evp_tile_extent_j(kEvpResidualCfg);
    const int tiles_j = (g.NY + TJ - 1) / TJ;
    const int n_tiles = ((g.NX + TI - 1) / TI) * tiles_j;
    for (int l = 0; l < nsub; ++l) {
      if (so) evp_launch_res<T, true>(ab, l, n_tiles, tiles_j, st);
      else evp_launch_res<T, false>(ab, l, n_tiles, tiles_j, st);
    }
    return;
  }
  const int K = kEvpTiles[cfg].k;
  int l = 0;
  for (int q = 0; q < nsub / K + nsub % K; ++q) {
    const int c = q < nsub / K ? cfg : kEvpRemainderCfg;
    const int TI = evp_tile_extent_i(c), TJ = evp_tile_extent_j(c);
    const int tiles_j = (g.NY + TJ - 1) / TJ;
    const int n_tiles = ((g.NX + TI - 1) / TI) * tiles_j;
    evp_launch_cfg<T, TD, false>(c, ab, 0, l++, n_tiles, tiles_j, st, so, plain);
  }
}

// Residual norms per substep from the per-workgroup partials [substep][tile][2]
// (stress, velocity): one workgroup per substep, fixed summation order.
__global__ __launch_bounds__(256) void evp_residual_reduce_kernel(const double* part, int n_tiles,
                                                                  double* resU, double* resSig) {
  __shared__ double s0[256], s1[256];
  const int k = blockIdx.x, t = threadIdx.x;
  const double* p = part + (size_t)k * n_tiles * 2;
  double a0 = 0.0, a1 = 0.0;
  for (int q = t; q < n_tiles; q += 256) {
    a0 += p[2 * q];
    a1 += p[2 * q + 1];
  }
  s0[t] = a0;
  s1[t] = a1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      s0[t] += s0[t + w];
      s1[t] += s1[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    resSig[k] = s0[0];
    resU[k] = s1[0];
  }
}

}  // namespace

// Workspace of one single-domain subcycle: the ping-pong state, the argument
// block and the packed constants.
template <typename T>
size_t evp_core_ws_bytes(const Geom& g) {
  return 5 * align256(field_elems(g) * sizeof(T)) + align256(sizeof(EvpArgBlock<T>)) +
         evp_pack_bytes<T>(field_elems(g), g.NY);
}

int evp_residual_tiles(const Geom& g) {
  const int TI = evp_tile_extent_i(kEvpResidualCfg), TJ = evp_tile_extent_j(kEvpResidualCfg);
  return ((g.NX + TI - 1) / TI) * ((g.NY + TJ - 1) / TJ);
}

// Subcycle of nsub substeps from the state `in` (u, v, sigma1, sigma2,
// sigma12) to `out`; fl->uIce/vIce are the previous-step velocities of the
// momentum equation.  Shared by vh_evp_solver and the checkpointed forward
// and segment recompute of the adjoint.
// `direct`: launch the substeps on `st` one by one, never through a graph
// (no capture, no instantiation: nothing that allocates).
template <typename T>
int evp_run_core(const Geom& g, const vh_settings* s, const vh_fields* fl, const T* const in[5],
                 T* const out[5], int nsub, void* ws, size_t ws_bytes, int tile_req, int dtype,
                 hipStream_t st, double* res = nullptr, bool direct = false) {
  const size_t fb = align256(field_elems(g) * sizeof(T));
  const size_t need = 5 * fb + align256(sizeof(EvpArgBlock<T>));
  // the substep kernel addresses the fields with 32-bit byte offsets
  VH_REQUIRE(field_elems(g) * sizeof(T) < ((size_t)1 << 32), VH_ESHAPE,
             "domain too large: %d x %d cells exceed 4 GiB per field", g.nx, g.ny);
  VH_REQUIRE(ws != nullptr && ws_bytes >= need, VH_EWORKSPACE,
             "evp workspace: need %zu bytes, got %zu", need, ws_bytes);
  // packed constants when the workspace has room for them (its full size,
  // vh_evp_workspace_bytes); a workspace of the minimal size reads full fields
  const bool pack = ws_bytes >= evp_core_ws_bytes<T>(g);
  const void* req[] = {fl->uIce, fl->vIce, fl->sigma1, fl->sigma2, fl->sigma12,
                       fl->SeaIceStrength, fl->SeaIceMassC, fl->SeaIceMassU, fl->SeaIceMassV,
                       fl->AreaW, fl->AreaS, fl->WindForcingX, fl->WindForcingY, fl->uOcean,
                       fl->vOcean, fl->fCori, fl->iceMask, fl->iceMaskU, fl->iceMaskV,
                       fl->maskInC, fl->maskInU, fl->maskInV, fl->recip_dxU, fl->recip_dyV,
                       fl->recip_dyU, fl->recip_dxV, fl->dxU, fl->dyU, fl->dxV, fl->dyV,
                       fl->recip_rAu, fl->recip_rAv, fl->rAz, fl->recip_rA};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "evp_solver: missing input field");
  for (int q = 0; q < 5; ++q) {
    VH_REQUIRE(in[q] != nullptr, VH_EINVAL, "evp_solver: null input state");
    VH_REQUIRE(out[q] != nullptr, VH_EINVAL, "evp_solver: null output");
    for (const void* p : req)
      VH_REQUIRE(out[q] != p, VH_EINVAL, "evp_solver: output aliases an input field");
    for (int r = 0; r < 5; ++r)
      VH_REQUIRE(out[q] != in[r], VH_EINVAL, "evp_solver: output aliases the input state");
    for (int r = 0; r < q; ++r)
      VH_REQUIRE(out[q] != out[r], VH_EINVAL, "evp_solver: two outputs alias");
  }
  if (s->noSlip == 0.0 && s->use_coastline != 0.0)
    VH_REQUIRE(fl->Fu && fl->Fv, VH_EINVAL, "lateral drag with use_coastline needs Fu, Fv");
  const int kt = (fl->k1AtC || fl->k2AtC || fl->k1AtZ || fl->k2AtZ) ? 1 : 0;
  VH_REQUIRE(!kt || (fl->k1AtC && fl->k2AtC && fl->k1AtZ && fl->k2AtZ), VH_EINVAL,
             "k1AtC/k2AtC/k1AtZ/k2AtZ must be all set or all NULL");
  const int basal = s->basalDragK2 != 0.0 ? 1 : 0;
  if (basal)
    VH_REQUIRE(fl->R_low && fl->Area && fl->hIceMean, VH_EINVAL,
               "basal drag needs R_low, Area, hIceMean");

  EvpArgBlock<T> ab{};
  fill_evp_scalars(s, kt, basal, &ab.sc);
  fill_evp_const(fl, &ab.c);

  const T* const* inp = in;
  T* const* outp = out;
  const size_t nb = field_elems(g) * sizeof(T);
  if (nsub == 0) {
    for (int q = 0; q < 5; ++q)
      VH_TRY_HIP(hipMemcpyAsync(outp[q], inp[q], nb, hipMemcpyDeviceToDevice, st));
    return VH_OK;
  }
  for (int q = 0; q < 5; ++q) {
    ab.in[q] = inp[q];
    ab.out[q] = outp[q];
    ab.alt[q] = (T*)((char*)ws + q * fb);
  }
  const int cfg = res ? kEvpResidualCfg : pick_tile(g, tile_req);
  ab.res = res;
  VH_REQUIRE(!ab.sc.sobc || kEvpTiles[cfg].k == 1, VH_EUNSUPPORTED,
             "secondOrderBC needs a single-substep tile shape (tile %d has k=%d)", cfg,
             kEvpTiles[cfg].k);
  const bool so = ab.sc.sobc != 0;
  const bool plain = !res && evp_plain_kernels_enabled() && evp_plain_settings(ab.sc);
  ab.g = g;
  ab.nlaunch = evp_n_launches(cfg, nsub);
  ab.members = nullptr;
  if (pack) evp_pack_attach(&ab, (char*)ws + need, field_elems(g), g.NY);
  EvpArgBlock<T>* ab_dev = (EvpArgBlock<T>*)((char*)ws + 5 * fb);
  hipLaunchKernelGGL(evp_write_args_kernel<T>, dim3(1), dim3(64), 0, st, ab_dev, ab);
  if (pack) evp_pack_launch(ab_dev, (long long)field_elems(g), 1, st);
  VH_TRY_HIP(hipGetLastError());
  if (nsub >= 8 && graphs_enabled() && !direct) {
    char key[160];
    snprintf(key, sizeof(key), "evp1/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%p", dtype, cfg, (int)so,
             res ? 1 : 0, (int)plain, nsub, g.NX, g.NY, g.nx, g.fill_x, g.fill_y,
             (const void*)ab_dev);
    int rc = graph_run(
        key,
        [&](hipStream_t cs) { launch_substeps<T, double>(cfg, ab_dev, nsub, g, so, res, plain, cs); },
        st);
    if (rc) return rc;
  } else {
    launch_substeps<T, double>(cfg, ab_dev, nsub, g, so, res, plain, st);
  }
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template int evp_run_core<double>(const Geom&, const vh_settings*, const vh_fields*,
                                  const double* const[5], double* const[5], int, void*, size_t,
                                  int, int, hipStream_t, double*, bool);
template int evp_run_core<float>(const Geom&, const vh_settings*, const vh_fields*,
                                 const float* const[5], float* const[5], int, void*, size_t, int,
                                 int, hipStream_t, double*, bool);

// Scratch of vh_evp_solver_noalloc: the workspace of one subcycle (ping-pong
// state, argument block, packed step constants) in a device array of the
// library, one copy per GPU, held until the process ends.  Host code never
// passes `scratch` itself (on the host it only names the symbol): it uses the
// device address hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchBytes = (size_t)VH_EVP_NOALLOC_MAX_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up once per
// device (with that device current) and kept.
int scratch_for_current_device(void** out) {
  int dev = -1;
  VH_TRY_HIP(hipGetDevice(&dev));
  VH_REQUIRE(dev >= 0 && dev < kMaxDevices, VH_EINVAL, "device %d beyond the library's %d", dev,
             kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    VH_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    VH_REQUIRE(p != nullptr, VH_EHIP, "hipGetSymbolAddress gave no address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return VH_OK;
}
}  // namespace

namespace {

template <typename T>
int run_evp(const Geom& g, const vh_settings* s, const vh_fields* fl, int nsub, void* uo, void* vo,
            void* s1o, void* s2o, void* s12o, void* ws, size_t ws_bytes, int tile_req,
            int dtype, hipStream_t st) {
  const T* in[5] = {evp_ptr<T>(fl->uIce), evp_ptr<T>(fl->vIce), evp_ptr<T>(fl->sigma1),
                    evp_ptr<T>(fl->sigma2), evp_ptr<T>(fl->sigma12)};
  T* const out[5] = {(T*)uo, (T*)vo, (T*)s1o, (T*)s2o, (T*)s12o};
  return evp_run_core<T>(g, s, fl, in, out, nsub, ws, ws_bytes, tile_req, dtype, st);
}

// Subcycle with residual norms: the workspace holds the plain subcycle's
// workspace followed by the residual partials.
template <typename T>
int run_evp_res(const Geom& g, const vh_settings* s, const vh_fields* fl, int nsub, void* const o[5],
                double* resU, double* resSig, void* ws, size_t ws_bytes, int dtype, hipStream_t st) {
  VH_REQUIRE(nsub >= 1, VH_EINVAL, "residuals need n_substeps >= 1");
  VH_REQUIRE(resU && resSig && resU != resSig, VH_EINVAL, "null or aliased residual outputs");
  const size_t base = evp_core_ws_bytes<T>(g);
  const int nt = evp_residual_tiles(g);
  const size_t need = base + (size_t)nsub * nt * 2 * sizeof(double);
  VH_REQUIRE(ws != nullptr && ws_bytes >= need, VH_EWORKSPACE,
             "evp residual workspace: need %zu bytes, got %zu", need, ws_bytes);
  double* part = (double*)((char*)ws + base);
  const T* in[5] = {evp_ptr<T>(fl->uIce), evp_ptr<T>(fl->vIce), evp_ptr<T>(fl->sigma1),
                    evp_ptr<T>(fl->sigma2), evp_ptr<T>(fl->sigma12)};
  T* const out[5] = {(T*)o[0], (T*)o[1], (T*)o[2], (T*)o[3], (T*)o[4]};
  int rc = evp_run_core<T>(g, s, fl, in, out, nsub, ws, base, 0, dtype, st, part);
  if (rc) return rc;
  hipLaunchKernelGGL(evp_residual_reduce_kernel, dim3(nsub), dim3(256), 0, st, part, nt, resU,
                     resSig);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

}  // namespace
}  // namespace vh

namespace vh {
namespace {

// Fields a multi-substep launch reads at neighbouring cells: every cell of a
// filled overlap must hold what the fill copies there.
constexpr int kPeriodicFull = 30;
template <typename T>
struct PeriodicCheck {
  const T* full[kPeriodicFull];  // null entries are skipped
  const T *s1, *s2, *s12;        // stresses: only the overlap cells that reach the interior
  int* flag;
};

__device__ inline bool same_bits(double a, double b) {
  return __double_as_longlong(a) == __double_as_longlong(b);
}
__device__ inline bool same_bits(float a, float b) { return __float_as_uint(a) == __float_as_uint(b); }

template <typename T>
__global__ __launch_bounds__(256) void overlaps_init_kernel(int* flag) {
  if (threadIdx.x == 0) *flag = 1;
}

// One thread per padded cell; cells the fill overwrites compare each field
// with the cell the fill copies from (x first, then y, as the fill does) and
// clear the flag on any difference in any bit.
template <typename T>
__global__ __launch_bounds__(256) void overlaps_periodic_kernel(Geom g, PeriodicCheck<T> c) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)g.NX * g.NY) return;
  const int i = (int)(e / g.NY), j = (int)(e - (size_t)i * g.NY);
  if (is_computed(g, i, j)) return;
  int si = i, sj = j;
  if (g.fill_x) si = i < g.ol ? i + g.nx : (i >= g.nx + g.ol ? i - g.nx : i);
  if (g.fill_y) sj = j < g.ol ? j + g.ny : (j >= g.ny + g.ol ? j - g.ny : j);
  const size_t src = (size_t)si * g.NY + sj;
  bool ok = true;
#pragma unroll 1
  for (int q = 0; q < kPeriodicFull; ++q)
    if (c.full[q] && !same_bits(c.full[q][e], c.full[q][src])) ok = false;
  // sigma1, sigma2 at c-points of overlap column (row) 1 and sigma12 at z-points
  // of column (row) NX-2 (NY-2) enter the interior's next substep; the other
  // overlap stresses are wrap-around values nothing interior reads
  const bool c1 = (g.fill_x && i == 1) || (g.fill_y && j == 1);
  const bool z2 = (g.fill_x && i == g.NX - 2) || (g.fill_y && j == g.NY - 2);
  if (c1 && (!same_bits(c.s1[e], c.s1[src]) || !same_bits(c.s2[e], c.s2[src]))) ok = false;
  if (z2 && !same_bits(c.s12[e], c.s12[src])) ok = false;
  if (!ok) *c.flag = 0;
}

template <typename T>
int overlaps_periodic(const Geom& g, const vh_settings* s, const vh_fields* f, int* flag,
                      hipStream_t st) {
  PeriodicCheck<T> c{};
  const void* full[] = {f->uIce,     f->vIce,      f->uOcean,    f->vOcean,    f->iceMask,
                        f->iceMaskU, f->iceMaskV,  f->maskInC,   f->maskInU,   f->maskInV,
                        f->recip_dyU, f->recip_dxV, f->rAz,      f->recip_dxU, f->recip_dyV,
                        f->recip_rA, f->dyV,       f->dxU,       f->dxV,       f->dyU,
                        f->SeaIceStrength, f->fCori, f->SeaIceMassC, f->k1AtC, f->k2AtC,
                        f->k1AtZ,    f->k2AtZ,     nullptr,      nullptr,      nullptr};
  static_assert(sizeof(full) / sizeof(full[0]) == kPeriodicFull, "field list");
  for (int q = 0; q < kPeriodicFull; ++q) c.full[q] = (const T*)full[q];
  if (s->basalDragK2 != 0.0) {  // the basal drag's inputs at the c-points
    c.full[27] = (const T*)f->R_low;
    c.full[28] = (const T*)f->Area;
    c.full[29] = (const T*)f->hIceMean;
  }
  const void* req[] = {f->uIce, f->vIce, f->sigma1, f->sigma2, f->sigma12, f->uOcean,
                       f->vOcean, f->iceMask, f->iceMaskU, f->iceMaskV, f->maskInC, f->maskInU,
                       f->maskInV, f->recip_dyU, f->recip_dxV, f->rAz, f->recip_dxU,
                       f->recip_dyV, f->recip_rA, f->dyV, f->dxU, f->dxV, f->dyU,
                       f->SeaIceStrength, f->fCori, f->SeaIceMassC};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "overlaps_periodic: missing field");
  if (s->basalDragK2 != 0.0)
    VH_REQUIRE(f->R_low && f->Area && f->hIceMean, VH_EINVAL,
               "overlaps_periodic: basal drag needs R_low, Area, hIceMean");
  c.s1 = (const T*)f->sigma1;
  c.s2 = (const T*)f->sigma2;
  c.s12 = (const T*)f->sigma12;
  c.flag = flag;
  hipLaunchKernelGGL(overlaps_init_kernel<T>, dim3(1), dim3(64), 0, st, flag);
  if (g.fill_x || g.fill_y)
    hipLaunchKernelGGL(overlaps_periodic_kernel<T>, dim3((unsigned)((field_elems(g) + 255) / 256)),
                       dim3(256), 0, st, g, c);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

}  // namespace
}  // namespace vh

extern "C" int vh_evp_overlaps_periodic(const vh_domain* d, const vh_settings* s,
                                        const vh_fields* f, int32_t* flag, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(s != nullptr && f != nullptr && flag != nullptr, VH_EINVAL,
             "null settings, fields or flag");
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64) return vh::overlaps_periodic<double>(g, s, f, (int*)flag, st);
  return vh::overlaps_periodic<float>(g, s, f, (int*)flag, st);
}

extern "C" int vh_evp_workspace_bytes(const vh_domain* d, size_t* bytes) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(bytes != nullptr, VH_EINVAL, "null bytes");
  *bytes = d->dtype == VH_F64 ? vh::evp_core_ws_bytes<double>(g) : vh::evp_core_ws_bytes<float>(g);
  return VH_OK;
}

extern "C" int vh_evp_tile_count(void) { return vh::kEvpNumTiles - 1; }

extern "C" int vh_evp_tile_shape(int32_t tile, int32_t* ri, int32_t* rj, int32_t* k) {
  VH_REQUIRE(tile >= 1 && tile < vh::kEvpNumTiles && ri && rj && k, VH_EINVAL, "tile %d", tile);
  *ri = vh::kEvpTiles[tile].ri;
  *rj = vh::kEvpTiles[tile].rj;
  *k = vh::kEvpTiles[tile].k;
  return VH_OK;
}

extern "C" int vh_evp_solver(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                             int32_t n_substeps, void* uIce, void* vIce, void* sigma1, void* sigma2,
                             void* sigma12, void* workspace, size_t workspace_bytes, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(s != nullptr && f != nullptr, VH_EINVAL, "null settings or fields");
  VH_REQUIRE(n_substeps >= 0, VH_EINVAL, "n_substeps=%d", n_substeps);
  VH_REQUIRE(d->tile >= 0 && d->tile < vh::kEvpNumTiles, VH_EINVAL, "tile %d", d->tile);
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64)
    return vh::run_evp<double>(g, s, f, n_substeps, uIce, vIce, sigma1, sigma2, sigma12, workspace,
                               workspace_bytes, d->tile, d->dtype, st);
  return vh::run_evp<float>(g, s, f, n_substeps, uIce, vIce, sigma1, sigma2, sigma12, workspace,
                            workspace_bytes, d->tile, d->dtype, st);
}

extern "C" int vh_evp_solver_noalloc(const vh_domain* d, const vh_settings* s,
                                     const vh_fields* f, int32_t n_substeps, void* uIce,
                                     void* vIce, void* sigma1, void* sigma2, void* sigma12,
                                     void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(s != nullptr && f != nullptr, VH_EINVAL, "null settings or fields");
  VH_REQUIRE(n_substeps >= 0, VH_EINVAL, "n_substeps=%d", n_substeps);
  VH_REQUIRE(d->tile >= 0 && d->tile < vh::kEvpNumTiles, VH_EINVAL, "tile %d", d->tile);
  // the full workspace of vh_evp_solver (packed constants included), so both
  // entry points run the same kernels on the same launch plan
  const size_t need = d->dtype == VH_F64 ? vh::evp_core_ws_bytes<double>(g)
                                         : vh::evp_core_ws_bytes<float>(g);
  VH_REQUIRE(need <= vh::kNoAllocScratchBytes, VH_EWORKSPACE,
             "vh_evp_solver_noalloc: a %d x %d %s domain needs %zu bytes of scratch, beyond its "
             "maximum of %zu (VH_EVP_NOALLOC_MAX_BYTES); use vh_evp_solver with a workspace",
             d->nx, d->ny, d->dtype == VH_F64 ? "FP64" : "FP32", need,
             vh::kNoAllocScratchBytes);
  void* ws = nullptr;
  rc = vh::scratch_for_current_device(&ws);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64) {
    const double* in[5] = {(const double*)f->uIce, (const double*)f->vIce,
                           (const double*)f->sigma1, (const double*)f->sigma2,
                           (const double*)f->sigma12};
    double* const out[5] = {(double*)uIce, (double*)vIce, (double*)sigma1, (double*)sigma2,
                            (double*)sigma12};
    return vh::evp_run_core<double>(g, s, f, in, out, n_substeps, ws, need, d->tile, d->dtype, st,
                                    nullptr, true);
  }
  const float* in[5] = {(const float*)f->uIce, (const float*)f->vIce, (const float*)f->sigma1,
                        (const float*)f->sigma2, (const float*)f->sigma12};
  float* const out[5] = {(float*)uIce, (float*)vIce, (float*)sigma1, (float*)sigma2,
                         (float*)sigma12};
  return vh::evp_run_core<float>(g, s, f, in, out, n_substeps, ws, need, d->tile, d->dtype, st,
                                 nullptr, true);
}

extern "C" int vh_evp_residuals_workspace_bytes(const vh_domain* d, int32_t n_substeps,
                                                size_t* bytes) {
  int rc = vh_evp_workspace_bytes(d, bytes);
  if (rc) return rc;
  VH_REQUIRE(n_substeps >= 1, VH_EINVAL, "n_substeps=%d", n_substeps);
  vh::Geom g;
  vh::make_geom(d, &g);
  *bytes += (size_t)n_substeps * vh::evp_residual_tiles(g) * 2 * sizeof(double);
  return VH_OK;
}

extern "C" int vh_evp_solver_residuals(const vh_domain* d, const vh_settings* s,
                                       const vh_fields* f, int32_t n_substeps, void* uIce,
                                       void* vIce, void* sigma1, void* sigma2, void* sigma12,
                                       double* resU, double* resSig, void* workspace,
                                       size_t workspace_bytes, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(s != nullptr && f != nullptr, VH_EINVAL, "null settings or fields");
  hipStream_t st = (hipStream_t)stream;
  void* const o[5] = {uIce, vIce, sigma1, sigma2, sigma12};
  if (d->dtype == VH_F64)
    return vh::run_evp_res<double>(g, s, f, n_substeps, o, resU, resSig, workspace,
                                   workspace_bytes, d->dtype, st);
  return vh::run_evp_res<float>(g, s, f, n_substeps, o, resU, resSig, workspace, workspace_bytes,
                                d->dtype, st);
}
