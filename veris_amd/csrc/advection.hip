// Flux-limited, dimensionally split C-grid advection of hIceMean, hSnowMean
// and Area (veris/advection.py:15-175), extensive fields only.
//
// Three launches for all three fields:
//   XF  zonal flux on u-faces i in [2, NX-1) (zero elsewhere), overlap-filled
//   YF  x-update of the field (recomputed where needed) and the meridional
//       flux of the updated field on v-faces j in [2, NY-1), overlap-filled
//   UP  y-update and land mask
// Built with -ffp-contract=off; expressions keep the reference's order.
#include "vh_common.h"

namespace vh {
namespace {

constexpr int BLK = 256;

template <typename T>
struct AdvArgs {
  const T* fin[3];   // hIceMean, hSnowMean, Area
  T* out[3];
  T* F[3];           // zonal fluxes (workspace)
  T* G[3];           // meridional fluxes (workspace)
  const T *u, *v, *dxG, *dyG, *MU, *MV, *mInU, *mInV, *mInC, *M, *rdxC, *rdyC, *rrA;
  T dtT, CrMax;
};

template <typename T>
__device__ inline T limited_ratio(T rjm, T rj, T rjp, T trans, T crmax) {
  T cr = trans > (T)0 ? rjm : rjp;
  cr = fabs(rj) * crmax > fabs(cr) ? cr / rj : (sgn(cr) * crmax) * sgn(rj);
  return fmax((T)0, fmax(fmin((T)1, (T)2 * cr), fmin((T)2, cr)));
}

template <typename T>
__device__ inline T face_flux(T trans, T f0, T fm, T cfl, T rj, T cr) {
  return (trans * (f0 + fm)) * (T)0.5 - ((fabs(trans) * (((T)1 - cr) + cfl * cr)) * rj) * (T)0.5;
}

#define AT(ii, jj) ((size_t)(ii) * g.NY + (jj))

template <typename T>
__global__ __launch_bounds__(BLK) void adv_xflux_kernel(Geom g, AdvArgs<T> a) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)g.NX * g.NY) return;
  const int i = (int)(idx / g.NY), j = (int)(idx - (size_t)i * g.NY);
  if (!is_computed(g, i, j)) return;
  const bool face = i >= 2 && i < g.NX - 1;
  T trans = 0, cfl = 0, mW0 = 0, mWp = 0, mWm = 0;
  if (face) {
    trans = a.u[idx] * (a.dyG[idx] * a.MU[idx]);
    cfl = fabs((a.u[idx] * a.dtT) * a.rdxC[idx]);
    mWp = a.MU[AT(i + 1, j)] * a.mInU[AT(i + 1, j)];
    mW0 = a.MU[idx] * a.mInU[idx];
    mWm = a.MU[AT(i - 1, j)] * a.mInU[AT(i - 1, j)];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    T F = 0;
    if (face) {
      const T* f = a.fin[q];
      const T fp = f[AT(i + 1, j)], f0 = f[idx], fm = f[AT(i - 1, j)], fmm = f[AT(i - 2, j)];
      const T rjp = (fp - f0) * mWp;
      const T rj = (f0 - fm) * mW0;
      const T rjm = (fm - fmm) * mWm;
      const T cr = limited_ratio(rjm, rj, rjp, trans, a.CrMax);
      F = face_flux(trans, f0, fm, cfl, rj, cr);
    }
    store_filled(g, a.F[q], i, j, F);
  }
}

// field after the zonal update at (i, j)
template <typename T>
__device__ inline T xupd(const Geom& g, const AdvArgs<T>& a, int q, int i, int j) {
  const int ip = i == g.NX - 1 ? 0 : i + 1;
  const size_t c = AT(i, j);
  return a.fin[q][c] - ((a.dtT * a.mInC[c]) * a.rrA[c]) * (a.F[q][AT(ip, j)] - a.F[q][c]);
}

template <typename T>
__global__ __launch_bounds__(BLK) void adv_yflux_kernel(Geom g, AdvArgs<T> a) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)g.NX * g.NY) return;
  const int i = (int)(idx / g.NY), j = (int)(idx - (size_t)i * g.NY);
  if (!is_computed(g, i, j)) return;
  const bool face = j >= 2 && j < g.NY - 1;
  T trans = 0, cfl = 0, mS0 = 0, mSp = 0, mSm = 0;
  if (face) {
    trans = a.v[idx] * (a.dxG[idx] * a.MV[idx]);
    cfl = fabs((a.v[idx] * a.dtT) * a.rdyC[idx]);
    mSp = a.MV[AT(i, j + 1)] * a.mInV[AT(i, j + 1)];
    mS0 = a.MV[idx] * a.mInV[idx];
    mSm = a.MV[AT(i, j - 1)] * a.mInV[AT(i, j - 1)];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    T G = 0;
    if (face) {
      const T fp = xupd(g, a, q, i, j + 1), f0 = xupd(g, a, q, i, j);
      const T fm = xupd(g, a, q, i, j - 1), fmm = xupd(g, a, q, i, j - 2);
      const T rjp = (fp - f0) * mSp;
      const T rj = (f0 - fm) * mS0;
      const T rjm = (fm - fmm) * mSm;
      const T cr = limited_ratio(rjm, rj, rjp, trans, a.CrMax);
      G = face_flux(trans, f0, fm, cfl, rj, cr);
    }
    store_filled(g, a.G[q], i, j, G);
  }
}

template <typename T>
__global__ __launch_bounds__(BLK) void adv_update_kernel(Geom g, AdvArgs<T> a) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)g.NX * g.NY) return;
  const int i = (int)(idx / g.NY), j = (int)(idx - (size_t)i * g.NY);
  const int jp = j == g.NY - 1 ? 0 : j + 1;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T f1 = xupd(g, a, q, i, j);
    const T f2 = f1 - ((a.dtT * a.mInC[idx]) * a.rrA[idx]) * (a.G[q][AT(i, jp)] - a.G[q][idx]);
    a.out[q][idx] = f2 * a.M[idx];
  }
}

#undef AT

template <typename T>
int advect(const Geom& g, const vh_settings* s, const vh_fields* f, void* h, void* hs, void* A,
           void* ws, size_t wsb, hipStream_t st) {
  VH_REQUIRE(s->extensiveFld != 0.0, VH_EUNSUPPORTED, "extensiveFld=False is not supported");
  const size_t fb = align256(field_elems(g) * sizeof(T));
  VH_REQUIRE(ws && wsb >= 6 * fb, VH_EWORKSPACE, "advection workspace: need %zu bytes", 6 * fb);
  AdvArgs<T> a;
  a.fin[0] = (const T*)f->hIceMean;
  a.fin[1] = (const T*)f->hSnowMean;
  a.fin[2] = (const T*)f->Area;
  a.out[0] = (T*)h;
  a.out[1] = (T*)hs;
  a.out[2] = (T*)A;
  for (int q = 0; q < 3; ++q) {
    a.F[q] = (T*)((char*)ws + q * fb);
    a.G[q] = (T*)((char*)ws + (3 + q) * fb);
  }
  a.u = (const T*)f->uIce;
  a.v = (const T*)f->vIce;
  a.dxG = (const T*)f->dxG;
  a.dyG = (const T*)f->dyG;
  a.MU = (const T*)f->iceMaskU;
  a.MV = (const T*)f->iceMaskV;
  a.mInU = (const T*)f->maskInU;
  a.mInV = (const T*)f->maskInV;
  a.mInC = (const T*)f->maskInC;
  a.M = (const T*)f->iceMask;
  a.rdxC = (const T*)f->recip_dxC;
  a.rdyC = (const T*)f->recip_dyC;
  a.rrA = (const T*)f->recip_rA;
  const void* req[] = {a.fin[0], a.fin[1], a.fin[2], a.out[0], a.out[1], a.out[2], a.u, a.v,
                       a.dxG, a.dyG, a.MU, a.MV, a.mInU, a.mInV, a.mInC, a.M, a.rdxC, a.rdyC,
                       a.rrA};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "advection: null pointer");
  for (int q = 0; q < 3; ++q)
    for (int r = 0; r < 3; ++r)
      VH_REQUIRE((const void*)a.out[q] != (const void*)a.fin[r], VH_EINVAL,
                 "advection: output aliases an input field");
  a.dtT = (T)s->deltatTherm;
  a.CrMax = (T)s->CrMax;
  const dim3 grid = cell_grid(g, BLK);
  hipLaunchKernelGGL(adv_xflux_kernel<T>, grid, dim3(BLK), 0, st, g, a);
  hipLaunchKernelGGL(adv_yflux_kernel<T>, grid, dim3(BLK), 0, st, g, a);
  hipLaunchKernelGGL(adv_update_kernel<T>, grid, dim3(BLK), 0, st, g, a);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

}  // namespace
}  // namespace vh

extern "C" int vh_advection_workspace_bytes(const vh_domain* d, size_t* bytes) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(bytes != nullptr, VH_EINVAL, "null bytes");
  *bytes = 6 * vh::align256(vh::field_elems(g) * vh::elem_bytes(d->dtype));
  return VH_OK;
}

extern "C" int vh_advection(const vh_domain* d, const vh_settings* s, const vh_fields* f, void* h,
                            void* hs, void* A, void* ws, size_t wsb, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(s != nullptr && f != nullptr, VH_EINVAL, "null settings or fields");
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64) return vh::advect<double>(g, s, f, h, hs, A, ws, wsb, st);
  return vh::advect<float>(g, s, f, h, hs, A, ws, wsb, st);
}
