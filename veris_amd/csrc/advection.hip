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

// ---------------------------------------------------------------- adjoint
//
// Reverse of the dimensionally split update (veris/advection.py:28-175) for
// the three fields: adjoints of the outputs -> adjoints of the input fields
// and of the ice velocities.  Every flux is a piecewise smooth function of
// the field at the four cells of its stencil and of the face velocity; its
// derivative follows the branch the forward evaluation takes: the upwind
// side picked by the sign of the transport, the ratio or the clipped
// (constant) slope ratio, and the superbee segment with slope 2 on (0, 1/2],
// 1 on (1, 2] and 0 elsewhere; d|x|/dx = sgn(x), 0 at 0.
//
// Passes (one thread per padded cell each):
//   AY  y-face adjoints: the flux adjoint from the output adjoints (with the
//       periodic-fill mirrors summed), times its derivatives -> the four
//       stencil contributions of the face, and the v adjoint
//   BY  adjoint of the x-updated field: output adjoint + the contributions
//       of the four y-faces whose stencil holds the cell
//   AX  x-face adjoints from the x-updated field's adjoint -> stencil
//       contributions and the u adjoint
//   BX  adjoint of the input field

template <typename T>
struct FluxGrad {
  T d0, dm, dp, dmm, dvel;
};

// derivatives of face_flux(trans, f0, fm, cfl, rj, superbee ratio) w.r.t.
// the field at the face's own cell (f0), the cells behind (fm, fmm) and
// ahead (fp) of it, and the face velocity (through trans and cfl, with
// d trans / d vel = dtr and d cfl / d vel = dcf)
template <typename T>
__device__ inline FluxGrad<T> face_flux_grad(T trans, T dtr, T cfl, T dcf, T fp, T f0, T fm,
                                             T fmm, T mWp, T mW0, T mWm, T crmax) {
  const T rjp = (fp - f0) * mWp;
  const T rj = (f0 - fm) * mW0;
  const T rjm = (fm - fmm) * mWm;
  const bool up = trans > (T)0;
  const T c0 = up ? rjm : rjp;
  const bool ratio = fabs(rj) * crmax > fabs(c0);
  const T cr = ratio ? c0 / rj : (sgn(c0) * crmax) * sgn(rj);
  T psi, dpsi;
  if (cr <= (T)0) {
    psi = (T)0;
    dpsi = (T)0;
  } else if (cr <= (T)0.5) {
    psi = (T)2 * cr;
    dpsi = (T)2;
  } else if (cr <= (T)1) {
    psi = (T)1;
    dpsi = (T)0;
  } else if (cr <= (T)2) {
    psi = cr;
    dpsi = (T)1;
  } else {
    psi = (T)2;
    dpsi = (T)0;
  }
  const T at = fabs(trans);
  const T W = ((T)1 - psi) + cfl * psi;
  const T gpsi = -(at * (cfl - (T)1) * rj) * (T)0.5;       // d flux / d psi
  const T gc0 = ratio ? gpsi * dpsi / rj : (T)0;           // via the other slope
  const T grj = -(at * W) * (T)0.5 + (ratio ? gpsi * dpsi * (-cr / rj) : (T)0);
  FluxGrad<T> d;
  d.d0 = trans * (T)0.5 + grj * mW0;
  d.dm = trans * (T)0.5 - grj * mW0;
  d.dp = (T)0;
  d.dmm = (T)0;
  if (up) {
    d.dm = d.dm + gc0 * mWm;
    d.dmm = -gc0 * mWm;
  } else {
    d.dp = gc0 * mWp;
    d.d0 = d.d0 - gc0 * mWp;
  }
  d.dvel = dtr * (((f0 + fm) * (T)0.5) - ((sgn(trans) * W) * rj) * (T)0.5) +
           dcf * (-((at * psi) * rj) * (T)0.5);
  return d;
}

template <typename T>
struct AdvAdjArgs {
  AdvArgs<T> a;          // forward arguments (F = recomputed zonal fluxes)
  const T* lout[3];      // adjoints of the outputs (NULL = 0)
  T* lf1[3];             // adjoints of the x-updated fields (workspace)
  T* c0[3];              // per-face stencil contributions (workspace)
  T* cm[3];
  T* cp[3];
  T* cmm[3];
  T *lin[3], *lu, *lv;   // results
};

// sum of a face adjoint over a computed cell and the overlap cells that the
// periodic fill copies its value to (the positions store_filled writes)
template <typename T, typename F>
__device__ inline T gather_mirrors(const Geom& g, int i, int j, F&& val) {
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
  T acc = (T)0;
  for (int p = 0; p < 2; ++p) {
    if (ix[p] < 0) continue;
    for (int r = 0; r < 2; ++r) {
      if (jy[r] < 0) continue;
      acc = acc + val(ix[p], jy[r]);
    }
  }
  return acc;
}

template <typename T>
__device__ inline T upd_coef(const Geom& g, const AdvArgs<T>& a, size_t c) {
  return (a.dtT * a.mInC[c]) * a.rrA[c];
}

template <typename T>
__global__ __launch_bounds__(BLK) void adv_adj_y_kernel(Geom g, AdvAdjArgs<T> b) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)g.NX * g.NY) return;
  const int i = (int)(idx / g.NY), j = (int)(idx - (size_t)i * g.NY);
  const AdvArgs<T>& a = b.a;
  const bool face = is_computed(g, i, j) && j >= 2 && j < g.NY - 1;
  T trans = 0, cfl = 0, dtr = 0, dcf = 0, mS0 = 0, mSp = 0, mSm = 0;
  if (face) {
    const T v = a.v[idx];
    trans = v * (a.dxG[idx] * a.MV[idx]);
    const T x = (v * a.dtT) * a.rdyC[idx];
    cfl = fabs(x);
    dtr = a.dxG[idx] * a.MV[idx];
    dcf = sgn(x) * (a.dtT * a.rdyC[idx]);
    mSp = a.MV[AT(i, j + 1)] * a.mInV[AT(i, j + 1)];
    mS0 = a.MV[idx] * a.mInV[idx];
    mSm = a.MV[AT(i, j - 1)] * a.mInV[AT(i, j - 1)];
  }
  T lv = (T)0;
  for (int q = 0; q < 3; ++q) {
    T w0 = 0, wm = 0, wp = 0, wmm = 0;
    if (face && b.lout[q]) {
      const T* lo = b.lout[q];
      // adjoint of G at (ii, jj): +c lambda_f2 there, -c lambda_f2 at (ii, jj-1)
      const T lG = gather_mirrors<T>(g, i, j, [&](int ii, int jj) {
        const int jm = jj == 0 ? g.NY - 1 : jj - 1;
        const size_t p = AT(ii, jj), pm = AT(ii, jm);
        return upd_coef(g, a, p) * (lo[p] * a.M[p]) - upd_coef(g, a, pm) * (lo[pm] * a.M[pm]);
      });
      const FluxGrad<T> d = face_flux_grad(trans, dtr, cfl, dcf, xupd(g, a, q, i, j + 1),
                                           xupd(g, a, q, i, j), xupd(g, a, q, i, j - 1),
                                           xupd(g, a, q, i, j - 2), mSp, mS0, mSm, a.CrMax);
      w0 = lG * d.d0;
      wm = lG * d.dm;
      wp = lG * d.dp;
      wmm = lG * d.dmm;
      lv = lv + lG * d.dvel;
    }
    b.c0[q][idx] = w0;
    b.cm[q][idx] = wm;
    b.cp[q][idx] = wp;
    b.cmm[q][idx] = wmm;
  }
  b.lv[idx] = lv;
}

// adjoint of the x-updated field: the output's (times the land mask) plus
// the stencil contributions of the y-faces j, j+1, j-1, j+2
template <typename T>
__global__ __launch_bounds__(BLK) void adv_adj_ygather_kernel(Geom g, AdvAdjArgs<T> b) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)g.NX * g.NY) return;
  const int i = (int)(idx / g.NY), j = (int)(idx - (size_t)i * g.NY);
  for (int q = 0; q < 3; ++q) {
    T l = b.lout[q] ? b.lout[q][idx] * b.a.M[idx] : (T)0;
    l = l + b.c0[q][idx];
    if (j + 1 < g.NY) l = l + b.cm[q][AT(i, j + 1)];
    if (j >= 1) l = l + b.cp[q][AT(i, j - 1)];
    if (j + 2 < g.NY) l = l + b.cmm[q][AT(i, j + 2)];
    b.lf1[q][idx] = l;
  }
}

template <typename T>
__global__ __launch_bounds__(BLK) void adv_adj_x_kernel(Geom g, AdvAdjArgs<T> b) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)g.NX * g.NY) return;
  const int i = (int)(idx / g.NY), j = (int)(idx - (size_t)i * g.NY);
  const AdvArgs<T>& a = b.a;
  const bool face = is_computed(g, i, j) && i >= 2 && i < g.NX - 1;
  T trans = 0, cfl = 0, dtr = 0, dcf = 0, mW0 = 0, mWp = 0, mWm = 0;
  if (face) {
    const T u = a.u[idx];
    trans = u * (a.dyG[idx] * a.MU[idx]);
    const T x = (u * a.dtT) * a.rdxC[idx];
    cfl = fabs(x);
    dtr = a.dyG[idx] * a.MU[idx];
    dcf = sgn(x) * (a.dtT * a.rdxC[idx]);
    mWp = a.MU[AT(i + 1, j)] * a.mInU[AT(i + 1, j)];
    mW0 = a.MU[idx] * a.mInU[idx];
    mWm = a.MU[AT(i - 1, j)] * a.mInU[AT(i - 1, j)];
  }
  T lu = (T)0;
  for (int q = 0; q < 3; ++q) {
    T w0 = 0, wm = 0, wp = 0, wmm = 0;
    if (face) {
      const T* l1 = b.lf1[q];
      // adjoint of F at (ii, jj): +c lambda_f1 there, -c lambda_f1 at (ii-1, jj)
      const T lF = gather_mirrors<T>(g, i, j, [&](int ii, int jj) {
        const int im = ii == 0 ? g.NX - 1 : ii - 1;
        const size_t p = AT(ii, jj), pm = AT(im, jj);
        return upd_coef(g, a, p) * l1[p] - upd_coef(g, a, pm) * l1[pm];
      });
      const T* f = a.fin[q];
      const FluxGrad<T> d =
          face_flux_grad(trans, dtr, cfl, dcf, f[AT(i + 1, j)], f[idx], f[AT(i - 1, j)],
                         f[AT(i - 2, j)], mWp, mW0, mWm, a.CrMax);
      w0 = lF * d.d0;
      wm = lF * d.dm;
      wp = lF * d.dp;
      wmm = lF * d.dmm;
      lu = lu + lF * d.dvel;
    }
    b.c0[q][idx] = w0;
    b.cm[q][idx] = wm;
    b.cp[q][idx] = wp;
    b.cmm[q][idx] = wmm;
  }
  b.lu[idx] = lu;
}

template <typename T>
__global__ __launch_bounds__(BLK) void adv_adj_xgather_kernel(Geom g, AdvAdjArgs<T> b) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)g.NX * g.NY) return;
  const int i = (int)(idx / g.NY), j = (int)(idx - (size_t)i * g.NY);
  for (int q = 0; q < 3; ++q) {
    T l = b.lf1[q][idx] + b.c0[q][idx];
    if (i + 1 < g.NX) l = l + b.cm[q][AT(i + 1, j)];
    if (i >= 1) l = l + b.cp[q][AT(i - 1, j)];
    if (i + 2 < g.NX) l = l + b.cmm[q][AT(i + 2, j)];
    b.lin[q][idx] = l;
  }
}

#undef AT

template <typename T>
int adv_args(const Geom& g, const vh_settings* s, const vh_fields* f, AdvArgs<T>* ap) {
  VH_REQUIRE(s->extensiveFld != 0.0, VH_EUNSUPPORTED, "extensiveFld=False is not supported");
  AdvArgs<T>& a = *ap;
  a.fin[0] = (const T*)f->hIceMean;
  a.fin[1] = (const T*)f->hSnowMean;
  a.fin[2] = (const T*)f->Area;
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
  const void* req[] = {a.fin[0], a.fin[1], a.fin[2], a.u, a.v, a.dxG, a.dyG, a.MU, a.MV,
                       a.mInU, a.mInV, a.mInC, a.M, a.rdxC, a.rdyC, a.rrA};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "advection: null pointer");
  a.dtT = (T)s->deltatTherm;
  a.CrMax = (T)s->CrMax;
  return VH_OK;
}

// Workspace fields of the adjoint: 3 zonal fluxes, 3 adjoints of the
// x-updated fields, 4 x 3 stencil contributions.
constexpr int kAdvAdjFields = 18;

template <typename T>
int advect_adjoint(const Geom& g, const vh_settings* s, const vh_fields* f,
                   const void* const lout[3], void* const lin[3], void* lu, void* lv, void* ws,
                   size_t wsb, hipStream_t st) {
  const size_t fb = align256(field_elems(g) * sizeof(T));
  VH_REQUIRE(ws && wsb >= kAdvAdjFields * fb, VH_EWORKSPACE,
             "advection adjoint workspace: need %zu bytes", kAdvAdjFields * fb);
  AdvAdjArgs<T> b{};
  int rc = adv_args<T>(g, s, f, &b.a);
  if (rc) return rc;
  VH_REQUIRE(lu && lv && lin[0] && lin[1] && lin[2], VH_EINVAL, "advection adjoint: null output");
  char* w = (char*)ws;
  for (int q = 0; q < 3; ++q) {
    b.a.F[q] = (T*)(w + q * fb);
    b.a.G[q] = nullptr;
    b.a.out[q] = nullptr;
    b.lf1[q] = (T*)(w + (3 + q) * fb);
    b.c0[q] = (T*)(w + (6 + q) * fb);
    b.cm[q] = (T*)(w + (9 + q) * fb);
    b.cp[q] = (T*)(w + (12 + q) * fb);
    b.cmm[q] = (T*)(w + (15 + q) * fb);
    b.lout[q] = (const T*)lout[q];
    b.lin[q] = (T*)lin[q];
  }
  b.lu = (T*)lu;
  b.lv = (T*)lv;
  const void* ins[] = {b.a.fin[0], b.a.fin[1], b.a.fin[2], b.a.u, b.a.v, lout[0], lout[1], lout[2]};
  void* outs[] = {lin[0], lin[1], lin[2], lu, lv};
  for (int q = 0; q < 5; ++q) {
    for (const void* p : ins) VH_REQUIRE(outs[q] != p, VH_EINVAL, "advection adjoint: output aliases an input");
    for (int r = 0; r < q; ++r) VH_REQUIRE(outs[q] != outs[r], VH_EINVAL, "advection adjoint: outputs alias");
  }
  const dim3 grid = cell_grid(g, BLK);
  hipLaunchKernelGGL(adv_xflux_kernel<T>, grid, dim3(BLK), 0, st, g, b.a);  // the forward zonal fluxes
  hipLaunchKernelGGL(adv_adj_y_kernel<T>, grid, dim3(BLK), 0, st, g, b);
  hipLaunchKernelGGL(adv_adj_ygather_kernel<T>, grid, dim3(BLK), 0, st, g, b);
  hipLaunchKernelGGL(adv_adj_x_kernel<T>, grid, dim3(BLK), 0, st, g, b);
  hipLaunchKernelGGL(adv_adj_xgather_kernel<T>, grid, dim3(BLK), 0, st, g, b);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

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

extern "C" int vh_advection_adjoint_workspace_bytes(const vh_domain* d, size_t* bytes) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(bytes != nullptr, VH_EINVAL, "null bytes");
  *bytes = vh::kAdvAdjFields * vh::align256(vh::field_elems(g) * vh::elem_bytes(d->dtype));
  return VH_OK;
}

extern "C" int vh_advection_adjoint(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                                    const void* adj_hIceMean, const void* adj_hSnowMean,
                                    const void* adj_Area, void* adj_hIceMean_in,
                                    void* adj_hSnowMean_in, void* adj_Area_in, void* adj_uIce,
                                    void* adj_vIce, void* ws, size_t wsb, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(s != nullptr && f != nullptr, VH_EINVAL, "null settings or fields");
  const void* lout[3] = {adj_hIceMean, adj_hSnowMean, adj_Area};
  void* const lin[3] = {adj_hIceMean_in, adj_hSnowMean_in, adj_Area_in};
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64)
    return vh::advect_adjoint<double>(g, s, f, lout, lin, adj_uIce, adj_vIce, ws, wsb, st);
  return vh::advect_adjoint<float>(g, s, f, lout, lin, adj_uIce, adj_vIce, ws, wsb, st);
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
