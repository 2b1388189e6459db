// Once-per-step stencil and pointwise kernels of the sea-ice step: pre-EVP
// forcing (veris/area_mass.py, veris/dynamics_routines.py:5-18,
// veris/dynsolver.py:7-84), the free-drift solver
// (veris/freedrift_solver.py:5-62), ocean stress (veris/ocean_stress.py:8-40),
// clean-up and ridging (veris/clean_up.py:5-47), the overlap fill
// (veris/fill_overlap.py:6-18) and an FP64 global sum (veris/global_sum.py).
//
// One thread per padded cell, neighbours wrapped like numpy.roll; these
// kernels run once per model step and move < 1% of the step's bytes, so they
// are plain coalesced sweeps.  Built with -ffp-contract=off: every expression
// keeps the reference's association order.
#include <vector>

#include "vh_common.h"

namespace vh {
namespace {

constexpr int BLK = 256;

#define CELL_PROLOGUE                                                   \
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;     \
  if (idx >= (size_t)g.NX * g.NY) return;                               \
  const int i = (int)(idx / g.NY);                                      \
  const int j = (int)(idx - (size_t)i * g.NY);                          \
  const int im = (i == 0 ? g.NX - 1 : i - 1), ip = (i == g.NX - 1 ? 0 : i + 1); \
  const int jm = (j == 0 ? g.NY - 1 : j - 1), jp = (j == g.NY - 1 ? 0 : j + 1); \
  (void)im; (void)ip; (void)jm; (void)jp;
#define AT(ii, jj) ((size_t)(ii) * g.NY + (jj))

template <typename T>
const T* C_(const void* p) { return static_cast<const T*>(p); }

// ---------------------------------------------------------------- pre-EVP

template <typename T>
__global__ __launch_bounds__(BLK) void seaice_mass_kernel(Geom g, const T* h, const T* hs, T rhoI,
                                                          T rhoS, T* mC, T* mU, T* mV) {
  CELL_PROLOGUE
  const T c = rhoI * h[idx] + rhoS * hs[idx];
  const T cim = rhoI * h[AT(im, j)] + rhoS * hs[AT(im, j)];
  const T cjm = rhoI * h[AT(i, jm)] + rhoS * hs[AT(i, jm)];
  mC[idx] = c;
  mU[idx] = (T)0.5 * (c + cim);
  mV[idx] = (T)0.5 * (c + cjm);
}

template <typename T>
__global__ __launch_bounds__(BLK) void area_ws_kernel(Geom g, const T* A, T* AW, T* AS) {
  CELL_PROLOGUE
  AW[idx] = (T)0.5 * (A[idx] + A[AT(im, j)]);
  AS[idx] = (T)0.5 * (A[idx] + A[AT(i, jm)]);
}

template <typename T>
__global__ __launch_bounds__(BLK) void strength_kernel(Geom g, const T* h, const T* A, const T* M,
                                                       const T* psf, T ps, T cStar, T* P) {
  CELL_PROLOGUE
  const T pstar = psf ? psf[idx] : ps;
  P[idx] = ((pstar * h[idx]) * exp((-cStar) * ((T)1 - A[idx]))) * M[idx];
}

template <typename T>
struct WindArgs {
  const T *u, *v, *uW, *vW, *fC, *MU, *MV, *AW, *AS, *mU, *mV, *rdxC, *rdyC, *ssh, *sp, *load;
  T sinWin, cosWin, wmin, wmin2, caN, caS, rhoAir, grav, rrho, loadFac;
  int relWind, realFW;
};

template <typename T>
__device__ inline void tau_c(const Geom& g, const WindArgs<T>& w, int i, int j, T* tx, T* ty) {
  const int ip = i == g.NX - 1 ? 0 : i + 1;
  const int jp = j == g.NY - 1 ? 0 : j + 1;
  const size_t c = AT(i, j);
  T urel = w.uW[c], vrel = w.vW[c];
  if (w.relWind) {
    urel = urel - (T)0.5 * (w.u[c] + w.u[AT(ip, j)]);
    vrel = vrel - (T)0.5 * (w.v[c] + w.v[AT(i, jp)]);
  }
  const T wsq = urel * urel + vrel * vrel;
  const T ws = wsq < w.wmin2 ? w.wmin : sqrt(wsq);
  const T f = w.fC[c];
  const T cda = ((f < (T)0 ? w.caS : w.caN) * w.rhoAir) * ws;
  const T sg = sgn(f);
  *tx = cda * (w.cosWin * urel - (sg * w.sinWin) * vrel);
  *ty = cda * (w.cosWin * vrel + (sg * w.sinWin) * urel);
}

template <typename T>
__device__ inline T phi_at(const WindArgs<T>& w, size_t c) {
  T phi = w.grav * w.ssh[c];
  if (w.realFW) phi = phi + (w.sp[c] + (w.load[c] * w.grav) * w.loadFac) * w.rrho;
  else phi = phi + w.sp[c] * w.rrho;
  return phi;
}

template <typename T>
__global__ __launch_bounds__(BLK) void wind_forcing_kernel(Geom g, WindArgs<T> w, T* WFX, T* WFY) {
  CELL_PROLOGUE
  T tx, ty, txim, tyim, txjm, tyjm;
  tau_c(g, w, i, j, &tx, &ty);
  tau_c(g, w, im, j, &txim, &tyim);
  tau_c(g, w, i, jm, &txjm, &tyjm);
  const T tauX = ((T)0.5 * (tx + txim)) * w.MU[idx];
  const T tauY = ((T)0.5 * (ty + tyjm)) * w.MV[idx];
  const T phi = phi_at(w, idx);
  T fx = tauX * w.AW[idx];
  T fy = tauY * w.AS[idx];
  fx = fx - (w.mU[idx] * w.rdxC[idx]) * (phi - phi_at(w, AT(im, j)));
  fy = fy - (w.mV[idx] * w.rdyC[idx]) * (phi - phi_at(w, AT(i, jm)));
  WFX[idx] = fx;
  WFY[idx] = fy;
}

// ---------------------------------------------------------------- ocean stress

template <typename T>
struct DragArgs {
  const T *u, *v, *uOc, *vOc, *fC, *mInU, *mInV, *M;
  T cwN, cwS, rhoSea, cmin, cmin2;
};

template <typename T>
__device__ inline T cdrag_c(const Geom& g, const DragArgs<T>& a, int i, int j) {
  const int ip = i == g.NX - 1 ? 0 : i + 1;
  const int jp = j == g.NY - 1 ? 0 : j + 1;
  const size_t c = AT(i, j), cip = AT(ip, j), cjp = AT(i, jp);
  const T D = (a.fC[c] < (T)0 ? a.cwS : a.cwN) * a.rhoSea;
  const T du = (a.u[c] - a.uOc[c]) * a.mInU[c];
  const T duip = (a.u[cip] - a.uOc[cip]) * a.mInU[cip];
  const T dv = (a.v[c] - a.vOc[c]) * a.mInV[c];
  const T dvjp = (a.v[cjp] - a.vOc[cjp]) * a.mInV[cjp];
  const T su = du + duip, sv = dv + dvjp;
  const T w = (T)0.25 * (su * su + sv * sv);
  const T cd = ((D * D) * w > a.cmin2) ? D * sqrt(w) : a.cmin;
  return cd * a.M[c];
}

template <typename T>
__global__ __launch_bounds__(BLK) void ocean_stress_kernel(Geom g, DragArgs<T> a, T sinW, T cosW,
                                                           T* tu, T* tv) {
  CELL_PROLOGUE
  if (!is_computed(g, i, j)) return;
  const T cd = cdrag_c(g, a, i, j);
  const T cdim = cdrag_c(g, a, im, j);
  const T cdjm = cdrag_c(g, a, i, jm);
  const T du = a.u[idx] - a.uOc[idx];
  const T dv = a.v[idx] - a.vOc[idx];
  const T sg = sgn(a.fC[idx]);
  // dvAtC at (i, j) and (i, j-1); duAtC at (i, j) and (i-1, j)
  const T dvC = (T)0.5 * (dv + (a.v[AT(i, jp)] - a.vOc[AT(i, jp)]));
  const T dvCjm = (T)0.5 * ((a.v[AT(i, jm)] - a.vOc[AT(i, jm)]) + dv);
  const T duC = (T)0.5 * (du + (a.u[AT(ip, j)] - a.uOc[AT(ip, j)]));
  const T duCim = (T)0.5 * ((a.u[AT(im, j)] - a.uOc[AT(im, j)]) + du);
  const T su = (((T)0.5 * (cd + cdim)) * cosW) * du -
               ((sg * sinW) * (T)0.5) * (cd * dvC + cdjm * dvCjm);
  const T sv = (((T)0.5 * (cd + cdjm)) * cosW) * dv +
               ((sg * sinW) * (T)0.5) * (cd * duC + cdim * duCim);
  store_filled(g, tu, i, j, su);
  store_filled(g, tv, i, j, sv);
}

// ---------------------------------------------------------------- free drift

template <typename T>
struct FreeDriftArgs {
  const T *WFX, *WFY, *h, *fC, *uOc, *vOc, *MU, *MV;
  T rhoIce, cwN, cwS, rhoSea;
};

// Steady free-drift velocity at c-point (i, j) (veris/freedrift_solver.py:12-52)
template <typename T>
__device__ inline void free_drift_c(const Geom& g, const FreeDriftArgs<T>& a, int i, int j, T* uc,
                                    T* vc) {
  const int ip = i == g.NX - 1 ? 0 : i + 1;
  const int jp = j == g.NY - 1 ? 0 : j + 1;
  const size_t c = AT(i, j), cip = AT(ip, j), cjp = AT(i, jp);
  const T tx = (T)0.5 * (a.WFX[c] + a.WFX[cip]);
  const T ty = (T)0.5 * (a.WFY[c] + a.WFY[cjp]);
  const T f = a.fC[c];
  const T mf = (a.rhoIce * a.h[c]) * f;
  const T uoc = (T)0.5 * (a.uOc[c] + a.uOc[cip]);
  const T voc = (T)0.5 * (a.vOc[c] + a.vOc[cjp]);
  const T rx = (-tx) - mf * voc;
  const T ry = (-ty) + mf * uoc;
  const T r2 = rx * rx + ry * ry;
  const T rN = r2 > (T)0 ? sqrt(rx * rx + ry * ry) : (T)0;
  const T rA = r2 > (T)0 ? atan2(ry, rx) : (T)0;
  const T t1 = (T)1 / ((f < (T)0 ? a.cwS : a.cwN) * a.rhoSea);
  const T t2 = (t1 * t1) * (mf * mf);
  const T t3 = (t1 * t1) * (rN * rN);
  const T t4 = t2 * t2 + (T)4 * t3;
  const T sN = t3 > (T)0 ? sqrt((T)0.5 * (sqrt(t4) - t2)) : (T)0;
  const T q1 = (T)1 / t1;
  const T q2 = q1 * (sN * sN);
  const T q3 = mf * sN;
  const T q4 = q2 * q2 + q3 * q3;
  const T ang = q4 > (T)0 ? rA - atan2(q3, q2) : (T)0;
  *uc = uoc - sN * cos(ang);
  *vc = voc - sN * sin(ang);
}

template <typename T>
__global__ __launch_bounds__(BLK) void free_drift_kernel(Geom g, FreeDriftArgs<T> a, T* u, T* v) {
  CELL_PROLOGUE
  T uc, vc, ucim, vcim, ucjm, vcjm;
  free_drift_c(g, a, i, j, &uc, &vc);
  free_drift_c(g, a, im, j, &ucim, &vcim);
  free_drift_c(g, a, i, jm, &ucjm, &vcjm);
  u[idx] = ((T)0.5 * (ucim + uc)) * a.MU[idx];
  v[idx] = ((T)0.5 * (vcjm + vc)) * a.MV[idx];
}

// ---------------------------------------------------------------- clean-up, ridging

template <typename T>
__global__ __launch_bounds__(BLK) void clean_up_kernel(Geom g, const T* hin, const T* hsin,
                                                       const T* Ain, const T* Tin, T hmin, T Amin,
                                                       T c2k, T* h, T* hs, T* A, T* Ts, T* osh,
                                                       T* oshs) {
  CELL_PROLOGUE
  const T h0 = hin[idx], hs0 = hsin[idx];
  osh[idx] = fmax(-h0, (T)0);
  oshs[idx] = fmax(-hs0, (T)0);
  T hh = fmax(h0, (T)0), hhs = fmax(hs0, (T)0), AA = fmax(Ain[idx], (T)0);
  const bool thin = hh <= hmin;
  hh = hh * (T)(!thin);
  hhs = hhs * (T)(!thin);
  Ts[idx] = thin ? c2k : Tin[idx];
  if (hh == (T)0 && hhs == (T)0) AA = (T)0;
  if (hh > (T)0 || hhs > (T)0) AA = fmax(AA, Amin);
  h[idx] = hh;
  hs[idx] = hhs;
  A[idx] = AA;
}

template <typename T>
__global__ __launch_bounds__(BLK) void ridging_kernel(Geom g, const T* Ain, T* A) {
  CELL_PROLOGUE
  A[idx] = fmin(Ain[idx], (T)1);
}

// ---------------------------------------------------------------- fill, sum

template <typename T>
__global__ __launch_bounds__(BLK) void fill_kernel(Geom g, T* a) {
  // one thread per overlap cell of the padded array
  CELL_PROLOGUE
  if (is_computed(g, i, j)) return;
  int si = i, sj = j;
  if (g.fill_x) {
    if (i < g.ol) si = i + g.nx;
    else if (i >= g.nx + g.ol) si = i - g.nx;
  }
  if (g.fill_y) {
    if (j < g.ol) sj = j + g.ny;
    else if (j >= g.ny + g.ol) sj = j - g.ny;
  }
  a[idx] = a[AT(si, sj)];
}

// several fields of one grid in one launch (the step's closing fills):
// every overlap cell reads only interior cells, so one pass per field equals
// the x-then-y fill
constexpr int kFillMax = 8;
template <typename T>
struct FillList {
  T* f[kFillMax];
  int n;
};

template <typename T>
__global__ __launch_bounds__(BLK) void fill_many_kernel(Geom g, FillList<T> fl) {
  CELL_PROLOGUE
  if (is_computed(g, i, j)) return;
  int si = i, sj = j;
  if (g.fill_x) {
    if (i < g.ol) si = i + g.nx;
    else if (i >= g.nx + g.ol) si = i - g.nx;
  }
  if (g.fill_y) {
    if (j < g.ol) sj = j + g.ny;
    else if (j >= g.ny + g.ol) sj = j - g.ny;
  }
  const size_t src = AT(si, sj);
  for (int q = 0; q < fl.n; ++q) fl.f[q][idx] = fl.f[q][src];
}

template <typename T>
__global__ __launch_bounds__(BLK) void sum_partial_kernel(Geom g, const T* x, const T* w,
                                                          double* part) {
  // interior cells only; fixed-order tree per block, so results are reproducible
  __shared__ double red[BLK / 64];
  const size_t n = (size_t)g.nx * g.ny;
  double acc = 0.0;
  for (size_t k = (size_t)blockIdx.x * BLK + threadIdx.x; k < n; k += (size_t)gridDim.x * BLK) {
    const int ii = (int)(k / g.ny) + g.ol, jj = (int)(k % g.ny) + g.ol;
    const size_t c = AT(ii, jj);
    double v = (double)x[c];
    if (w) v = v * (double)w[c];
    acc += v;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < BLK / 64; ++q) s += red[q];
    part[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(BLK) void sum_final_kernel(const double* part, int np, double* out) {
  __shared__ double red[BLK / 64];
  double acc = 0.0;
  for (int k = threadIdx.x; k < np; k += BLK) acc += part[k];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < BLK / 64; ++q) s += red[q];
    *out = s;
  }
}

constexpr int SUM_BLOCKS = 512;

// ---------------------------------------------------------------- host side

template <typename T>
int seaice_mass(const Geom& g, const vh_settings* s, const vh_fields* f, void* C, void* U, void* V,
                hipStream_t st) {
  VH_REQUIRE(f->hIceMean && f->hSnowMean && C && U && V, VH_EINVAL, "seaice_mass: null pointer");
  hipLaunchKernelGGL(seaice_mass_kernel<T>, cell_grid(g, BLK), dim3(BLK), 0, st, g,
                     C_<T>(f->hIceMean), C_<T>(f->hSnowMean), (T)s->rhoIce, (T)s->rhoSnow, (T*)C,
                     (T*)U, (T*)V);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int area_ws(const Geom& g, const vh_fields* f, void* W, void* S, hipStream_t st) {
  VH_REQUIRE(f->Area && W && S, VH_EINVAL, "area_ws: null pointer");
  hipLaunchKernelGGL(area_ws_kernel<T>, cell_grid(g, BLK), dim3(BLK), 0, st, g, C_<T>(f->Area),
                     (T*)W, (T*)S);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int strength(const Geom& g, const vh_settings* s, const vh_fields* f, const void* psf, void* P,
             hipStream_t st) {
  VH_REQUIRE(f->hIceMean && f->Area && f->iceMask && P, VH_EINVAL, "seaice_strength: null pointer");
  hipLaunchKernelGGL(strength_kernel<T>, cell_grid(g, BLK), dim3(BLK), 0, st, g,
                     C_<T>(f->hIceMean), C_<T>(f->Area), C_<T>(f->iceMask), C_<T>(psf),
                     (T)s->pStar, (T)s->cStar, (T*)P);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int wind_forcing(const Geom& g, const vh_settings* s, const vh_fields* f, void* X, void* Y,
                 hipStream_t st) {
  WindArgs<T> w;
  w.u = C_<T>(f->uIce);
  w.v = C_<T>(f->vIce);
  w.uW = C_<T>(f->uWind);
  w.vW = C_<T>(f->vWind);
  w.fC = C_<T>(f->fCori);
  w.MU = C_<T>(f->iceMaskU);
  w.MV = C_<T>(f->iceMaskV);
  w.AW = C_<T>(f->AreaW);
  w.AS = C_<T>(f->AreaS);
  w.mU = C_<T>(f->SeaIceMassU);
  w.mV = C_<T>(f->SeaIceMassV);
  w.rdxC = C_<T>(f->recip_dxC);
  w.rdyC = C_<T>(f->recip_dyC);
  w.ssh = C_<T>(f->ssh_an);
  w.sp = C_<T>(f->surfPress);
  w.load = C_<T>(f->SeaIceLoad);
  w.relWind = s->useRelativeWind != 0.0;
  w.realFW = s->useRealFreshWaterFlux != 0.0;
  const void* req[] = {w.uW, w.vW, w.fC, w.MU, w.MV, w.AW, w.AS, w.mU, w.mV, w.rdxC, w.rdyC, w.ssh,
                       w.sp, X, Y};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "wind_forcing_xy: null pointer");
  if (w.relWind) VH_REQUIRE(w.u && w.v, VH_EINVAL, "wind_forcing_xy: null uIce/vIce");
  if (w.realFW) VH_REQUIRE(w.load, VH_EINVAL, "wind_forcing_xy: null SeaIceLoad");
  const double ang = s->airTurnAngle * (M_PI / 180.0);
  w.sinWin = (T)sin(ang);
  w.cosWin = (T)cos(ang);
  w.wmin = (T)s->wSpeedMin;
  w.wmin2 = (T)(s->wSpeedMin * s->wSpeedMin);
  w.caN = (T)s->airIceDrag;
  w.caS = (T)s->airIceDrag_south;
  w.rhoAir = (T)s->rhoAir;
  w.grav = (T)s->gravity;
  w.rrho = (T)s->recip_rhoSea;
  w.loadFac = (T)s->seaIceLoadFac;
  hipLaunchKernelGGL(wind_forcing_kernel<T>, cell_grid(g, BLK), dim3(BLK), 0, st, g, w, (T*)X,
                     (T*)Y);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int ocean_stress(const Geom& g, const vh_settings* s, const vh_fields* f, void* U, void* V,
                 hipStream_t st) {
  DragArgs<T> a;
  a.u = C_<T>(f->uIce);
  a.v = C_<T>(f->vIce);
  a.uOc = C_<T>(f->uOcean);
  a.vOc = C_<T>(f->vOcean);
  a.fC = C_<T>(f->fCori);
  a.mInU = C_<T>(f->maskInU);
  a.mInV = C_<T>(f->maskInV);
  a.M = C_<T>(f->iceMask);
  const void* req[] = {a.u, a.v, a.uOc, a.vOc, a.fC, a.mInU, a.mInV, a.M, U, V};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "ocean_stress_uv: null pointer");
  a.cwN = (T)s->waterIceDrag;
  a.cwS = (T)s->waterIceDrag_south;
  a.rhoSea = (T)s->rhoSea;
  a.cmin = (T)s->cDragMin;
  a.cmin2 = (T)(s->cDragMin * s->cDragMin);
  const double ang = s->waterTurnAngle * (M_PI / 180.0);
  hipLaunchKernelGGL(ocean_stress_kernel<T>, cell_grid(g, BLK), dim3(BLK), 0, st, g, a,
                     (T)sin(ang), (T)cos(ang), (T*)U, (T*)V);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int clean_up(const Geom& g, const vh_settings* s, const vh_fields* f, void* h, void* hs, void* A,
             void* Ts, void* osh, void* oshs, hipStream_t st) {
  const void* req[] = {f->hIceMean, f->hSnowMean, f->Area, f->TSurf, h, hs, A, Ts, osh, oshs};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "clean_up_advection: null pointer");
  hipLaunchKernelGGL(clean_up_kernel<T>, cell_grid(g, BLK), dim3(BLK), 0, st, g,
                     C_<T>(f->hIceMean), C_<T>(f->hSnowMean), C_<T>(f->Area), C_<T>(f->TSurf),
                     (T)s->hIce_min, (T)s->Area_min, (T)s->celsius2K, (T*)h, (T*)hs, (T*)A, (T*)Ts,
                     (T*)osh, (T*)oshs);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int free_drift(const Geom& g, const vh_settings* s, const vh_fields* f, void* U, void* V,
               hipStream_t st) {
  FreeDriftArgs<T> a;
  a.WFX = C_<T>(f->WindForcingX);
  a.WFY = C_<T>(f->WindForcingY);
  a.h = C_<T>(f->hIceMean);
  a.fC = C_<T>(f->fCori);
  a.uOc = C_<T>(f->uOcean);
  a.vOc = C_<T>(f->vOcean);
  a.MU = C_<T>(f->iceMaskU);
  a.MV = C_<T>(f->iceMaskV);
  const void* req[] = {a.WFX, a.WFY, a.h, a.fC, a.uOc, a.vOc, a.MU, a.MV};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "freedrift_solver: null pointer");
  VH_REQUIRE(U && V && U != V, VH_EINVAL, "freedrift_solver: null or aliased output");
  for (const void* p : req)
    VH_REQUIRE(U != p && V != p, VH_EINVAL, "freedrift_solver: output aliases an input");
  a.rhoIce = (T)s->rhoIce;
  a.cwN = (T)s->waterIceDrag;
  a.cwS = (T)s->waterIceDrag_south;
  a.rhoSea = (T)s->rhoSea;
  hipLaunchKernelGGL(free_drift_kernel<T>, cell_grid(g, BLK), dim3(BLK), 0, st, g, a, (T*)U,
                     (T*)V);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int ridging(const Geom& g, const vh_fields* f, void* A, hipStream_t st) {
  VH_REQUIRE(f->Area && A, VH_EINVAL, "ridging: null pointer");
  hipLaunchKernelGGL(ridging_kernel<T>, cell_grid(g, BLK), dim3(BLK), 0, st, g, C_<T>(f->Area),
                     (T*)A);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int fill(const Geom& g, void* a, hipStream_t st) {
  VH_REQUIRE(a != nullptr, VH_EINVAL, "fill_overlap: null field");
  if (!g.fill_x && !g.fill_y) return VH_OK;
  hipLaunchKernelGGL(fill_kernel<T>, cell_grid(g, BLK), dim3(BLK), 0, st, g, (T*)a);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int fill_many(const Geom& g, int n, void* const* fields, hipStream_t st) {
  VH_REQUIRE(n >= 0 && n <= kFillMax, VH_EINVAL, "fill_overlaps: %d fields (at most %d)", n,
             kFillMax);
  VH_REQUIRE(n == 0 || fields != nullptr, VH_EINVAL, "fill_overlaps: null field list");
  FillList<T> fl{};
  fl.n = n;
  for (int q = 0; q < n; ++q) {
    VH_REQUIRE(fields[q] != nullptr, VH_EINVAL, "fill_overlaps: null field %d", q);
    for (int r = 0; r < q; ++r)
      VH_REQUIRE(fields[q] != fields[r], VH_EINVAL, "fill_overlaps: field %d repeats field %d", q, r);
    fl.f[q] = (T*)fields[q];
  }
  if (n == 0 || (!g.fill_x && !g.fill_y)) return VH_OK;
  hipLaunchKernelGGL(fill_many_kernel<T>, cell_grid(g, BLK), dim3(BLK), 0, st, g, fl);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int gsum(const Geom& g, const void* x, const void* w, double* out, void* ws, size_t wsb,
         hipStream_t st) {
  VH_REQUIRE(x && out, VH_EINVAL, "global_sum: null pointer");
  VH_REQUIRE(ws && wsb >= SUM_BLOCKS * sizeof(double), VH_EWORKSPACE,
             "global_sum workspace: need %zu bytes", SUM_BLOCKS * sizeof(double));
  hipLaunchKernelGGL(sum_partial_kernel<T>, dim3(SUM_BLOCKS), dim3(BLK), 0, st, g, C_<T>(x),
                     C_<T>(w), (double*)ws);
  hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(BLK), 0, st, (const double*)ws, SUM_BLOCKS,
                     out);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

// ---------------------------------------------------------------- fused pre-EVP

// SeaIceMass, AreaWS, WindForcingXY and SeaIceStrength in one sweep; every
// output is the same expression, in the same order, as the separate kernels
// (masses and areas of the neighbours are recomputed instead of re-read).
template <typename T>
struct PreEvpIO {
  const T *h, *hs, *A, *M, *psf;
  T ps, cStar, rhoI, rhoS;
  T *mC, *mU, *mV, *AW, *AS, *WX, *WY, *P;
};

template <typename T>
__device__ inline void pre_evp_cell(const Geom& g, const WindArgs<T>& w, const PreEvpIO<T>& o,
                                    size_t idx) {
  const int i = (int)(idx / g.NY);
  const int j = (int)(idx - (size_t)i * g.NY);
  const int im = (i == 0 ? g.NX - 1 : i - 1);
  const int jm = (j == 0 ? g.NY - 1 : j - 1);
  const T c = o.rhoI * o.h[idx] + o.rhoS * o.hs[idx];
  const T cim = o.rhoI * o.h[AT(im, j)] + o.rhoS * o.hs[AT(im, j)];
  const T cjm = o.rhoI * o.h[AT(i, jm)] + o.rhoS * o.hs[AT(i, jm)];
  const T mU = (T)0.5 * (c + cim);
  const T mV = (T)0.5 * (c + cjm);
  const T a = o.A[idx];
  const T AW = (T)0.5 * (a + o.A[AT(im, j)]);
  const T AS = (T)0.5 * (a + o.A[AT(i, jm)]);
  T tx, ty, txim, tyim, txjm, tyjm;
  tau_c(g, w, i, j, &tx, &ty);
  tau_c(g, w, im, j, &txim, &tyim);
  tau_c(g, w, i, jm, &txjm, &tyjm);
  const T tauX = ((T)0.5 * (tx + txim)) * w.MU[idx];
  const T tauY = ((T)0.5 * (ty + tyjm)) * w.MV[idx];
  const T phi = phi_at(w, idx);
  T fx = tauX * AW;
  T fy = tauY * AS;
  fx = fx - (mU * w.rdxC[idx]) * (phi - phi_at(w, AT(im, j)));
  fy = fy - (mV * w.rdyC[idx]) * (phi - phi_at(w, AT(i, jm)));
  const T pstar = o.psf ? o.psf[idx] : o.ps;
  o.mC[idx] = c;
  o.mU[idx] = mU;
  o.mV[idx] = mV;
  o.AW[idx] = AW;
  o.AS[idx] = AS;
  o.WX[idx] = fx;
  o.WY[idx] = fy;
  o.P[idx] = ((pstar * o.h[idx]) * exp((-o.cStar) * ((T)1 - a))) * o.M[idx];
}

template <typename T>
__global__ __launch_bounds__(BLK) void pre_evp_kernel(Geom g, WindArgs<T> w, PreEvpIO<T> o) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)g.NX * g.NY) return;
  pre_evp_cell(g, w, o, idx);
}

// Members of a batch (packed fields, each member's padded box at its element
// offset), passed by value: up to kBatchChunk per launch.
struct BatchMember {
  long long off;
  Geom g;
};
constexpr int kBatchChunk = 64;
struct BatchChunk {
  int count;
  BatchMember m[kBatchChunk];
};

template <typename P>
__device__ inline P* at_off(P* p, long long off) {
  return p ? p + off : p;
}

template <typename T>
__device__ inline WindArgs<T> shifted(WindArgs<T> w, long long off) {
  w.u = at_off(w.u, off);
  w.v = at_off(w.v, off);
  w.uW = at_off(w.uW, off);
  w.vW = at_off(w.vW, off);
  w.fC = at_off(w.fC, off);
  w.MU = at_off(w.MU, off);
  w.MV = at_off(w.MV, off);
  w.AW = at_off(w.AW, off);
  w.AS = at_off(w.AS, off);
  w.mU = at_off(w.mU, off);
  w.mV = at_off(w.mV, off);
  w.rdxC = at_off(w.rdxC, off);
  w.rdyC = at_off(w.rdyC, off);
  w.ssh = at_off(w.ssh, off);
  w.sp = at_off(w.sp, off);
  w.load = at_off(w.load, off);
  return w;
}

template <typename T>
__device__ inline PreEvpIO<T> shifted(PreEvpIO<T> o, long long off) {
  o.h = at_off(o.h, off);
  o.hs = at_off(o.hs, off);
  o.A = at_off(o.A, off);
  o.M = at_off(o.M, off);
  o.psf = at_off(o.psf, off);
  o.mC = at_off(o.mC, off);
  o.mU = at_off(o.mU, off);
  o.mV = at_off(o.mV, off);
  o.AW = at_off(o.AW, off);
  o.AS = at_off(o.AS, off);
  o.WX = at_off(o.WX, off);
  o.WY = at_off(o.WY, off);
  o.P = at_off(o.P, off);
  return o;
}

// The pre-EVP sweep of every member of a chunk in one launch: blockIdx.y is
// the member, blockIdx.x runs over its padded cells.
template <typename T>
__global__ __launch_bounds__(BLK) void pre_evp_batched_kernel(BatchChunk ch, WindArgs<T> w,
                                                              PreEvpIO<T> o) {
  const BatchMember& m = ch.m[blockIdx.y];
  const Geom g = m.g;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)g.NX * g.NY) return;
  pre_evp_cell(g, shifted(w, m.off), shifted(o, m.off), idx);
}

// FP64 sums over each member's interior of up to kMaxSumFields packed fields
// (times `w` when not null): one workgroup per (member, field), a fixed
// summation order, result [field * n_members + first + member].
constexpr int kMaxSumFields = 8;
template <typename T>
struct SumFields {
  const T* x[kMaxSumFields];
  const T* w;
  double* out;
  int n_members, first;
};

template <typename T>
__global__ __launch_bounds__(BLK) void sums_batched_kernel(BatchChunk ch, SumFields<T> a) {
  const BatchMember& m = ch.m[blockIdx.x];
  const Geom g = m.g;
  const T* x = a.x[blockIdx.y] + m.off;
  const T* w = a.w ? a.w + m.off : nullptr;
  __shared__ double red[BLK / 64];
  const int n = g.nx * g.ny;
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += BLK) {
    const int ii = k / g.ny + g.ol, jj = k - (k / g.ny) * g.ny + g.ol;
    const size_t c = AT(ii, jj);
    double v = (double)x[c];
    if (w) v = v * (double)w[c];
    acc += v;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < BLK / 64; ++q) s += red[q];
    a.out[(size_t)blockIdx.y * a.n_members + a.first + blockIdx.x] = s;
  }
}

// Arguments of the pre-EVP sweep (validated), shared by the single-domain and
// the batched launch.
template <typename T>
int pre_evp_args(const vh_settings* s, const vh_fields* f, const void* psf,
                 const vh_pre_evp_out* out, WindArgs<T>* wp, PreEvpIO<T>* op) {
  VH_REQUIRE(out != nullptr, VH_EINVAL, "pre_evp: null output struct");
  WindArgs<T>& w = *wp;
  PreEvpIO<T>& o = *op;
  w.u = C_<T>(f->uIce);
  w.v = C_<T>(f->vIce);
  w.uW = C_<T>(f->uWind);
  w.vW = C_<T>(f->vWind);
  w.fC = C_<T>(f->fCori);
  w.MU = C_<T>(f->iceMaskU);
  w.MV = C_<T>(f->iceMaskV);
  w.AW = nullptr;
  w.AS = nullptr;
  w.mU = nullptr;
  w.mV = nullptr;
  w.rdxC = C_<T>(f->recip_dxC);
  w.rdyC = C_<T>(f->recip_dyC);
  w.ssh = C_<T>(f->ssh_an);
  w.sp = C_<T>(f->surfPress);
  w.load = C_<T>(f->SeaIceLoad);
  w.relWind = s->useRelativeWind != 0.0;
  w.realFW = s->useRealFreshWaterFlux != 0.0;
  o.h = C_<T>(f->hIceMean);
  o.hs = C_<T>(f->hSnowMean);
  o.A = C_<T>(f->Area);
  o.M = C_<T>(f->iceMask);
  o.psf = C_<T>(psf);
  o.mC = (T*)out->SeaIceMassC;
  o.mU = (T*)out->SeaIceMassU;
  o.mV = (T*)out->SeaIceMassV;
  o.AW = (T*)out->AreaW;
  o.AS = (T*)out->AreaS;
  o.WX = (T*)out->WindForcingX;
  o.WY = (T*)out->WindForcingY;
  o.P = (T*)out->SeaIceStrength;
  const void* req[] = {w.uW, w.vW, w.fC, w.MU, w.MV, w.rdxC, w.rdyC, w.ssh, w.sp, o.h, o.hs, o.A,
                       o.M, o.mC, o.mU, o.mV, o.AW, o.AS, o.WX, o.WY, o.P};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "pre_evp: null pointer");
  if (w.relWind) VH_REQUIRE(w.u && w.v, VH_EINVAL, "pre_evp: null uIce/vIce");
  if (w.realFW) VH_REQUIRE(w.load, VH_EINVAL, "pre_evp: null SeaIceLoad");
  T* outs[] = {o.mC, o.mU, o.mV, o.AW, o.AS, o.WX, o.WY, o.P};
  const void* ins[] = {w.u, w.v, w.uW, w.vW, w.fC, w.MU, w.MV, w.rdxC, w.rdyC, w.ssh, w.sp, w.load,
                       o.h, o.hs, o.A, o.M, o.psf};
  for (int q = 0; q < 8; ++q) {
    for (int r = 0; r < q; ++r) VH_REQUIRE(outs[q] != outs[r], VH_EINVAL, "pre_evp: outputs alias");
    for (const void* p : ins) VH_REQUIRE(outs[q] != p, VH_EINVAL, "pre_evp: output aliases an input");
  }
  const double ang = s->airTurnAngle * (M_PI / 180.0);
  w.sinWin = (T)sin(ang);
  w.cosWin = (T)cos(ang);
  w.wmin = (T)s->wSpeedMin;
  w.wmin2 = (T)(s->wSpeedMin * s->wSpeedMin);
  w.caN = (T)s->airIceDrag;
  w.caS = (T)s->airIceDrag_south;
  w.rhoAir = (T)s->rhoAir;
  w.grav = (T)s->gravity;
  w.rrho = (T)s->recip_rhoSea;
  w.loadFac = (T)s->seaIceLoadFac;
  o.ps = (T)s->pStar;
  o.cStar = (T)s->cStar;
  o.rhoI = (T)s->rhoIce;
  o.rhoS = (T)s->rhoSnow;
  return VH_OK;
}

template <typename T>
int pre_evp(const Geom& g, const vh_settings* s, const vh_fields* f, const void* psf,
            const vh_pre_evp_out* out, hipStream_t st) {
  WindArgs<T> w;
  PreEvpIO<T> o;
  int rc = pre_evp_args<T>(s, f, psf, out, &w, &o);
  if (rc) return rc;
  hipLaunchKernelGGL(pre_evp_kernel<T>, cell_grid(g, BLK), dim3(BLK), 0, st, g, w, o);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

// Members of a batch in chunks of kBatchChunk (validated like the batched EVP).
int batch_chunks(const vh_batch* b, std::vector<BatchChunk>* chunks, long long* max_cells) {
  VH_REQUIRE(b != nullptr && b->members != nullptr && b->n_members > 0, VH_EINVAL, "empty batch");
  VH_REQUIRE(b->dtype == VH_F64 || b->dtype == VH_F32, VH_EDTYPE, "dtype %d", b->dtype);
  *max_cells = 0;
  chunks->clear();
  for (int m = 0; m < b->n_members; ++m) {
    if (m % kBatchChunk == 0) chunks->push_back(BatchChunk{});
    BatchChunk& ch = chunks->back();
    BatchMember& bm = ch.m[ch.count++];
    const vh_member& mm = b->members[m];
    vh_domain d{mm.nx, mm.ny, 2, 2, mm.cyclic_x, mm.fill_mode, b->dtype, 0};
    int rc = make_geom(&d, &bm.g);
    if (rc) return rc;
    bm.off = mm.offset;
    const long long n = (long long)bm.g.NX * bm.g.NY;
    VH_REQUIRE(mm.offset >= 0 && mm.offset + n <= b->total_elems, VH_ESHAPE,
               "member %d box [%lld, %lld) outside the packed fields (%lld)", m,
               (long long)mm.offset, (long long)mm.offset + n, (long long)b->total_elems);
    *max_cells = n > *max_cells ? n : *max_cells;
  }
  return VH_OK;
}

template <typename T>
int pre_evp_batched(const vh_batch* b, const vh_settings* s, const vh_fields* f, const void* psf,
                    const vh_pre_evp_out* out, hipStream_t st) {
  std::vector<BatchChunk> chunks;
  long long max_cells = 0;
  int rc = batch_chunks(b, &chunks, &max_cells);
  if (rc) return rc;
  WindArgs<T> w;
  PreEvpIO<T> o;
  rc = pre_evp_args<T>(s, f, psf, out, &w, &o);
  if (rc) return rc;
  for (const BatchChunk& ch : chunks)
    hipLaunchKernelGGL(pre_evp_batched_kernel<T>,
                       dim3((unsigned)((max_cells + BLK - 1) / BLK), ch.count), dim3(BLK), 0, st,
                       ch, w, o);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int sums_batched(const vh_batch* b, int nf, const void* const* fields, const void* wgt,
                 double* result, hipStream_t st) {
  std::vector<BatchChunk> chunks;
  long long max_cells = 0;
  int rc = batch_chunks(b, &chunks, &max_cells);
  if (rc) return rc;
  VH_REQUIRE(nf >= 1 && nf <= kMaxSumFields, VH_EINVAL, "n_fields %d (1..%d)", nf, kMaxSumFields);
  VH_REQUIRE(fields != nullptr && result != nullptr, VH_EINVAL, "global sums: null pointer");
  SumFields<T> a{};
  for (int q = 0; q < nf; ++q) {
    VH_REQUIRE(fields[q] != nullptr, VH_EINVAL, "global sums: field %d is null", q);
    a.x[q] = C_<T>(fields[q]);
  }
  a.w = C_<T>(wgt);
  a.out = result;
  a.n_members = b->n_members;
  int first = 0;
  for (const BatchChunk& ch : chunks) {
    a.first = first;
    hipLaunchKernelGGL(sums_batched_kernel<T>, dim3(ch.count, nf), dim3(BLK), 0, st, ch, a);
    first += ch.count;
  }
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

// ---------------------------------------------------------------- ocean coupling

// Fields the host ocean hands to the sea ice before a step
// (veris/setup/seaice_global_4deg/seaice_global_4deg.py:847, 868-876); wSpeed is formed
// from uWind/vWind before their overlap fill (:865), as there.
template <typename T>
struct OceanToIce {
  const T *uS, *vS, *tS, *sS, *ftemp, *SWdown, *uW, *vW;
  T *uO, *vO, *theta, *salt, *Qnet, *Qsw, *wSpeed;
  T c2K, cp, rho;
};

template <typename T>
__global__ __launch_bounds__(BLK) void ocean_to_ice_kernel(Geom g, OceanToIce<T> a) {
  CELL_PROLOGUE
  a.uO[idx] = a.uS[idx];
  a.vO[idx] = a.vS[idx];
  a.theta[idx] = a.tS[idx] + a.c2K;
  a.salt[idx] = a.sS[idx];
  a.Qnet[idx] = ((-a.ftemp[idx]) * a.cp) * a.rho;
  a.Qsw[idx] = -a.SWdown[idx];
  const T u = a.uW[idx], v = a.vW[idx];
  a.wSpeed[idx] = sqrt(u * u + v * v);
}

// Fluxes the sea ice hands back to the ocean after a step
// (veris/setup/seaice_global_4deg/seaice_global_4deg.py:933-940).
template <typename T>
struct IceToOcean {
  const T *tx, *ty, *AW, *AS, *OSU, *OSV, *Qnet, *fssi, *fssr;
  T *txo, *tyo, *ftemp, *fsalt;
  T cpRho;
};

template <typename T>
__global__ __launch_bounds__(BLK) void ice_to_ocean_kernel(Geom g, IceToOcean<T> a) {
  CELL_PROLOGUE
  const T aw = a.AW[idx], as = a.AS[idx];
  a.txo[idx] = a.tx[idx] * ((T)1 - aw) + a.OSU[idx] * aw;
  a.tyo[idx] = a.ty[idx] * ((T)1 - as) + a.OSV[idx] * as;
  a.ftemp[idx] = (-a.Qnet[idx]) / a.cpRho;
  a.fsalt[idx] = a.fssi[idx] + a.fssr[idx];
}

}  // namespace
}  // namespace vh

#define VH_DISPATCH(d, call_f64, call_f32) ((d)->dtype == VH_F64 ? (call_f64) : (call_f32))
#define VH_PRE(d, s, f)                                                    \
  vh::Geom g;                                                              \
  {                                                                        \
    int rc_ = vh::make_geom(d, &g);                                        \
    if (rc_) return rc_;                                                   \
  }                                                                        \
  VH_REQUIRE((s) != nullptr && (f) != nullptr, VH_EINVAL, "null settings or fields"); \
  hipStream_t st = (hipStream_t)stream;

extern "C" int vh_seaice_mass(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                              void* C, void* U, void* V, void* stream) {
  VH_PRE(d, s, f)
  return VH_DISPATCH(d, vh::seaice_mass<double>(g, s, f, C, U, V, st),
                     vh::seaice_mass<float>(g, s, f, C, U, V, st));
}

extern "C" int vh_area_ws(const vh_domain* d, const vh_settings* s, const vh_fields* f, void* W,
                          void* S, void* stream) {
  VH_PRE(d, s, f)
  return VH_DISPATCH(d, vh::area_ws<double>(g, f, W, S, st), vh::area_ws<float>(g, f, W, S, st));
}

extern "C" int vh_seaice_strength(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                                  const void* psf, void* P, void* stream) {
  VH_PRE(d, s, f)
  return VH_DISPATCH(d, vh::strength<double>(g, s, f, psf, P, st),
                     vh::strength<float>(g, s, f, psf, P, st));
}

extern "C" int vh_wind_forcing_xy(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                                  void* X, void* Y, void* stream) {
  VH_PRE(d, s, f)
  return VH_DISPATCH(d, vh::wind_forcing<double>(g, s, f, X, Y, st),
                     vh::wind_forcing<float>(g, s, f, X, Y, st));
}

extern "C" int vh_ocean_stress_uv(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                                  void* U, void* V, void* stream) {
  VH_PRE(d, s, f)
  return VH_DISPATCH(d, vh::ocean_stress<double>(g, s, f, U, V, st),
                     vh::ocean_stress<float>(g, s, f, U, V, st));
}

extern "C" int vh_freedrift_solver(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                                   void* uIce, void* vIce, void* stream) {
  VH_PRE(d, s, f)
  return VH_DISPATCH(d, vh::free_drift<double>(g, s, f, uIce, vIce, st),
                     vh::free_drift<float>(g, s, f, uIce, vIce, st));
}

extern "C" int vh_clean_up_advection(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                                     void* h, void* hs, void* A, void* Ts, void* osh, void* oshs,
                                     void* stream) {
  VH_PRE(d, s, f)
  return VH_DISPATCH(d, vh::clean_up<double>(g, s, f, h, hs, A, Ts, osh, oshs, st),
                     vh::clean_up<float>(g, s, f, h, hs, A, Ts, osh, oshs, st));
}

extern "C" int vh_ridging(const vh_domain* d, const vh_settings* s, const vh_fields* f, void* A,
                          void* stream) {
  VH_PRE(d, s, f)
  return VH_DISPATCH(d, vh::ridging<double>(g, f, A, st), vh::ridging<float>(g, f, A, st));
}

extern "C" int vh_fill_overlap(const vh_domain* d, void* a, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return VH_DISPATCH(d, vh::fill<double>(g, a, st), vh::fill<float>(g, a, st));
}

extern "C" int vh_fill_overlaps(const vh_domain* d, int32_t n, void* const* fields, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return VH_DISPATCH(d, vh::fill_many<double>(g, n, fields, st),
                     vh::fill_many<float>(g, n, fields, st));
}

extern "C" int vh_global_sum_workspace_bytes(const vh_domain* d, size_t* bytes) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(bytes != nullptr, VH_EINVAL, "null bytes");
  *bytes = vh::SUM_BLOCKS * sizeof(double);
  return VH_OK;
}

extern "C" int vh_global_sum_f64(const vh_domain* d, const void* x, const void* w, double* out,
                                 void* ws, size_t wsb, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  return VH_DISPATCH(d, vh::gsum<double>(g, x, w, out, ws, wsb, st),
                     vh::gsum<float>(g, x, w, out, ws, wsb, st));
}

extern "C" int vh_pre_evp(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                          const void* pStarField, const vh_pre_evp_out* out, void* stream) {
  VH_PRE(d, s, f)
  return VH_DISPATCH(d, vh::pre_evp<double>(g, s, f, pStarField, out, st),
                     vh::pre_evp<float>(g, s, f, pStarField, out, st));
}

extern "C" int vh_pre_evp_batched(const vh_batch* b, const vh_settings* s, const vh_fields* f,
                                  const void* pStarField, const vh_pre_evp_out* out,
                                  void* stream) {
  VH_REQUIRE(b != nullptr && s != nullptr && f != nullptr, VH_EINVAL, "null batch, settings or fields");
  hipStream_t st = (hipStream_t)stream;
  if (b->dtype == VH_F32) return vh::pre_evp_batched<float>(b, s, f, pStarField, out, st);
  return vh::pre_evp_batched<double>(b, s, f, pStarField, out, st);
}

extern "C" int vh_global_sums_batched(const vh_batch* b, int32_t n_fields,
                                      const void* const* fields, const void* weight,
                                      double* result, void* stream) {
  VH_REQUIRE(b != nullptr, VH_EINVAL, "null batch");
  hipStream_t st = (hipStream_t)stream;
  if (b->dtype == VH_F32) return vh::sums_batched<float>(b, n_fields, fields, weight, result, st);
  return vh::sums_batched<double>(b, n_fields, fields, weight, result, st);
}

namespace {

template <typename T>
int ocean_to_ice_t(const vh::Geom& g, const vh_settings* s, const vh_fields* f, const void* u,
                   const void* v, const void* temp, const void* salt, const void* ftemp,
                   const vh_ocean_to_ice_out* o, hipStream_t st) {
  vh::OceanToIce<T> a;
  a.uS = (const T*)u;
  a.vS = (const T*)v;
  a.tS = (const T*)temp;
  a.sS = (const T*)salt;
  a.ftemp = (const T*)ftemp;
  a.SWdown = (const T*)f->SWdown;
  a.uW = (const T*)f->uWind;
  a.vW = (const T*)f->vWind;
  a.uO = (T*)o->uOcean;
  a.vO = (T*)o->vOcean;
  a.theta = (T*)o->theta;
  a.salt = (T*)o->ocSalt;
  a.Qnet = (T*)o->Qnet;
  a.Qsw = (T*)o->Qsw;
  a.wSpeed = (T*)o->wSpeed;
  const void* req[] = {a.uS, a.vS, a.tS, a.sS, a.ftemp, a.SWdown, a.uW, a.vW, a.uO, a.vO,
                       a.theta, a.salt, a.Qnet, a.Qsw, a.wSpeed};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "ocean_to_ice: null pointer");
  a.c2K = (T)s->celsius2K;
  a.cp = (T)s->cpWater;
  a.rho = (T)s->rhoSea;
  hipLaunchKernelGGL(vh::ocean_to_ice_kernel<T>, vh::cell_grid(g, 256), dim3(256), 0, st, g, a);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int ice_to_ocean_t(const vh::Geom& g, const vh_settings* s, const vh_fields* f, const void* taux,
                   const void* tauy, const void* fsi, const void* fsr, const vh_ice_to_ocean_out* o,
                   hipStream_t st) {
  vh::IceToOcean<T> a;
  a.tx = (const T*)taux;
  a.ty = (const T*)tauy;
  a.AW = (const T*)f->AreaW;
  a.AS = (const T*)f->AreaS;
  a.OSU = (const T*)f->OceanStressU;
  a.OSV = (const T*)f->OceanStressV;
  a.Qnet = (const T*)f->Qnet;
  a.fssi = (const T*)fsi;
  a.fssr = (const T*)fsr;
  a.txo = (T*)o->surface_taux;
  a.tyo = (T*)o->surface_tauy;
  a.ftemp = (T*)o->forc_temp_surface;
  a.fsalt = (T*)o->forc_salt_surface;
  const void* req[] = {a.tx, a.ty, a.AW, a.AS, a.OSU, a.OSV, a.Qnet, a.fssi, a.fssr, a.txo, a.tyo,
                       a.ftemp, a.fsalt};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "ice_to_ocean: null pointer");
  a.cpRho = (T)(s->cpWater * s->rhoSea);
  hipLaunchKernelGGL(vh::ice_to_ocean_kernel<T>, vh::cell_grid(g, 256), dim3(256), 0, st, g, a);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

}  // namespace

extern "C" int vh_ocean_to_ice(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                               const void* u_surface, const void* v_surface,
                               const void* temp_surface, const void* salt_surface,
                               const void* forc_temp_surface, const vh_ocean_to_ice_out* out,
                               void* stream) {
  VH_PRE(d, s, f)
  VH_REQUIRE(out != nullptr, VH_EINVAL, "ocean_to_ice: null output struct");
  return VH_DISPATCH(d,
                     ocean_to_ice_t<double>(g, s, f, u_surface, v_surface, temp_surface,
                                            salt_surface, forc_temp_surface, out, st),
                     ocean_to_ice_t<float>(g, s, f, u_surface, v_surface, temp_surface,
                                           salt_surface, forc_temp_surface, out, st));
}

extern "C" int vh_ice_to_ocean(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                               const void* surface_taux, const void* surface_tauy,
                               const void* forc_salt_surface_ice,
                               const void* forc_salt_surface_res, const vh_ice_to_ocean_out* out,
                               void* stream) {
  VH_PRE(d, s, f)
  VH_REQUIRE(out != nullptr, VH_EINVAL, "ice_to_ocean: null output struct");
  return VH_DISPATCH(d,
                     ice_to_ocean_t<double>(g, s, f, surface_taux, surface_tauy,
                                            forc_salt_surface_ice, forc_salt_surface_res, out, st),
                     ice_to_ocean_t<float>(g, s, f, surface_taux, surface_tauy,
                                           forc_salt_surface_ice, forc_salt_surface_res, out, st));
}
