// Hand-written adjoint of the EVP subcycle (veris/evp_solver.py:22-319) with
// segment checkpointing, and of the pre-EVP forcing it depends on
// (veris/dynamics_routines.py:5-18, veris/area_mass.py:17-28,
// veris/dynsolver.py:7-84).
//
// Forward: vh_evp_forward_checkpointed runs the fused substep kernel and keeps
// the state every `ckpt_every` substeps in the workspace.  Reverse: for each
// segment, last first, the states inside it are recomputed from its
// checkpoint (one fused substep launch each), then every substep is reversed
// by six per-cell kernels in gather form -- each cell sums what its
// neighbours contribute, so no atomics are needed:
//   I   forward intermediates of the substep from its input state
//       (e12, zeta, Delta, e11+e22, e11-e22, water drag)
//   R1  u/v-points: momentum update and fill (overlap copies fold back)
//   R2  c-points: stress divergence, stress relaxation, drag, Coriolis;
//       z-points: sigma12 relaxation
//   R3  c-points: zeta and Delta
//   R4  z-points: shear strain
//   R5  u/v-points: adjoint of the input velocities
// Parameter adjoints (strength, masses, areas, wind forcing, previous-step
// velocities, water drag per hemisphere) accumulate cell by cell.
//
// Non-smooth points: d sqrt(x)/dx is taken as 0 at x = 0 (Delta = 0 on
// motionless or masked cells, zero relative velocity), the denominator guard
// and the drag regime switch contribute no derivative.  Supported settings:
// non-adaptive EVP, noSlip, no turning angle (sinWat = 0), no basal drag, no
// metric terms -- the defaults.
#include <algorithm>
#include <cstdio>

#include "evp_kernel.h"

namespace vh {

template <typename T>
int evp_run_core(const Geom& g, const vh_settings* s, const vh_fields* fl, const T* const in[5],
                 T* const out[5], int nsub, void* ws, size_t ws_bytes, int tile_req, int dtype,
                 hipStream_t st);

namespace {

constexpr int BLK = 256;

template <typename T>
struct AdjConst {
  const T *M, *MU, *MV, *mInC, *mInU, *mInV, *rdxU, *rdyV, *rdyU, *rdxV, *dxU, *dyU, *dxV, *dyV;
  const T *rrAu, *rrAv, *rAz, *rrA, *fC, *uOc, *vOc, *P, *mC, *mU, *mV, *AW, *AS, *uNm1, *vNm1;
};

template <typename T>
struct AdjWork {
  T *E12, *ZETA, *DELTA, *EP, *EM, *CD;                     // forward intermediates
  T *X, *Y, *FX, *FY, *CDU, *CDV, *UDIR, *VDIR;             // R1
  T *ZBP, *EPB, *EMB, *DBP, *AB, *BB, *FUB, *FVB, *ZZB, *E12P;  // R2
  T *E11B, *E22B, *QB;                                      // R3
  T *E12B;                                                  // R4
};

template <typename T>
struct AdjAcc {  // parameter adjoints, accumulated over substeps
  T *P, *mC, *mU, *mV, *AW, *AS, *WFX, *WFY, *u0, *v0, *DN, *DS;
};

#define CELL_SETUP                                                                  \
  const int n_ = g.NX * g.NY;                                                       \
  const int c = blockIdx.x * blockDim.x + threadIdx.x;                              \
  if (c >= n_) return;                                                              \
  const int i = c / g.NY, j = c - (c / g.NY) * g.NY;                                \
  const int im = i == 0 ? g.NX - 1 : i - 1, ip = i == g.NX - 1 ? 0 : i + 1;         \
  const int jm = j == 0 ? g.NY - 1 : j - 1, jp = j == g.NY - 1 ? 0 : j + 1;         \
  (void)im; (void)ip; (void)jm; (void)jp;
#define AT(ii, jj) ((ii) * g.NY + (jj))

template <typename T>
__device__ inline T e12_at(const Geom& g, const AdjConst<T>& k, const T* u, const T* v, int ii,
                           int jj) {
  const int im = ii == 0 ? g.NX - 1 : ii - 1, jm = jj == 0 ? g.NY - 1 : jj - 1;
  const int c = AT(ii, jj);
  const T u0 = u[c], ujm = u[AT(ii, jm)], v0 = v[c], vim = v[AT(im, jj)];
  const T rdyU = k.rdyU[c], rdxV = k.rdxV[c];
  const T mz = (k.M[c] * k.M[AT(im, jj)]) * (k.M[AT(ii, jm)] * k.M[AT(im, jm)]);
  const T hu = k.MU[c] - k.MU[AT(ii, jm)], hv = k.MV[c] - k.MV[AT(im, jj)];
  return ((T)0.5 * ((u0 - ujm) * rdyU + (v0 - vim) * rdxV)) * mz +
         ((((T)2 * ((u0 + ujm) * (T)0.5)) * rdyU) * hu +
          (((T)2 * ((v0 + vim) * (T)0.5)) * rdxV) * hv);
}

template <typename T>
__device__ inline T znorm(const Geom& g, const AdjConst<T>& k, int ii, int jj) {
  const int im = ii == 0 ? g.NX - 1 : ii - 1, jm = jj == 0 ? g.NY - 1 : jj - 1;
  const T n = (k.M[AT(ii, jj)] + k.M[AT(im, jj)]) + (k.M[AT(ii, jm)] + k.M[AT(im, jm)]);
  return n > (T)0 ? (T)1 / n : (T)0;
}

// sum of a u/v adjoint over a computed cell and the overlap cells the fill
// copies it to (the adjoint of the periodic fill)
template <typename T>
__device__ inline T gather_filled(const Geom& g, const T* a, int i, int j) {
  int ix[2] = {i, -1}, jy[2] = {j, -1};
  if (g.fill_x) {
    if (i < 2 * g.ol) ix[1] = i + g.nx;
    else if (i >= g.nx) ix[1] = i - g.nx;
  }
  if (g.fill_y) {
    if (j < 2 * g.ol) jy[1] = j + g.ny;
    else if (j >= g.ny) jy[1] = j - g.ny;
  }
  T s = 0;
  for (int p = 0; p < 2; ++p)
    for (int q = 0; q < 2; ++q)
      if (ix[p] >= 0 && jy[q] >= 0) s += a[AT(ix[p], jy[q])];
  return s;
}

// ---- I: forward intermediates from the substep's input state
template <typename T>
__global__ __launch_bounds__(BLK) void adj_interm_kernel(Geom g, AdjConst<T> k, EvpScalars sc,
                                                         const T* u, const T* v, AdjWork<T> w) {
  CELL_SETUP
  const T e12 = e12_at(g, k, u, v, i, j);
  w.E12[c] = e12;
  auto qz = [&](int ii, int jj) {
    const T e = e12_at(g, k, u, v, ii, jj);
    return k.rAz[AT(ii, jj)] * (e * e);
  };
  const T e12c = ((T)0.25 * k.rrA[c]) * ((qz(i, j) + qz(ip, j)) + (qz(i, jp) + qz(ip, jp)));
  const T e11 = ((u[AT(ip, j)] - u[c]) * k.rdxU[c]) * k.mInC[c];
  const T e22 = ((v[AT(i, jp)] - v[c]) * k.rdyV[c]) * k.mInC[c];
  const T ep = e11 + e22, em = e11 - e22;
  const T dsq = ep * ep + (T)sc.rpd * (em * em + (T)4 * e12c);
  const T dC = sqrt(dsq);
  const T zeta = ((T)0.5 * (k.P[c] * (T)sc.tens1p)) / (dC + (T)sc.deltaMin);
  w.EP[c] = ep;
  w.EM[c] = em;
  w.DELTA[c] = dC;
  w.ZETA[c] = zeta;
  const T fc = k.fC[c];
  const T D = (T)(fc < (T)0 ? sc.cwS : sc.cwN) * (T)sc.rhoSea;
  const T a = (u[c] - k.uOc[c]) * k.mInU[c] + (u[AT(ip, j)] - k.uOc[AT(ip, j)]) * k.mInU[AT(ip, j)];
  const T b = (v[c] - k.vOc[c]) * k.mInV[c] + (v[AT(i, jp)] - k.vOc[AT(i, jp)]) * k.mInV[AT(i, jp)];
  const T ww = (T)0.25 * (a * a + b * b);
  const T cd = ((D * D) * ww > (T)sc.cDragMin2) ? D * sqrt(ww) : (T)sc.cDragMin;
  w.CD[c] = cd * k.M[c];
}

// ---- R1: u/v-points (momentum update, fill)
template <typename T>
__global__ __launch_bounds__(BLK) void adj_r1_kernel(Geom g, AdjConst<T> k, EvpScalars sc,
                                                     const T* u, const T* v, const T* un,
                                                     const T* vn, const T* lu_in, const T* lv_in,
                                                     AdjWork<T> w, AdjAcc<T> acc) {
  CELL_SETUP
  if (!is_computed(g, i, j)) {
    w.X[c] = w.Y[c] = w.FX[c] = w.FY[c] = w.CDU[c] = w.CDV[c] = w.UDIR[c] = w.VDIR[c] = (T)0;
    return;
  }
  const T lu = gather_filled(g, lu_in, i, j), lv = gather_filled(g, lv_in, i, j);
  const T cw = (T)sc.cosWat, bU = (T)sc.betaU, bP1 = (T)sc.betaP1, rdt = (T)sc.rdt;
  {  // u-point
    const T cdU = (T)0.5 * (w.CD[c] + w.CD[AT(im, j)]);
    const T mU = k.mU[c], AW = k.AW[c], uo = k.uOc[c];
    const T den = mU * bP1 + AW * (cdU * cw);
    const bool guard = den == (T)0;
    const T de = guard ? (T)1 : den;
    const T Nb = lu * k.MU[c] / de;
    const T db = guard ? (T)0 : -lu * un[c] / de;
    w.X[c] = Nb * k.rrAu[c];
    w.FX[c] = Nb;
    acc.mU[c] += Nb * (bU * u[c] + rdt * k.uNm1[c]) + db * bP1;
    acc.u0[c] += Nb * mU * rdt;
    acc.WFX[c] += Nb;
    acc.AW[c] += Nb * (cdU * cw * uo) + db * (cdU * cw);
    w.CDU[c] = Nb * cw * uo * AW + db * AW * cw;
    w.UDIR[c] = Nb * bU * mU;
  }
  {  // v-point
    const T cdV = (T)0.5 * (w.CD[c] + w.CD[AT(i, jm)]);
    const T mV = k.mV[c], AS = k.AS[c], vo = k.vOc[c];
    const T den = mV * bP1 + AS * (cdV * cw);
    const bool guard = den == (T)0;
    const T de = guard ? (T)1 : den;
    const T Nb = lv * k.MV[c] / de;
    const T db = guard ? (T)0 : -lv * vn[c] / de;
    w.Y[c] = Nb * k.rrAv[c];
    w.FY[c] = Nb;
    acc.mV[c] += Nb * (bU * v[c] + rdt * k.vNm1[c]) + db * bP1;
    acc.v0[c] += Nb * mV * rdt;
    acc.WFY[c] += Nb;
    acc.AS[c] += Nb * (cdV * cw * vo) + db * (cdV * cw);
    w.CDV[c] = Nb * cw * vo * AS + db * AS * cw;
    w.VDIR[c] = Nb * bU * mV;
  }
}

// ---- R2: c-points (stresses, drag, Coriolis) and z-points (sigma12)
template <typename T>
__global__ __launch_bounds__(BLK) void adj_r2_kernel(Geom g, AdjConst<T> k, EvpScalars sc,
                                                     const T* u, const T* v, const T* ls1,
                                                     const T* ls2, const T* ls12, T* ls1o,
                                                     T* ls2o, T* ls12o, AdjWork<T> w,
                                                     AdjAcc<T> acc) {
  CELL_SETUP
  const T Mc = k.M[c], fc = k.fC[c];
  // drag and Coriolis
  const T cdb = (T)0.5 * (w.CDU[c] + w.CDU[AT(ip, j)]) + (T)0.5 * (w.CDV[c] + w.CDV[AT(i, jp)]);
  const T fvb = (T)0.5 * (w.FX[c] + w.FX[AT(ip, j)]);
  const T fub = -(T)0.5 * (w.FY[c] + w.FY[AT(i, jp)]);
  acc.mC[c] += fvb * fc * (T)0.5 * (v[c] + v[AT(i, jp)]) + fub * fc * (T)0.5 * (u[c] + u[AT(ip, j)]);
  w.FVB[c] = fvb;
  w.FUB[c] = fub;
  const T cbe = cdb * Mc;
  const T D = (T)(fc < (T)0 ? sc.cwS : sc.cwN) * (T)sc.rhoSea;
  const T a = (u[c] - k.uOc[c]) * k.mInU[c] + (u[AT(ip, j)] - k.uOc[AT(ip, j)]) * k.mInU[AT(ip, j)];
  const T b = (v[c] - k.vOc[c]) * k.mInV[c] + (v[AT(i, jp)] - k.vOc[AT(i, jp)]) * k.mInV[AT(i, jp)];
  const T ww = (T)0.25 * (a * a + b * b);
  T wb = 0, Db = 0;
  if ((D * D) * ww > (T)sc.cDragMin2) {
    const T sw = sqrt(ww);
    wb = sw > (T)0 ? cbe * D * (T)0.5 / sw : (T)0;
    Db = cbe * sw;
  }
  w.AB[c] = wb * (T)0.5 * a;
  w.BB[c] = wb * (T)0.5 * b;
  if (fc < (T)0) acc.DS[c] += Db * (T)sc.rhoSea;
  else acc.DN[c] += Db * (T)sc.rhoSea;
  // stresses at c
  const T s11b = k.dyV[c] * (w.X[c] - w.X[AT(ip, j)]);
  const T s22b = k.dxU[c] * (w.Y[c] - w.Y[AT(i, jp)]);
  const T s1t = ls1[c] + (T)0.5 * (s11b + s22b);
  const T s2t = ls2[c] + (T)0.5 * (s11b - s22b);
  const T H1 = s1t * (T)sc.den1 * Mc, H2 = s2t * (T)sc.den2 * Mc;
  ls1o[c] = H1 * (T)sc.alphaM1;
  ls2o[c] = H2 * (T)sc.alphaM1;
  const T zeta = w.ZETA[c], dC = w.DELTA[c], ep = w.EP[c], em = w.EM[c];
  T zb = H1 * (T)2 * ep + H2 * (T)0.5 * em;
  const T pb = -H1;
  acc.P[c] += pb * (T)sc.tens1m * (T)sc.pr1m;
  zb += pb * (T)sc.tens1m * (T)2 * dC * (T)sc.pr / (T)sc.tens1p;
  w.ZBP[c] = zb;
  w.EPB[c] = H1 * (T)2 * zeta;
  w.EMB[c] = H2 * (T)0.5 * zeta;
  w.DBP[c] = pb * (T)sc.tens1m * (T)2 * zeta * (T)sc.pr / (T)sc.tens1p;
  // sigma12 at z
  const T s12t = ls12[c] + k.dxV[c] * (w.X[AT(i, jm)] - w.X[c]) + k.dyU[c] * (w.Y[AT(im, j)] - w.Y[c]);
  const T G = s12t * (T)sc.den2;
  ls12o[c] = G * (T)sc.alphaM1;
  const T zz = znorm(g, k, i, j) * ((w.ZETA[c] + w.ZETA[AT(im, j)]) +
                                    (w.ZETA[AT(i, jm)] + w.ZETA[AT(im, jm)]));
  w.ZZB[c] = G * (T)0.5 * w.E12[c];
  w.E12P[c] = G * (T)0.5 * zz;
}

// ---- R3: c-points (zeta, Delta)
template <typename T>
__global__ __launch_bounds__(BLK) void adj_r3_kernel(Geom g, AdjConst<T> k, EvpScalars sc,
                                                     AdjWork<T> w, AdjAcc<T> acc) {
  CELL_SETUP
  const T zb = w.ZBP[c] + znorm(g, k, i, j) * w.ZZB[c] + znorm(g, k, ip, j) * w.ZZB[AT(ip, j)] +
               znorm(g, k, i, jp) * w.ZZB[AT(i, jp)] + znorm(g, k, ip, jp) * w.ZZB[AT(ip, jp)];
  const T dC = w.DELTA[c], zeta = w.ZETA[c];
  const T dreg = dC + (T)sc.deltaMin;
  acc.P[c] += zb * (T)0.5 * (T)sc.tens1p / dreg;
  const T Db = w.DBP[c] - zb * zeta / dreg;
  const T dsqb = dC > (T)0 ? Db * (T)0.5 / dC : (T)0;
  const T ep = w.EP[c], em = w.EM[c];
  const T epb = w.EPB[c] + dsqb * (T)2 * ep;
  const T emb = w.EMB[c] + dsqb * (T)sc.rpd * (T)2 * em;
  const T e12cb = dsqb * (T)sc.rpd * (T)4;
  w.E11B[c] = (epb + emb) * k.rdxU[c] * k.mInC[c];
  w.E22B[c] = (epb - emb) * k.rdyV[c] * k.mInC[c];
  w.QB[c] = e12cb * (T)0.25 * k.rrA[c];
}

// ---- R4: z-points (shear strain)
template <typename T>
__global__ __launch_bounds__(BLK) void adj_r4_kernel(Geom g, AdjConst<T> k, AdjWork<T> w) {
  CELL_SETUP
  const T qb = (w.QB[c] + w.QB[AT(im, j)]) + (w.QB[AT(i, jm)] + w.QB[AT(im, jm)]);
  w.E12B[c] = w.E12P[c] + qb * (T)2 * k.rAz[c] * w.E12[c];
}

// coefficients of e12(z) = A1 u + A2 u[j-1] + A3 v + A4 v[i-1] at z-point (ii, jj)
template <typename T>
__device__ inline void e12_coeffs(const Geom& g, const AdjConst<T>& k, int ii, int jj, T* A1,
                                  T* A2, T* A3, T* A4) {
  const int im = ii == 0 ? g.NX - 1 : ii - 1, jm = jj == 0 ? g.NY - 1 : jj - 1;
  const int c = AT(ii, jj);
  const T mz = (k.M[c] * k.M[AT(im, jj)]) * (k.M[AT(ii, jm)] * k.M[AT(im, jm)]);
  const T hu = k.MU[c] - k.MU[AT(ii, jm)], hv = k.MV[c] - k.MV[AT(im, jj)];
  const T ry = k.rdyU[c], rx = k.rdxV[c];
  *A1 = (T)0.5 * mz * ry + hu * ry;
  *A2 = -(T)0.5 * mz * ry + hu * ry;
  *A3 = (T)0.5 * mz * rx + hv * rx;
  *A4 = -(T)0.5 * mz * rx + hv * rx;
}

// ---- R5: adjoint of the substep's input velocities
template <typename T>
__global__ __launch_bounds__(BLK) void adj_r5_kernel(Geom g, AdjConst<T> k, AdjWork<T> w,
                                                     T* luo, T* lvo) {
  CELL_SETUP
  T A1, A2, A3, A4, B1, B2, B3, B4, C1, C2, C3, C4;
  e12_coeffs(g, k, i, j, &A1, &A2, &A3, &A4);
  e12_coeffs(g, k, i, jp, &B1, &B2, &B3, &B4);
  e12_coeffs(g, k, ip, j, &C1, &C2, &C3, &C4);
  const T fmc = k.fC[c] * k.mC[c];
  luo[c] = w.UDIR[c] + (w.AB[c] + w.AB[AT(im, j)]) * k.mInU[c] + (w.E11B[AT(im, j)] - w.E11B[c]) +
           w.E12B[c] * A1 + w.E12B[AT(i, jp)] * B2 + (T)0.5 * fmc * w.FUB[c] +
           (T)0.5 * k.fC[AT(im, j)] * k.mC[AT(im, j)] * w.FUB[AT(im, j)];
  lvo[c] = w.VDIR[c] + (w.BB[c] + w.BB[AT(i, jm)]) * k.mInV[c] + (w.E22B[AT(i, jm)] - w.E22B[c]) +
           w.E12B[c] * A3 + w.E12B[AT(ip, j)] * C4 + (T)0.5 * fmc * w.FVB[c] +
           (T)0.5 * k.fC[AT(i, jm)] * k.mC[AT(i, jm)] * w.FVB[AT(i, jm)];
}

template <typename T>
__global__ __launch_bounds__(BLK) void copy_kernel(int n, const T* a, T* b) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) b[c] = a ? a[c] : (T)0;
}

// ---- pre-EVP adjoint: strength, mass, wind forcing
template <typename T>
struct PreAdjArgs {
  // forward inputs
  const T *h, *A, *M, *MU, *MV, *pstarF, *u, *v, *uW, *vW, *fC, *rdxC, *rdyC, *ssh, *sp, *AW, *AS;
  // adjoints of the EVP inputs
  const T *aP, *amC, *amU, *amV, *aWFX, *aWFY;
  // outputs
  T *apstar, *ah, *ahs, *aCaN, *aCaS;
  T pstar, cStar, rhoI, rhoS, rhoAir, caN, caS, wmin, wmin2, grav, rrho, sinWin, cosWin;
  int relWind;
};

template <typename T>
__device__ inline void tau_c_and_dca(const Geom& g, const PreAdjArgs<T>& a, int i, int j, T* dtx,
                                     T* dty) {
  const int ip = i == g.NX - 1 ? 0 : i + 1, jp = j == g.NY - 1 ? 0 : j + 1;
  const int c = AT(i, j);
  T urel = a.uW[c], vrel = a.vW[c];
  if (a.relWind) {
    urel = urel - (T)0.5 * (a.u[c] + a.u[AT(ip, j)]);
    vrel = vrel - (T)0.5 * (a.v[c] + a.v[AT(i, jp)]);
  }
  const T wsq = urel * urel + vrel * vrel;
  const T ws = wsq < a.wmin2 ? a.wmin : sqrt(wsq);
  const T sg = sgn(a.fC[c]);
  // d tau / d Ca (Ca of this cell's hemisphere)
  *dtx = (a.rhoAir * ws) * (a.cosWin * urel - (sg * a.sinWin) * vrel);
  *dty = (a.rhoAir * ws) * (a.cosWin * vrel + (sg * a.sinWin) * urel);
}

template <typename T>
__global__ __launch_bounds__(BLK) void pre_evp_adjoint_kernel(Geom g, PreAdjArgs<T> a) {
  CELL_SETUP
  // strength: P = P* h exp(-C*(1-A)) M
  const T E = exp((-a.cStar) * ((T)1 - a.A[c]));
  const T ps = a.pstarF ? a.pstarF[c] : a.pstar;
  a.apstar[c] = a.aP[c] * a.h[c] * E * a.M[c];
  T hb = a.aP[c] * ps * E * a.M[c];
  // wind forcing: tilt term -mU rdxC (phi - phi[i-1]) adds to the mass adjoints
  const T phi = a.grav * a.ssh[c] + a.sp[c] * a.rrho;
  auto phi_at = [&](int ii, int jj) { return a.grav * a.ssh[AT(ii, jj)] + a.sp[AT(ii, jj)] * a.rrho; };
  const T amU_t = a.amU[c] - a.aWFX[c] * a.rdxC[c] * (phi - phi_at(im, j));
  const T amU_ip = a.amU[AT(ip, j)] - a.aWFX[AT(ip, j)] * a.rdxC[AT(ip, j)] * (phi_at(ip, j) - phi);
  const T amV_t = a.amV[c] - a.aWFY[c] * a.rdyC[c] * (phi - phi_at(i, jm));
  const T amV_jp = a.amV[AT(i, jp)] - a.aWFY[AT(i, jp)] * a.rdyC[AT(i, jp)] * (phi_at(i, jp) - phi);
  // mass: mC = rhoI h + rhoS hs, mU = (mC + mC[i-1])/2, mV = (mC + mC[j-1])/2
  const T amC = a.amC[c] + (T)0.5 * (amU_t + amU_ip) + (T)0.5 * (amV_t + amV_jp);
  hb += a.rhoI * amC;
  a.ah[c] = hb;
  a.ahs[c] = a.rhoS * amC;
  // air drag: WFX = 0.5 (tau_c + tau_c[i-1]) MU AW + tilt, tau_c proportional to Ca
  T dtx, dty;
  tau_c_and_dca(g, a, i, j, &dtx, &dty);
  const T gx = (T)0.5 * (a.aWFX[c] * a.AW[c] * a.MU[c] + a.aWFX[AT(ip, j)] * a.AW[AT(ip, j)] * a.MU[AT(ip, j)]);
  const T gy = (T)0.5 * (a.aWFY[c] * a.AS[c] * a.MV[c] + a.aWFY[AT(i, jp)] * a.AS[AT(i, jp)] * a.MV[AT(i, jp)]);
  const T contrib = gx * dtx + gy * dty;
  a.aCaN[c] = a.fC[c] < (T)0 ? (T)0 : contrib;
  a.aCaS[c] = a.fC[c] < (T)0 ? contrib : (T)0;
}

// ---- host side
template <typename T>
struct AdjLayout {
  size_t fb, abb, total;
  int nck, ce;
  size_t off_ck, off_seg, off_work, off_lam, off_ab, off_fwd;
};

template <typename T>
AdjLayout<T> adj_layout(const Geom& g, int nsub, int ce) {
  AdjLayout<T> L;
  L.fb = align256(field_elems(g) * sizeof(T));
  L.abb = align256(sizeof(EvpArgBlock<T>));
  L.ce = ce;
  L.nck = (nsub + ce - 1) / ce;
  L.off_ck = 0;
  L.off_seg = L.off_ck + (size_t)5 * L.nck * L.fb;
  L.off_work = L.off_seg + (size_t)5 * (ce + 1) * L.fb;
  L.off_lam = L.off_work + 28 * L.fb;
  L.off_ab = L.off_lam + 10 * L.fb;
  L.off_fwd = L.off_ab + L.abb;  // workspace of the fused forward launches
  L.total = L.off_fwd + 5 * L.fb + L.abb;  // = the single-domain solver's workspace
  return L;
}

template <typename T>
T* fld(void* ws, size_t off, size_t fb, int q) {
  return (T*)((char*)ws + off + (size_t)q * fb);
}

template <typename T>
AdjConst<T> adj_const(const vh_fields* f) {
  AdjConst<T> k;
  k.M = evp_ptr<T>(f->iceMask);
  k.MU = evp_ptr<T>(f->iceMaskU);
  k.MV = evp_ptr<T>(f->iceMaskV);
  k.mInC = evp_ptr<T>(f->maskInC);
  k.mInU = evp_ptr<T>(f->maskInU);
  k.mInV = evp_ptr<T>(f->maskInV);
  k.rdxU = evp_ptr<T>(f->recip_dxU);
  k.rdyV = evp_ptr<T>(f->recip_dyV);
  k.rdyU = evp_ptr<T>(f->recip_dyU);
  k.rdxV = evp_ptr<T>(f->recip_dxV);
  k.dxU = evp_ptr<T>(f->dxU);
  k.dyU = evp_ptr<T>(f->dyU);
  k.dxV = evp_ptr<T>(f->dxV);
  k.dyV = evp_ptr<T>(f->dyV);
  k.rrAu = evp_ptr<T>(f->recip_rAu);
  k.rrAv = evp_ptr<T>(f->recip_rAv);
  k.rAz = evp_ptr<T>(f->rAz);
  k.rrA = evp_ptr<T>(f->recip_rA);
  k.fC = evp_ptr<T>(f->fCori);
  k.uOc = evp_ptr<T>(f->uOcean);
  k.vOc = evp_ptr<T>(f->vOcean);
  k.P = evp_ptr<T>(f->SeaIceStrength);
  k.mC = evp_ptr<T>(f->SeaIceMassC);
  k.mU = evp_ptr<T>(f->SeaIceMassU);
  k.mV = evp_ptr<T>(f->SeaIceMassV);
  k.AW = evp_ptr<T>(f->AreaW);
  k.AS = evp_ptr<T>(f->AreaS);
  k.uNm1 = evp_ptr<T>(f->uIce);
  k.vNm1 = evp_ptr<T>(f->vIce);
  return k;
}

int check_adj_settings(const vh_domain* d, const vh_settings* s, const vh_fields* f) {
  VH_REQUIRE(d->tile >= 0 && d->tile < kEvpNumTiles && (d->tile == 0 || kEvpTiles[d->tile].k == 1),
             VH_EINVAL, "adjoint: tile %d (needs a single-substep tile shape or 0)", d->tile);
  VH_REQUIRE(s->useAdaptiveEVP == 0.0 && s->noSlip != 0.0 && s->secondOrderBC == 0.0 &&
                 s->useEVP != 0.0,
             VH_EUNSUPPORTED, "adjoint: only non-adaptive noSlip EVP is supported");
  VH_REQUIRE(s->sinWat == 0.0, VH_EUNSUPPORTED, "adjoint: turning angle (sinWat != 0) not supported");
  VH_REQUIRE(s->basalDragK2 == 0.0, VH_EUNSUPPORTED, "adjoint: basal drag not supported");
  VH_REQUIRE(!f->k1AtC && !f->k2AtC && !f->k1AtZ && !f->k2AtZ, VH_EUNSUPPORTED,
             "adjoint: metric terms k1/k2 not supported (pass NULL for all-zero fields)");
  return VH_OK;
}

template <typename T>
int forward_ckpt(const Geom& g, const vh_settings* s, const vh_fields* f, int nsub, int ce,
                 void* const outs[5], void* ws, size_t wsb, int tile, int dtype, hipStream_t st) {
  const AdjLayout<T> L = adj_layout<T>(g, nsub, ce);
  VH_REQUIRE(ws && wsb >= L.total, VH_EWORKSPACE, "adjoint workspace: need %zu bytes, got %zu",
             L.total, wsb);
  const T* in0[5] = {evp_ptr<T>(f->uIce), evp_ptr<T>(f->vIce), evp_ptr<T>(f->sigma1),
                     evp_ptr<T>(f->sigma2), evp_ptr<T>(f->sigma12)};
  const size_t nb = field_elems(g) * sizeof(T);
  for (int q = 0; q < 5; ++q)
    VH_TRY_HIP(hipMemcpyAsync(fld<T>(ws, L.off_ck, L.fb, q), in0[q], nb, hipMemcpyDeviceToDevice, st));
  for (int sgm = 0; sgm < L.nck; ++sgm) {
    const int k0 = sgm * ce, n = std::min(ce, nsub - k0);
    const T* in[5];
    T* out[5];
    for (int q = 0; q < 5; ++q) {
      in[q] = fld<T>(ws, L.off_ck, L.fb, 5 * sgm + q);
      out[q] = (sgm + 1 < L.nck) ? fld<T>(ws, L.off_ck, L.fb, 5 * (sgm + 1) + q) : (T*)outs[q];
    }
    int rc = evp_run_core<T>(g, s, f, in, out, n, (char*)ws + L.off_fwd, L.total - L.off_fwd,
                             tile, dtype, st);
    if (rc) return rc;
  }
  return VH_OK;
}

template <typename T>
int adjoint(const Geom& g, const vh_settings* s, const vh_fields* f, int nsub, int ce,
            const void* const lam_final[5], const vh_evp_adjoint_out* o, void* ws, size_t wsb,
            int tile, int dtype, hipStream_t st) {
  const AdjLayout<T> L = adj_layout<T>(g, nsub, ce);
  VH_REQUIRE(ws && wsb >= L.total, VH_EWORKSPACE, "adjoint workspace: need %zu bytes, got %zu",
             L.total, wsb);
  const int n = g.NX * g.NY;
  const dim3 grid((n + BLK - 1) / BLK);
  AdjConst<T> k = adj_const<T>(f);
  EvpScalars sc;
  fill_evp_scalars(s, 0, 0, &sc);
  AdjWork<T> w;
  T** wp[] = {&w.E12, &w.ZETA, &w.DELTA, &w.EP, &w.EM, &w.CD, &w.X, &w.Y, &w.FX, &w.FY,
              &w.CDU, &w.CDV, &w.UDIR, &w.VDIR, &w.ZBP, &w.EPB, &w.EMB, &w.DBP, &w.AB, &w.BB,
              &w.FUB, &w.FVB, &w.ZZB, &w.E12P, &w.E11B, &w.E22B, &w.QB, &w.E12B};
  for (int q = 0; q < 28; ++q) *wp[q] = fld<T>(ws, L.off_work, L.fb, q);
  AdjAcc<T> acc{(T*)o->adj_SeaIceStrength, (T*)o->adj_SeaIceMassC, (T*)o->adj_SeaIceMassU,
                (T*)o->adj_SeaIceMassV, (T*)o->adj_AreaW, (T*)o->adj_AreaS,
                (T*)o->adj_WindForcingX, (T*)o->adj_WindForcingY, (T*)o->adj_uNm1,
                (T*)o->adj_vNm1, (T*)o->adj_waterDrag_north, (T*)o->adj_waterDrag_south};
  T* accp[] = {acc.P, acc.mC, acc.mU, acc.mV, acc.AW, acc.AS, acc.WFX, acc.WFY, acc.u0, acc.v0,
               acc.DN, acc.DS};
  for (T* p : accp) {
    VH_REQUIRE(p != nullptr, VH_EINVAL, "adjoint: null parameter-adjoint output");
    hipLaunchKernelGGL(copy_kernel<T>, grid, dim3(BLK), 0, st, n, (const T*)nullptr, p);
  }
  // lambda ping-pong: set A holds the adjoint of the substep output
  T* lamA[5];
  T* lamB[5];
  for (int q = 0; q < 5; ++q) {
    lamA[q] = fld<T>(ws, L.off_lam, L.fb, q);
    lamB[q] = fld<T>(ws, L.off_lam, L.fb, 5 + q);
    hipLaunchKernelGGL(copy_kernel<T>, grid, dim3(BLK), 0, st, n, (const T*)lam_final[q], lamA[q]);
  }
  for (int sgm = L.nck - 1; sgm >= 0; --sgm) {
    const int k0 = sgm * ce, len = std::min(ce, nsub - k0);
    // recompute the segment's states seg[0..len] from its checkpoint
    for (int q = 0; q < 5; ++q)
      VH_TRY_HIP(hipMemcpyAsync(fld<T>(ws, L.off_seg, L.fb, q), fld<T>(ws, L.off_ck, L.fb, 5 * sgm + q),
                                field_elems(g) * sizeof(T), hipMemcpyDeviceToDevice, st));
    for (int kk = 0; kk < len; ++kk) {
      const T* in[5];
      T* out[5];
      for (int q = 0; q < 5; ++q) {
        in[q] = fld<T>(ws, L.off_seg, L.fb, 5 * kk + q);
        out[q] = fld<T>(ws, L.off_seg, L.fb, 5 * (kk + 1) + q);
      }
      int rc = evp_run_core<T>(g, s, f, in, out, 1, (char*)ws + L.off_fwd, L.total - L.off_fwd,
                               tile, dtype, st);
      if (rc) return rc;
    }
    // reverse the segment's substeps
    for (int kk = len - 1; kk >= 0; --kk) {
      const T* xu = fld<T>(ws, L.off_seg, L.fb, 5 * kk + 0);
      const T* xv = fld<T>(ws, L.off_seg, L.fb, 5 * kk + 1);
      const T* yu = fld<T>(ws, L.off_seg, L.fb, 5 * (kk + 1) + 0);
      const T* yv = fld<T>(ws, L.off_seg, L.fb, 5 * (kk + 1) + 1);
      hipLaunchKernelGGL(adj_interm_kernel<T>, grid, dim3(BLK), 0, st, g, k, sc, xu, xv, w);
      hipLaunchKernelGGL(adj_r1_kernel<T>, grid, dim3(BLK), 0, st, g, k, sc, xu, xv, yu, yv,
                         (const T*)lamA[0], (const T*)lamA[1], w, acc);
      hipLaunchKernelGGL(adj_r2_kernel<T>, grid, dim3(BLK), 0, st, g, k, sc, xu, xv,
                         (const T*)lamA[2], (const T*)lamA[3], (const T*)lamA[4], lamB[2], lamB[3],
                         lamB[4], w, acc);
      hipLaunchKernelGGL(adj_r3_kernel<T>, grid, dim3(BLK), 0, st, g, k, sc, w, acc);
      hipLaunchKernelGGL(adj_r4_kernel<T>, grid, dim3(BLK), 0, st, g, k, w);
      hipLaunchKernelGGL(adj_r5_kernel<T>, grid, dim3(BLK), 0, st, g, k, w, lamB[0], lamB[1]);
      for (int q = 0; q < 5; ++q) std::swap(lamA[q], lamB[q]);
    }
  }
  T* outs[5] = {(T*)o->adj_uIce, (T*)o->adj_vIce, (T*)o->adj_sigma1, (T*)o->adj_sigma2,
                (T*)o->adj_sigma12};
  for (int q = 0; q < 5; ++q) {
    VH_REQUIRE(outs[q] != nullptr, VH_EINVAL, "adjoint: null state-adjoint output");
    hipLaunchKernelGGL(copy_kernel<T>, grid, dim3(BLK), 0, st, n, (const T*)lamA[q], outs[q]);
  }
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int pre_evp_adjoint(const Geom& g, const vh_settings* s, const vh_fields* f, const void* pstarF,
                    const vh_evp_adjoint_out* in, const vh_pre_evp_adjoint_out* o, hipStream_t st) {
  PreAdjArgs<T> a;
  a.h = evp_ptr<T>(f->hIceMean);
  a.A = evp_ptr<T>(f->Area);
  a.M = evp_ptr<T>(f->iceMask);
  a.MU = evp_ptr<T>(f->iceMaskU);
  a.MV = evp_ptr<T>(f->iceMaskV);
  a.pstarF = evp_ptr<T>(pstarF);
  a.u = evp_ptr<T>(f->uIce);
  a.v = evp_ptr<T>(f->vIce);
  a.uW = evp_ptr<T>(f->uWind);
  a.vW = evp_ptr<T>(f->vWind);
  a.fC = evp_ptr<T>(f->fCori);
  a.rdxC = evp_ptr<T>(f->recip_dxC);
  a.rdyC = evp_ptr<T>(f->recip_dyC);
  a.ssh = evp_ptr<T>(f->ssh_an);
  a.sp = evp_ptr<T>(f->surfPress);
  a.AW = evp_ptr<T>(f->AreaW);
  a.AS = evp_ptr<T>(f->AreaS);
  a.aP = evp_ptr<T>(in->adj_SeaIceStrength);
  a.amC = evp_ptr<T>(in->adj_SeaIceMassC);
  a.amU = evp_ptr<T>(in->adj_SeaIceMassU);
  a.amV = evp_ptr<T>(in->adj_SeaIceMassV);
  a.aWFX = evp_ptr<T>(in->adj_WindForcingX);
  a.aWFY = evp_ptr<T>(in->adj_WindForcingY);
  a.apstar = (T*)o->adj_pStar;
  a.ah = (T*)o->adj_hIceMean;
  a.ahs = (T*)o->adj_hSnowMean;
  a.aCaN = (T*)o->adj_airDrag_north;
  a.aCaS = (T*)o->adj_airDrag_south;
  const void* req[] = {a.h, a.A, a.M, a.MU, a.MV, a.u, a.v, a.uW, a.vW, a.fC, a.rdxC, a.rdyC,
                       a.ssh, a.sp, a.AW, a.AS, a.aP, a.amC, a.amU, a.amV, a.aWFX, a.aWFY,
                       a.apstar, a.ah, a.ahs, a.aCaN, a.aCaS};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "pre_evp_adjoint: null pointer");
  VH_REQUIRE(s->useRealFreshWaterFlux == 0.0, VH_EUNSUPPORTED,
             "pre_evp_adjoint: useRealFreshWaterFlux not supported");
  a.pstar = (T)s->pStar;
  a.cStar = (T)s->cStar;
  a.rhoI = (T)s->rhoIce;
  a.rhoS = (T)s->rhoSnow;
  a.rhoAir = (T)s->rhoAir;
  a.caN = (T)s->airIceDrag;
  a.caS = (T)s->airIceDrag_south;
  a.wmin = (T)s->wSpeedMin;
  a.wmin2 = (T)(s->wSpeedMin * s->wSpeedMin);
  a.grav = (T)s->gravity;
  a.rrho = (T)s->recip_rhoSea;
  const double ang = s->airTurnAngle * (M_PI / 180.0);
  a.sinWin = (T)sin(ang);
  a.cosWin = (T)cos(ang);
  a.relWind = s->useRelativeWind != 0.0;
  const int n = g.NX * g.NY;
  hipLaunchKernelGGL(pre_evp_adjoint_kernel<T>, dim3((n + BLK - 1) / BLK), dim3(BLK), 0, st, g, a);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

#undef CELL_SETUP
#undef AT

}  // namespace
}  // namespace vh

extern "C" int vh_evp_adjoint_workspace_bytes(const vh_domain* d, int32_t n_substeps,
                                              int32_t ckpt_every, size_t* bytes) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(bytes && n_substeps >= 1 && ckpt_every >= 1, VH_EINVAL, "bad arguments");
  *bytes = d->dtype == VH_F64 ? vh::adj_layout<double>(g, n_substeps, ckpt_every).total
                              : vh::adj_layout<float>(g, n_substeps, ckpt_every).total;
  return VH_OK;
}

extern "C" int vh_evp_forward_checkpointed(const vh_domain* d, const vh_settings* s,
                                           const vh_fields* f, int32_t n_substeps,
                                           int32_t ckpt_every, void* uIce, void* vIce,
                                           void* sigma1, void* sigma2, void* sigma12,
                                           void* workspace, size_t workspace_bytes, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(d && s && f && n_substeps >= 1 && ckpt_every >= 1, VH_EINVAL, "bad arguments");
  rc = vh::check_adj_settings(d, s, f);
  if (rc) return rc;
  void* outs[5] = {uIce, vIce, sigma1, sigma2, sigma12};
  for (void* p : outs) VH_REQUIRE(p != nullptr, VH_EINVAL, "null output");
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64)
    return vh::forward_ckpt<double>(g, s, f, n_substeps, ckpt_every, outs, workspace,
                                    workspace_bytes, d->tile, d->dtype, st);
  return vh::forward_ckpt<float>(g, s, f, n_substeps, ckpt_every, outs, workspace,
                                 workspace_bytes, d->tile, d->dtype, st);
}

extern "C" int vh_evp_adjoint(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                              int32_t n_substeps, int32_t ckpt_every, const void* adj_uIce,
                              const void* adj_vIce, const void* adj_sigma1, const void* adj_sigma2,
                              const void* adj_sigma12, const vh_evp_adjoint_out* out,
                              void* workspace, size_t workspace_bytes, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(s && f && out && n_substeps >= 1 && ckpt_every >= 1, VH_EINVAL, "bad arguments");
  rc = vh::check_adj_settings(d, s, f);
  if (rc) return rc;
  const void* lam[5] = {adj_uIce, adj_vIce, adj_sigma1, adj_sigma2, adj_sigma12};
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64)
    return vh::adjoint<double>(g, s, f, n_substeps, ckpt_every, lam, out, workspace,
                               workspace_bytes, d->tile, d->dtype, st);
  return vh::adjoint<float>(g, s, f, n_substeps, ckpt_every, lam, out, workspace, workspace_bytes,
                            d->tile, d->dtype, st);
}

extern "C" int vh_pre_evp_adjoint(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                                  const void* pStarField, const vh_evp_adjoint_out* evp_adj,
                                  const vh_pre_evp_adjoint_out* out, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(s && f && evp_adj && out, VH_EINVAL, "bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64) return vh::pre_evp_adjoint<double>(g, s, f, pStarField, evp_adj, out, st);
  return vh::pre_evp_adjoint<float>(g, s, f, pStarField, evp_adj, out, st);
}
