// Fused EVP substep kernel template (gfx950): K substeps of
// veris/evp_solver.py:22-254 (non-adaptive, noSlip) per launch, shared by the
// single-domain drop-in (evp.hip, FP64/FP32 in the reference's operation
// order) and the batched ensemble path (evp_batched_fast.hip).
//
// Work decomposition.  A workgroup of RI x RJ threads owns a TI x TJ tile
// (TI = RI - 4K, TJ = RJ - 4K, j contiguous) plus a ring of 2K cells, ONE
// THREAD PER CELL.  Every thread issues all global loads its cell needs at
// once (at 1 degree the grid is ~1 wave per SIMD, so memory latency, not
// bandwidth, sets the time) and keeps its velocity and stresses in registers
// across the K substeps.  Each substep runs in LDS-exchange stages separated
// by barriers:
//   A  u, v, masks, u-uOcean, v-vOcean -> LDS
//   B  e12 and q = rAz*e12^2 on z-points
//   C  e11, e22, Delta, zeta, press on c-points; sigma1, sigma2, water and
//      basal drag, Coriolis terms
//   D  zeta at z-points, sigma12
//   E  stress divergence, forcing, denominators, new u and v
// Substep s is valid on the tile grown by 2(K-1-s) cells, so the ring shrinks
// by two cells per substep and the last one covers exactly the tile.  Ring
// cells are recomputed by every tile that needs them; they depend only on the
// state at the start of the launch, so the state ping-pongs between two buffer
// sets and no workgroup reads another's output within a launch.
//
// Neighbour indexing.  K = 1 wraps over the padded (NX, NY) array exactly as
// numpy.roll does in the reference, and mirrors new velocities into the
// periodic overlap (the reference's fill_overlap); the result equals the
// reference on every cell.  K > 1 cannot mirror between its in-register
// substeps, so in a periodically filled direction it indexes the interior
// periodically (overlap cell i is read as interior cell i +- n), which is what
// the fill makes the overlap hold; non-filled directions keep the roll wrap.
// That reproduces velocities on every cell and stresses everywhere except in
// the filled overlap columns/rows whose reference values are roll wrap-around
// artefacts no other value depends on (there it returns the periodic value).
//
// TD is the type of the Delta-regularised viscosity (Delta, zeta, press): the
// FP32 path keeps it in FP64.
#pragma once

#include "vh_common.h"

namespace vh {

template <typename T>
struct EvpConst {
  const T *uNm1, *vNm1, *P, *massC, *massU, *massV, *AreaW, *AreaS, *WFX, *WFY, *uOc, *vOc, *fC;
  const T *M, *MU, *MV, *mInC, *mInU, *mInV;
  const T *rdxU, *rdyV, *rdyU, *rdxV, *dxU, *dyU, *dxV, *dyV, *rrAu, *rrAv, *rAz, *rrA;
  const T *k1C, *k2C, *k1Z, *k2Z;  // metric terms (read when kterms)
  const T *Rlow, *Area, *hIce;     // basal drag inputs (read when basal)
};

struct EvpScalars {
  double rpd;             // 1 / PlasDefCoeff**2
  double deltaMin;
  double tens1p, tens1m;  // 1 + tensileStrFac, 1 - tensileStrFac
  double pr, pr1m;        // pressReplFac, 1 - pressReplFac
  double alphaM1;         // evpAlpha - evpRevFac (evpRevFac = 1)
  double den1, den2;      // 1 / evpAlpha
  double cosWat, sinWat;
  double cwN, cwS, rhoSea, cDragMin, cDragMin2;
  double betaU, betaP1, rdt;  // evpBeta/deltatDyn, betaU + 1/deltatDyn, 1/deltatDyn
  double K2, U0sq, K1, cBasalStar;
  int kterms, turn, basal, pad;
};

// Per-member geometry of a batched launch: element offset of the member's
// padded box inside every packed field, and its extents.
struct MemberDesc {
  long long off;
  Geom g;
  int pad;
};

// One tile of a batched launch.
struct TileDesc {
  int member, i0, j0, pad;
};

// Everything a launch reads, in device memory (written once per solver call,
// so a captured graph of launches stays valid when the caller's buffers move).
template <typename T>
struct EvpArgBlock {
  EvpConst<T> c;
  const T* in[5];  // u, v, sigma1, sigma2, sigma12 at the start of the subcycle
  T* out[5];       // final state (written by the last launch)
  T* alt[5];       // ping-pong partner
  EvpScalars sc;
  Geom g;          // single-domain geometry (unused when batched)
  int nlaunch;     // launches in the subcycle
  int pad;
  const MemberDesc* members;  // batched only
  const TileDesc* tiles;      // batched only
};

template <typename T>
__global__ void evp_write_args_kernel(EvpArgBlock<T>* dst, EvpArgBlock<T> src) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = src;
}

// padded index of region coordinate p along an axis of padded length N,
// interior length n, overlap 2: roll wrap, or periodic interior if filled
template <int K>
__device__ inline int map_axis(int p, int N, int n, int filled) {
  if (K > 1 && filled) return 2 + wrap(p - 2, n);
  return wrap(p, N);
}

template <typename T, typename TD, int RI, int RJ, int K, bool BATCHED>
__global__ __launch_bounds__(RI* RJ) void evp_substep_kernel(const EvpArgBlock<T>* __restrict__ ab,
                                                             int launch, int n_tiles, int tiles_j,
                                                             int per_xcd) {
  constexpr int R = 2 * K;
  constexpr int TI = RI - 2 * R, TJ = RJ - 2 * R, NT = RI * RJ;
  static_assert((RJ & (RJ - 1)) == 0, "RJ must be a power of two");
  static_assert(TI >= 1 && TJ >= 1, "tile too small for K");
  __shared__ T sU[NT], sV[NT], sM[NT], sMU[NT], sMV[NT], sDU[NT], sDV[NT], sE12[NT], sQ[NT], sZ[NT];
  __shared__ T sS11[NT], sS22[NT], sCD[NT], sCB[NT], sFV[NT], sFU[NT];

  // tile of this workgroup: workgroups are dealt round-robin to the 8 XCDs,
  // so give each XCD a contiguous run of tiles (neighbours share L2 lines)
  const int bid = blockIdx.x;
  const int tile = (bid & 7) * per_xcd + (bid >> 3);
  if (tile >= n_tiles) return;
  Geom g;
  long long off = 0;
  int i0, j0;
  if (BATCHED) {
    const TileDesc td = ab->tiles[tile];
    const MemberDesc md = ab->members[td.member];
    g = md.g;
    off = md.off;
    i0 = td.i0;
    j0 = td.j0;
  } else {
    g = ab->g;
    const int ti = tile / tiles_j;
    i0 = ti * TI;
    j0 = (tile - ti * tiles_j) * TJ;
  }
  const int nl = ab->nlaunch;
  const T* const* src = (launch == 0) ? ab->in
                                      : (((nl - launch) % 2 == 0) ? (const T* const*)ab->out
                                                                  : (const T* const*)ab->alt);
  T* const* dst = ((nl - 1 - launch) % 2 == 0) ? ab->out : ab->alt;
  const EvpConst<T>& f = ab->c;
  const EvpScalars& sc = ab->sc;

  const int t = threadIdx.x;
  const int ar = t / RJ, br = t & (RJ - 1);  // region coordinates
  const int i = i0 + ar - R, j = j0 + br - R;  // padded position
  const int gi = map_axis<K>(i, g.NX, g.nx, g.fill_x);
  const int gj = map_axis<K>(j, g.NY, g.ny, g.fill_y);
  const long long c = off + (long long)gi * g.NY + gj;

  // load predicates: union over the substeps (substep 0 has the widest sets)
  const int a0 = ar - 2, b0 = br - 2;  // coordinates relative to the substep-0 tile
  const int TI0 = RI - 4, TJ0 = RJ - 4;
  const bool zc0 = (a0 >= -1) && (b0 >= -1);
  const bool cc0 = zc0 && (a0 <= TI0) && (b0 <= TJ0);
  const bool cs0 = zc0 && (a0 < TI0) && (b0 < TJ0);
  const bool zs0 = (a0 >= 0) && (a0 <= TI0) && (b0 >= 0) && (b0 <= TJ0);
  const bool up0 = (a0 >= 0) && (a0 < TI0) && (b0 >= 0) && (b0 < TJ0);
  // final output cell of this thread
  const bool own = (ar >= R) && (ar < R + TI) && (br >= R) && (br < R + TJ) && (i < g.NX) &&
                   (j < g.NY);

  // ---- all global loads of this cell, issued together
  T u = src[0][c], v = src[1][c];
  const T Mc = f.M[c], MUc = f.MU[c], MVc = f.MV[c];
  const T uoc = f.uOc[c], voc = f.vOc[c];
  const T mInUc = f.mInU[c], mInVc = f.mInV[c];
  T rdyU = 0, rdxV = 0, rAz = 0, k1Z = 0, k2Z = 0;
  if (zc0) {
    rdyU = f.rdyU[c];
    rdxV = f.rdxV[c];
    rAz = f.rAz[c];
    if (sc.kterms) {
      k1Z = f.k1Z[c];
      k2Z = f.k2Z[c];
    }
  }
  T rdxU = 0, rdyV = 0, mInC = 0, rrA = 0, P = 0, k1C = 0, k2C = 0;
  if (cc0) {
    rdxU = f.rdxU[c];
    rdyV = f.rdyV[c];
    mInC = f.mInC[c];
    rrA = f.rrA[c];
    P = f.P[c];
    if (sc.kterms) {
      k1C = f.k1C[c];
      k2C = f.k2C[c];
    }
  }
  T s1 = 0, s2 = 0, dyV = 0, dxU = 0, fc = 0, mC = 0;
  if (cs0) {
    s1 = src[2][c];
    s2 = src[3][c];
    dyV = f.dyV[c];
    dxU = f.dxU[c];
    fc = f.fC[c];
    mC = f.massC[c];
  }
  T s12 = 0, dxV = 0, dyU = 0;
  if (zs0) {
    s12 = src[4][c];
    dxV = f.dxV[c];
    dyU = f.dyU[c];
  }
  const bool upd_any = up0 && (K > 1 || is_computed(g, i, j));
  T rrAu = 0, rrAv = 0, mU = 0, mV = 0, AW = 0, AS = 0, wfx = 0, wfy = 0, uNm1 = 0, vNm1 = 0;
  if (upd_any) {
    rrAu = f.rrAu[c];
    rrAv = f.rrAv[c];
    mU = f.massU[c];
    mV = f.massV[c];
    AW = f.AreaW[c];
    AS = f.AreaS[c];
    wfx = f.WFX[c];
    wfy = f.WFY[c];
    uNm1 = f.uNm1[c];
    vNm1 = f.vNm1[c];
  }
  T cbc = (T)0;  // basal drag coefficient inputs are step constants: load lazily
  const T sg = sgn(fc);

#pragma unroll 1
  for (int s = 0; s < K; ++s) {
    // coordinates relative to the substep-s tile, and its extent
    const int a = ar - 2 - 2 * s, b = br - 2 - 2 * s;
    const int TIs = RI - 4 - 4 * s, TJs = RJ - 4 - 4 * s;
    const bool zc = (a >= -1) && (b >= -1) && (a < TIs + 2) && (b < TJs + 2);
    const bool cc = (a >= -1) && (b >= -1) && (a <= TIs) && (b <= TJs);
    const bool cs = (a >= -1) && (b >= -1) && (a < TIs) && (b < TJs);
    const bool zs = (a >= 0) && (b >= 0) && (a <= TIs) && (b <= TJs);
    const bool upd = (a >= 0) && (b >= 0) && (a < TIs) && (b < TJs) && upd_any;

    // ---- A
    if (s > 0) __syncthreads();  // E of the previous substep still reads LDS
    sU[t] = u;
    sV[t] = v;
    sM[t] = Mc;
    sMU[t] = MUc;
    sMV[t] = MVc;
    sDU[t] = (u - uoc) * mInUc;
    sDV[t] = (v - voc) * mInVc;
    __syncthreads();

    // ---- B: strain rate at z-points (veris/dynamics_routines.py:160-175)
    if (zc) {
      const T ujm = sU[t - 1], vim = sV[t - RJ];
      const T dudy = (u - ujm) * rdyU;
      const T uavz = (u + ujm) * (T)0.5;
      const T dvdx = (v - vim) * rdxV;
      const T vavz = (v + vim) * (T)0.5;
      const T mz = (Mc * sM[t - RJ]) * (sM[t - 1] * sM[t - RJ - 1]);
      T inner = dudy + dvdx;
      if (sc.kterms) inner = (inner - k1Z * vavz) - k2Z * uavz;
      T e12 = ((T)0.5 * inner) * mz;
      const T hu = MUc - sMU[t - 1];
      const T hv = MVc - sMV[t - RJ];
      e12 = e12 + ((((T)2.0 * uavz) * rdyU) * hu + (((T)2.0 * vavz) * rdxV) * hv);
      sE12[t] = e12;
      sQ[t] = rAz * (e12 * e12);
    }
    __syncthreads();

    // ---- C: c-point rheology (veris/dynamics_routines.py:151-158, 220-268),
    //         stresses (veris/evp_solver.py:79-94), drags (:108-110), Coriolis (:142-143)
    if (cc) {
      const T uip = sU[t + RJ], vjp = sV[t + 1];
      const T dudx = (uip - u) * rdxU;
      const T uave = (uip + u) * (T)0.5;
      const T dvdy = (vjp - v) * rdyV;
      const T vave = (vjp + v) * (T)0.5;
      T e11 = dudx, e22 = dvdy;
      if (sc.kterms) {
        e11 = dudx + vave * k2C;
        e22 = dvdy + uave * k1C;
      }
      e11 = e11 * mInC;
      e22 = e22 * mInC;
      const T qa = sQ[t] + sQ[t + RJ];
      const T qb = sQ[t + 1] + sQ[t + RJ + 1];
      const T e12c = ((T)0.25 * rrA) * (qa + qb);
      const T ep = e11 + e22, em = e11 - e22;
      const TD dsq = (TD)ep * (TD)ep + (TD)sc.rpd * ((TD)em * (TD)em + (TD)4.0 * (TD)e12c);
      const TD dC = sqrt(dsq);
      const TD dreg = dC + (TD)sc.deltaMin;
      const TD zeta = ((TD)0.5 * ((TD)P * (TD)sc.tens1p)) / dreg;
      sZ[t] = (T)zeta;
      if (cs) {
        const TD pressD =
            (((TD)P * (TD)sc.pr1m) + ((((TD)2.0 * zeta) * dC) * (TD)sc.pr) / (TD)sc.tens1p) *
            (TD)sc.tens1m;
        const T zt = (T)zeta, press = (T)pressD;
        const T divg = (((T)2 * zt) * ep) - press;
        const T tens = ((T)2 * zt) * em;
        s1 = (((s1 * (T)sc.alphaM1) + divg) * (T)sc.den1) * Mc;
        s2 = (((s2 * (T)sc.alphaM1) + tens * (T)0.25) * (T)sc.den2) * Mc;
        sS11[t] = ((T)0.5 * (s1 + s2)) * dyV;
        sS22[t] = ((T)0.5 * (s1 - s2)) * dxU;
        const T D = (T)(fc < (T)0 ? sc.cwS : sc.cwN) * (T)sc.rhoSea;
        const T su = sDU[t] + sDU[t + RJ], sv = sDV[t] + sDV[t + 1];
        const T w = (T)0.25 * (su * su + sv * sv);
        const T cd = ((D * D) * w > (T)sc.cDragMin2) ? D * sqrt(w) : (T)sc.cDragMin;
        sCD[t] = cd * Mc;
        if (sc.basal) {
          const long long cip =
              off + (long long)map_axis<K>(i + 1, g.NX, g.nx, g.fill_x) * g.NY + gj;
          const long long cjp =
              off + (long long)gi * g.NY + map_axis<K>(j + 1, g.NY, g.ny, g.fill_y);
          const T um = u * mInUc, umip = uip * f.mInU[cip];
          const T vm = v * mInVc, vmjp = vjp * f.mInV[cjp];
          T tb = (T)0.25 * (((um * um + umip * umip) + vm * vm) + vmjp * vmjp);
          tb = (T)sc.K2 / sqrt(tb + (T)sc.U0sq);
          const T Ar = f.Area[c];
          const T hc = (fabs(f.Rlow[c]) * Ar) / (T)sc.K1;
          const T val = ((tb * log(exp((T)10.0 * (f.hIce[c] - hc)) + (T)1.0)) * (T)0.1) *
                        exp((T)(-sc.cBasalStar) * ((T)1.0 - Ar));
          cbc = Ar > (T)0.01 ? val : (T)0;
        }
        sCB[t] = cbc;
        const T mf = (mC * fc) * (T)0.5;
        sFV[t] = mf * (v + vjp);
        sFU[t] = mf * (u + uip);
      }
    }
    if (sc.turn) {  // ocean velocity for the turning term (sMU/sMV are free after B)
      sMU[t] = uoc;
      sMV[t] = voc;
    }
    __syncthreads();

    // ---- D: sigma12 on z-points (veris/averaging.py:4-19, veris/evp_solver.py:86,103)
    if (zs) {
      const T n = (Mc + sM[t - RJ]) + (sM[t - 1] + sM[t - RJ - 1]);
      const T nrm = n > (T)0 ? (T)1.0 / n : (T)0.0;
      const T zz = nrm * ((sZ[t] + sZ[t - RJ]) + (sZ[t - 1] + sZ[t - RJ - 1]));
      const T shear = ((T)2 * zz) * sE12[t];
      s12 = ((s12 * (T)sc.alphaM1) + shear * (T)0.25) * (T)sc.den2;
      sQ[t] = s12 * dxV;   // sigma12*dxV (q is dead after C)
      sDU[t] = s12 * dyU;  // sigma12*dyU (du is dead after C)
    }
    __syncthreads();

    // ---- E: stress divergence (veris/dynamics_routines.py:284-302), forcing
    //         (veris/evp_solver.py:112-145) and momentum update (:152-187)
    if (upd) {
      const T divx = (((sS11[t] - sS11[t - RJ]) - sQ[t]) + sQ[t + 1]) * rrAu;
      const T divy = (((sS22[t] - sS22[t - 1]) - sDU[t]) + sDU[t + RJ]) * rrAv;
      const T cdc = sCD[t];
      const T cdU = (T)0.5 * (cdc + sCD[t - RJ]);
      const T cdV = (T)0.5 * (cdc + sCD[t - 1]);
      T inx = (cdU * (T)sc.cosWat) * uoc;
      T iny = (cdV * (T)sc.cosWat) * voc;
      if (sc.turn) {
        const T lmU = mU != (T)0 ? (T)1 : mU;
        const T lmV = mV != (T)0 ? (T)1 : mV;
        const T X = (cdc * (T)0.5) * ((uoc - u) + (sMU[t + 1] - sU[t + 1])) +
                    (sCD[t - RJ] * (T)0.5) *
                        ((sMU[t - RJ] - sU[t - RJ]) + (sMU[t - RJ + 1] - sU[t - RJ + 1]));
        const T Y = (cdc * (T)0.5) * ((voc - v) + (sMV[t + RJ] - sV[t + RJ])) +
                    (sCD[t - 1] * (T)0.5) *
                        ((sMV[t - 1] - sV[t - 1]) + (sMV[t + RJ - 1] - sV[t + RJ - 1]));
        inx = inx - (((sg * (T)sc.sinWat) * (T)0.5) * X) * lmU;
        iny = iny + (((sg * (T)sc.sinWat) * (T)0.5) * Y) * lmV;
      }
      T fx = wfx + inx * AW;
      T fy = wfy + iny * AS;
      fx = fx + (T)0.5 * (sFV[t] + sFV[t - RJ]);
      fy = fy - (T)0.5 * (sFU[t] + sFU[t - 1]);
      T denU = mU * (T)sc.betaP1 + AW * ((cdU * (T)sc.cosWat) + (T)0.5 * (sCB[t] + sCB[t - RJ]));
      T denV = mV * (T)sc.betaP1 + AS * ((cdV * (T)sc.cosWat) + (T)0.5 * (sCB[t] + sCB[t - 1]));
      denU = denU == (T)0 ? (T)1 : denU;
      denV = denV == (T)0 ? (T)1 : denV;
      const T un =
          (MUc * ((((T)sc.betaU * mU) * u + (mU * (T)sc.rdt) * uNm1 + fx) + divx)) / denU;
      const T vn =
          (MVc * ((((T)sc.betaU * mV) * v + (mV * (T)sc.rdt) * vNm1 + fy) + divy)) / denV;
      u = un;  // other threads read sU/sV, never these registers
      v = vn;
    }
  }

  // ---- results of the launch, at the cell's own padded position
  if (own) {
    const long long cp = off + (long long)i * g.NY + j;
    if (K == 1) {
      if (upd_any) {
        store_filled(g, dst[0] + off, i, j, u);
        store_filled(g, dst[1] + off, i, j, v);
      }
    } else {
      dst[0][cp] = u;
      dst[1][cp] = v;
    }
    dst[2][cp] = s1;
    dst[3][cp] = s2;
    dst[4][cp] = s12;
  }
}

}  // namespace vh
