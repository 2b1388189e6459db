// Thermodynamic growth (veris/growth.py:8-397) with the surface energy
// balance solve4temp (veris/solve4temp.py:5-174): pointwise, one thread per
// padded cell, all nITC thickness categories and their Newton iterations in
// registers (the kernel is bound by the arithmetic of the flux evaluations;
// a variant with one wave per category was no faster).  Built
// with -ffp-contract=off; expressions keep the reference's association order
// (exp/pow come from the device math library, so results agree with an IEEE
// host evaluation to a few ulp, not bit for bit).
#include <cmath>

#include "vh_common.h"

namespace vh {
namespace {

constexpr int BLK = 256;

struct GrowthScalars {
  double qi, qs;
  double Area_reg, hIce_reg, recip_nITC, nITC_d;
  int nITC;
  double tempFrz, dtempFrz_dS, celsius2K;
  double snowRate_fac;  // rhoFresh2rhoSnow
  double dtT, rdtT;
  double t0, t1, t2;    // mixed-layer constants 0.4, 7/0.4, stantonNr*uStarBase*rhoSea*cpWater
  double McPheeTaperFac, recip_h0, recip_h0_south;
  double rhoSnow, rhoIce, recip_rhoSea, rhoIce2rhoSnow, rhoIce2rhoFresh, rhoFresh,
      rhoFresh2rhoSnow, saltIce_ref;
  // solve4temp
  double d1, d1i, Tmelt, Tmeltw, minLWdown, minTAir_K, wSpeedMin, snowEmiss, iceEmiss, stefBoltz;
  double dryIceAlb, drySnowAlb, wetIceAlb, wetSnowAlb, dryIceAlb_south, drySnowAlb_south,
      wetIceAlb_south, wetSnowAlb_south, hCut, shortwave, kIS /*iceConduct*snowConduct*/,
      snowConduct, iceConduct, minTIce_K, lhSublim;
  double aa1, aa2, p10aa2 /*10^aa2*/, bb1, bb1m /*1-bb1*/, ppa, cc1, cc2;
};

template <typename T>
struct GrowthIO {
  const T *h, *hs, *A, *Ts, *ocSalt, *theta, *snowfall, *precip, *Qnet, *Qsw, *evap, *runoff,
      *osh, *oshs, *M, *fC, *LWdown, *ATemp, *wSpeed, *aqh, *SWdown;
  T *oh, *ohs, *oA, *oTs, *oEmPmR, *ofss, *oQsw, *oQnet, *oload, *open, *orh;
};

template <typename T>
struct Fl {
  T Fc, Flh, Fia, dF;
};

template <typename T>
__device__ inline Fl<T> fluxes(const GrowthScalars& k, T t1, bool ice, T kEff, T Tfrz, T ug,
                               T aqh, T d3, T lw, T absSW, T ta) {
  Fl<T> r;
  const T t2 = t1 * t1;
  const T t3 = t2 * t1;
  const T t4 = t2 * t2;
  // c3 = 10^(aa1/t1) (veris/solve4temp.py:120) and the saturation vapour
  // pressure 10^(-aa1/t1 + aa2) (:114) as 10^aa2 / c3: one power of ten per
  // evaluation instead of two (the powers were almost all of the kernel's
  // arithmetic); within a few ulp of the reference's two powers
  const T c3 = exp10((T)k.aa1 / t1);
  const T svp = (T)k.p10aa2 / c3;
  const T qsurf = ice ? ((T)k.bb1 * svp) / ((T)k.ppa - (T)k.bb1m * svp) : (T)0;
  const T cd = (T)k.cc2 - c3 * (T)k.ppa;
  const T dq = ice ? ((T)k.cc1 * c3) / ((cd * cd) * t2) : (T)0;
  r.Fc = ice ? kEff * (Tfrz - t1) : (T)0;
  r.Flh = ice ? ((T)k.d1i * ug) * (qsurf - aqh) : (T)0;
  const T Flw = ice ? t4 * d3 : (T)0;
  const T Fs = ice ? ((T)k.d1 * ug) * (t1 - ta) : (T)0;
  r.Fia = ice ? ((((-lw) - absSW) + Flw) + Fs) + r.Flh : (T)0;
  r.dF = ice ? (((T)4 * d3) * t3 + (T)k.d1 * ug) + ((T)k.d1i * ug) * dq : (T)0;
  return r;
}

// returns T_out; writes F_io_net, F_ia_net, IcePenetSW of one category
template <typename T>
__device__ inline T solve4temp(const GrowthScalars& k, T hI, T hS, T Tin, T Tfrz, T fc, T LWd,
                               T ATemp, T wSpeed, T aqh, T SWd, T* Fio, T* Fia, T* penSW) {
  T Tl = Tin;
  const T lwc = fmax((T)k.minLWdown, LWd);
  const T ta = fmax((T)k.minTAir_K, ATemp);
  const T ug = fmax((T)k.wSpeedMin, wSpeed);
  const bool ice = hI > (T)0;
  const bool snow = hS > (T)0;
  const T emis = snow ? (T)k.snowEmiss : (T)k.iceEmiss;
  const T d3 = emis * (T)k.stefBoltz;
  const T lw = emis * lwc;
  T aI = ice ? (T)k.dryIceAlb : (T)0;
  T aS = ice ? (T)k.drySnowAlb : (T)0;
  const bool wet = (hI > (T)0) && (Tl >= (T)k.Tmeltw);
  if (wet) {
    aI = (T)k.wetIceAlb;
    aS = (T)k.wetSnowAlb;
  }
  const bool south = (hI > (T)0) && (fc < (T)0);
  if (south) {
    aI = (T)k.dryIceAlb_south;
    aS = (T)k.drySnowAlb_south;
  }
  if (south && (Tl >= (T)k.Tmeltw)) {
    aI = (T)k.wetIceAlb_south;
    aS = (T)k.wetSnowAlb_south;
  }
  T alb = ice ? aI + (hS / (T)k.hCut) * (aS - aI) : (T)0;
  if (hS > (T)k.hCut) alb = aS;
  if (hS == (T)0) alb = aI;
  T pen = ice ? (T)k.shortwave * exp((T)(-1.5) * hI) : (T)0;
  if (snow) pen = (T)0;
  *penSW = ice ? ((-((T)1 - alb)) * pen) * SWd : (T)0;
  const T absSW = ice ? (((T)1 - alb) * ((T)1 - pen)) * SWd : (T)0;
  const T kEff = ice ? (T)k.kIS / ((T)k.snowConduct * hI + (T)k.iceConduct * hS) : (T)0;
  for (int it = 0; it < 6; ++it) {
    const Fl<T> r = fluxes(k, Tl, ice, kEff, Tfrz, ug, aqh, d3, lw, absSW, ta);
    Tl = ice ? Tl + (r.Fc - r.Fia) / (kEff + r.dF) : (T)0;
    Tl = fmin(Tl, (T)k.Tmelt);
    Tl = fmax(Tl, (T)k.minTIce_K);
  }
  const Fl<T> r = fluxes(k, Tl, ice, kEff, Tfrz, ug, aqh, d3, lw, absSW, ta);
  const bool up = r.Fc > (T)0;
  *Fio = up ? r.Fc : (T)0;
  *Fia = up ? (T)0 : r.Fia;
  return ice ? Tl : Tin;
}

// Storage type T, arithmetic in C = double for both storage types: the
// solve4temp flux derivative (cc1 c3 / (cd^2 T^2)) exceeds the FP32 range at
// the minTIce clamp.
//
// A cell's work splits into the inputs of its categories (cell_in), the
// surface energy balance of each category (category: six Newton iterations,
// two powers of ten per flux evaluation, almost all of a cell's arithmetic)
// and the budgets that combine them (tail).
template <typename C>
struct CellIn {
  C hP, hsP, AP, rregA, rhIA, hIA, hSA, sal, Tfrz, Ts0, fc, LWd, ATm, wsp, aqh, SWd;
};

template <typename T, typename C>
__device__ inline CellIn<C> cell_in(const GrowthScalars& k, const GrowthIO<T>& io, size_t c) {
  CellIn<C> x;
  const C h0 = io.h[c], hs0 = io.hs[c], A0 = io.A[c];
  const bool noIce0 = (h0 == (C)0) || (A0 == (C)0);
  const C hP = h0 * (C)(!noIce0);
  const C hsP = hs0 * (C)(!noIce0);
  const C AP = A0 * (C)(!noIce0);
  const bool isIce = hP > (C)0;
  const C regA = sqrt(AP * AP + (C)k.Area_reg);
  const C rregA = (C)1 / regA;
  C hIA = isIce ? hP * rregA : (C)0;
  const C rhIA = AP / sqrt(hP * hP + (C)k.hIce_reg);
  const C hSA = isIce ? hsP * rregA : (C)0;
  hIA = fmax(hIA, (C)0.05);
  const C sal = io.ocSalt[c];
  const C Tfrz = ((C)k.tempFrz + (C)k.dtempFrz_dS * sal) + (C)k.celsius2K;
  x.hP = hP;
  x.hsP = hsP;
  x.AP = AP;
  x.rregA = rregA;
  x.rhIA = rhIA;
  x.hIA = hIA;
  x.hSA = hSA;
  x.sal = sal;
  x.Tfrz = Tfrz;
  x.Ts0 = io.Ts[c];
  x.fc = io.fC[c];
  x.LWd = io.LWdown[c];
  x.ATm = io.ATemp[c];
  x.wsp = io.wSpeed[c];
  x.aqh = io.aqh[c];
  x.SWd = io.SWdown[c];
  return x;
}

template <typename C>
struct CatOut {
  C T, Fio, Fia, pen;
};

// category l of nITC (veris/growth.py:63-91)
template <typename C>
__device__ inline CatOut<C> category(const GrowthScalars& k, const CellIn<C>& x, int l) {
  const C p = (C)(2 * (l + 1) - 1) * (C)k.recip_nITC;
  CatOut<C> r;
  r.T = solve4temp(k, x.hIA * p, x.hSA * p, x.Ts0, x.Tfrz, x.fc, x.LWd, x.ATm, x.wsp, x.aqh, x.SWd,
                   &r.Fio, &r.Fia, &r.pen);
  return r;
}

// category means onwards (veris/growth.py:94-397), from the category sums
// accumulated in category order
template <typename T, typename C>
__device__ inline void tail(const GrowthScalars& k, const GrowthIO<T>& io, size_t c,
                            const CellIn<C>& x, C Tsum, C Fio_s, C Fia_s, C pen_s) {
  const C hP = x.hP, hsP = x.hsP, AP = x.AP, rregA = x.rregA, rhIA = x.rhIA, sal = x.sal,
          Tfrz = x.Tfrz, fc = x.fc;
  const C TSurf = Tsum * (C)k.recip_nITC;
  const C Fio = (Fio_s * (C)k.recip_nITC) * AP;
  C Fia = (Fia_s * (C)k.recip_nITC) * AP;
  const C pen = (pen_s * (C)k.recip_nITC) * AP;
  const bool cold = (AP > (C)0) && (Tsum / (C)k.nITC_d < (C)k.celsius2K);
  const C snowf = io.snowfall[c], prec = io.precip[c];
  const C snowRate = (cold ? snowf + prec : snowf) * (C)k.rhoFresh2rhoSnow;
  const C rainToSea = cold ? (C)0 : prec;
  const C snowAcc = (snowRate * AP) * (C)k.dtT;
  C gUnder = Fio * (C)k.qi;
  gUnder = AP == (C)0 ? (C)0 : gUnder;
  const C potMeltRate = (-Fia) * (C)k.qs;
  const C potMelt = potMeltRate * (C)k.dtT;
  const bool allMelt = potMelt >= hsP;
  const C melt = allMelt ? hsP : potMelt;
  const C meltRate = allMelt ? melt * (C)k.rdtT : potMeltRate;
  const C toMelt = allMelt ? ((-hsP) * (C)k.rdtT) / (C)k.qs : Fia;
  Fia = Fia - toMelt;
  const C gSurf = Fia * (C)k.qi;
  const C gNet = (gUnder + gSurf) * rregA;
  const C sth = fmax(io.theta[c], Tfrz);
  const C mltf = (C)1 + ((C)k.McPheeTaperFac - (C)1) / ((C)1 + exp((AP - (C)k.t0) * (C)k.t1));
  const C Foi = ((C)(-k.t2) * (sth - Tfrz)) * mltf;
  const C gML = Foi * (C)k.qi;
  const C qnet_in = io.Qnet[c];
  const C gOW = (C)k.qi * qnet_in;
  const C dh = (gNet * AP + gOW * ((C)1 - AP)) + gML;
  const C dhs = (snowRate - meltRate) * AP;
  const C half = (C)0.5 * rhIA;
  C dAoa = (half * gOW) * ((C)1 - AP);
  const bool grow = (gOW > (C)0) && ((AP > (C)0) || (dh > (C)0));
  if (grow) dAoa = gOW * ((C)1 - AP);
  dAoa = (grow && (fc < (C)0)) ? dAoa * (C)k.recip_h0_south : dAoa * (C)k.recip_h0;
  const C dAoi = gML <= (C)0 ? half * gML : (C)0;
  const C dAia = ((gNet <= (C)0) && (hP > (C)0)) ? (half * gNet) * AP : (C)0;
  const C dA = (dAoa + dAoi) + dAia;
  const C Mc = io.M[c];
  C A = AP + (dA * Mc) * (C)k.dtT;
  C h = hP + (dh * Mc) * (C)k.dtT;
  C hs = hsP + (dhs * Mc) * (C)k.dtT;
  A = fmin(fmax(A, (C)0), (C)1);
  h = fmax(h, (C)0);
  hs = fmax(hs, (C)0);
  const bool noIce = (h == (C)0) || (A == (C)0);
  A = A * (C)(!noIce);
  hs = hs * (C)(!noIce);
  h = h * (C)(!noIce);
  const C hsub = (hs * (C)k.rhoSnow + h * (C)k.rhoIce) * (C)k.recip_rhoSea;
  const C flood = fmax((C)0, hsub - h);
  h = h + flood;
  hs = hs - flood * (C)k.rhoIce2rhoSnow;
  const C rh = (C)1 / sqrt(h * h + (C)k.hIce_reg);
  const C Qsw = pen * AP + io.Qsw[c] * ((C)1 - AP);
  const C dH = h - hP;
  const C meltedSnow = (hsP + snowAcc) - hs;
  const C eIce = dH / (C)k.qi;
  const C eOut = (AP * ((Fia + Fio) + pen) + ((C)1 - AP) * qnet_in) * (C)k.dtT;
  const C Qnet = (eOut - eIce) * (C)k.rdtT;
  const C salFac = ((sal > (C)0) && (sal > (C)k.saltIce_ref)) ? (C)1 - (C)k.saltIce_ref / sal : (C)1;
  const C fwIce = ((-dH) * (C)k.rhoIce2rhoFresh) * salFac;
  const C t0s = fmin((C)k.saltIce_ref, sal);
  const C osh = io.osh[c], oshs = io.oshs[c];
  const C saltflux = ((((dH + osh) * t0s) * Mc) * (C)k.rhoIce) * (C)k.rdtT;
  const C fwSnow = meltedSnow / (C)k.rhoFresh2rhoSnow;
  const C emp = (Mc * (((((-io.evap[c]) - prec) * ((C)1 - AP) - rainToSea * AP) - io.runoff[c]) -
                       (fwIce + fwSnow) / (C)k.dtT)) * (C)k.rhoFresh +
                ((Mc * (osh * (C)k.rhoIce + oshs * (C)k.rhoSnow)) * (C)k.rdtT);
  const C fss = (emp * sal) / (C)k.rhoFresh + saltflux;
  io.oh[c] = (T)h;
  io.ohs[c] = (T)hs;
  io.oA[c] = (T)A;
  io.oTs[c] = (T)TSurf;
  io.oEmPmR[c] = (T)emp;
  io.ofss[c] = (T)fss;
  io.oQsw[c] = (T)Qsw;
  io.oQnet[c] = (T)Qnet;
  io.oload[c] = (T)(h * (C)k.rhoIce + hs * (C)k.rhoSnow);
  io.open[c] = (T)pen;
  io.orh[c] = (T)rh;
}

// one thread per padded cell, its categories in turn
template <typename T, typename C = double>
__global__ __launch_bounds__(BLK) void growth_kernel(GrowthScalars k, GrowthIO<T> io, size_t n) {
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const CellIn<C> x = cell_in<T, C>(k, io, c);
  C Tsum = 0, Fio_s = 0, Fia_s = 0, pen_s = 0;
  for (int l = 0; l < k.nITC; ++l) {
    const CatOut<C> o = category(k, x, l);
    if (l == 0) {
      Tsum = o.T;
      Fio_s = o.Fio;
      Fia_s = o.Fia;
      pen_s = o.pen;
    } else {
      Tsum = Tsum + o.T;
      Fio_s = Fio_s + o.Fio;
      Fia_s = Fia_s + o.Fia;
      pen_s = pen_s + o.pen;
    }
  }
  tail<T, C>(k, io, c, x, Tsum, Fio_s, Fia_s, pen_s);
}

template <typename T>
int growth(const Geom& g, const vh_settings* s, const vh_fields* f, const vh_growth_out* o,
           hipStream_t st) {
  VH_REQUIRE(o != nullptr, VH_EINVAL, "growth: null outputs");
  GrowthIO<T> io;
  io.h = (const T*)f->hIceMean;
  io.hs = (const T*)f->hSnowMean;
  io.A = (const T*)f->Area;
  io.Ts = (const T*)f->TSurf;
  io.ocSalt = (const T*)f->ocSalt;
  io.theta = (const T*)f->theta;
  io.snowfall = (const T*)f->snowfall;
  io.precip = (const T*)f->precip;
  io.Qnet = (const T*)f->Qnet;
  io.Qsw = (const T*)f->Qsw;
  io.evap = (const T*)f->evap;
  io.runoff = (const T*)f->runoff;
  io.osh = (const T*)f->os_hIceMean;
  io.oshs = (const T*)f->os_hSnowMean;
  io.M = (const T*)f->iceMask;
  io.fC = (const T*)f->fCori;
  io.LWdown = (const T*)f->LWdown;
  io.ATemp = (const T*)f->ATemp;
  io.wSpeed = (const T*)f->wSpeed;
  io.aqh = (const T*)f->aqh;
  io.SWdown = (const T*)f->SWdown;
  io.oh = (T*)o->hIceMean;
  io.ohs = (T*)o->hSnowMean;
  io.oA = (T*)o->Area;
  io.oTs = (T*)o->TSurf;
  io.oEmPmR = (T*)o->EmPmR;
  io.ofss = (T*)o->forc_salt_surface;
  io.oQsw = (T*)o->Qsw;
  io.oQnet = (T*)o->Qnet;
  io.oload = (T*)o->SeaIceLoad;
  io.open = (T*)o->IcePenetSW;
  io.orh = (T*)o->recip_hIceMean;
  const void* in[] = {io.h, io.hs, io.A, io.Ts, io.ocSalt, io.theta, io.snowfall, io.precip,
                      io.Qnet, io.Qsw, io.evap, io.runoff, io.osh, io.oshs, io.M, io.fC,
                      io.LWdown, io.ATemp, io.wSpeed, io.aqh, io.SWdown};
  const void* out[] = {io.oh, io.ohs, io.oA, io.oTs, io.oEmPmR, io.ofss, io.oQsw, io.oQnet,
                       io.oload, io.open, io.orh};
  for (const void* p : in) VH_REQUIRE(p != nullptr, VH_EINVAL, "growth: null input field");
  for (const void* p : out) {
    VH_REQUIRE(p != nullptr, VH_EINVAL, "growth: null output");
    for (const void* q : in) VH_REQUIRE(p != q, VH_EINVAL, "growth: output aliases an input");
  }
  const int nITC = (int)s->nITC;
  VH_REQUIRE(nITC >= 1 && nITC <= 64, VH_EINVAL, "nITC=%d", nITC);
  GrowthScalars k;
  k.qi = 1 / (s->rhoIce * s->lhFusion);
  k.qs = 1 / (s->rhoSnow * s->lhFusion);
  k.Area_reg = s->Area_reg;
  k.hIce_reg = s->hIce_reg;
  k.recip_nITC = s->recip_nITC;
  k.nITC = nITC;
  k.nITC_d = (double)nITC;
  k.tempFrz = s->tempFrz;
  k.dtempFrz_dS = s->dtempFrz_dS;
  k.celsius2K = s->celsius2K;
  k.dtT = s->deltatTherm;
  k.rdtT = s->recip_deltatTherm;
  k.t0 = 0.4;
  k.t1 = 7 / 0.4;
  k.t2 = s->stantonNr * s->uStarBase * s->rhoSea * s->cpWater;
  k.McPheeTaperFac = s->McPheeTaperFac;
  k.recip_h0 = s->recip_h0;
  k.recip_h0_south = s->recip_h0_south;
  k.rhoSnow = s->rhoSnow;
  k.rhoIce = s->rhoIce;
  k.recip_rhoSea = s->recip_rhoSea;
  k.rhoIce2rhoSnow = s->rhoIce2rhoSnow;
  k.rhoIce2rhoFresh = s->rhoIce2rhoFresh;
  k.rhoFresh = s->rhoFresh;
  k.rhoFresh2rhoSnow = s->rhoFresh2rhoSnow;
  k.snowRate_fac = s->rhoFresh2rhoSnow;
  k.saltIce_ref = s->saltIce_ref;
  k.d1 = s->dalton * s->cpAir * s->rhoAir;
  k.d1i = s->dalton * s->lhSublim * s->rhoAir;
  k.Tmelt = s->celsius2K;
  k.Tmeltw = k.Tmelt + s->wetAlbTemp;
  k.minLWdown = s->minLWdown;
  k.minTAir_K = s->celsius2K + s->minTAir;
  k.wSpeedMin = s->wSpeedMin;
  k.snowEmiss = s->snowEmiss;
  k.iceEmiss = s->iceEmiss;
  k.stefBoltz = s->stefBoltz;
  k.dryIceAlb = s->dryIceAlb;
  k.drySnowAlb = s->drySnowAlb;
  k.wetIceAlb = s->wetIceAlb;
  k.wetSnowAlb = s->wetSnowAlb;
  k.dryIceAlb_south = s->dryIceAlb_south;
  k.drySnowAlb_south = s->drySnowAlb_south;
  k.wetIceAlb_south = s->wetIceAlb_south;
  k.wetSnowAlb_south = s->wetSnowAlb_south;
  k.hCut = s->hCut;
  k.shortwave = s->shortwave;
  k.kIS = s->iceConduct * s->snowConduct;
  k.snowConduct = s->snowConduct;
  k.iceConduct = s->iceConduct;
  k.minTIce_K = s->celsius2K + s->minTIce;
  k.lhSublim = s->lhSublim;
  k.aa1 = 2663.5;
  k.aa2 = 12.537;
  k.bb1 = 0.622;
  k.bb1m = 1 - k.bb1;
  k.ppa = 100000;
  const double cc0 = std::pow(10.0, k.aa2);
  k.p10aa2 = cc0;
  k.cc1 = cc0 * k.aa1 * k.bb1 * k.ppa * std::log(10.0);
  k.cc2 = cc0 * (1 - k.bb1);
  const size_t n = field_elems(g);
  hipLaunchKernelGGL(growth_kernel<T>, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, st, k,
                     io, n);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

}  // namespace
}  // namespace vh

extern "C" int vh_growth(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                         const vh_growth_out* out, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(s != nullptr && f != nullptr, VH_EINVAL, "null settings or fields");
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64) return vh::growth<double>(g, s, f, out, st);
  return vh::growth<float>(g, s, f, out, st);
}
