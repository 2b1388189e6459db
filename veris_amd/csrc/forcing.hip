// Forcing producer feeding the step: the open-ocean bulk fluxes of
// veris/heat_flux_MITgcm.py:5-123 (bulkf_formula_lanl), one thread per
// padded cell.  Pointwise, so a plain coalesced sweep: 5 fields in, 9 out
// (112 B per cell in FP64); the five Monin-Obukhov iterations stay in
// registers.  Built with -ffp-contract=off: the expressions keep the
// reference's association order (log, exp, atan and pow come from the device
// math library, so results equal the reference to rounding).
#include "vh_common.h"

namespace vh {
namespace {

constexpr int BLK = 256;

template <typename T>
struct BulkIn {
  const T *uw, *vw, *ta, *qa, *tsf;
};

template <typename T>
struct BulkOut {
  T *flwupa, *flha, *fsha, *df0dT, *ust, *vst, *evp, *ssq, *dEvdT;
};

template <typename T>
struct BulkConst {
  T karman, czol, latvap, zvir, gamma_blk, rhoAir, cpdair, emis, stefBoltz;
};

template <typename T>
__global__ __launch_bounds__(BLK) void bulk_flux_kernel(size_t n, BulkIn<T> in, BulkConst<T> c,
                                                        BulkOut<T> o) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const T ht = (T)2.0, zref = (T)10.0, zice = (T)0.0005;
  const T aln = log(ht / zref);
  const T lath = c.latvap;
  const T uw = in.uw[idx], vw = in.vw[idx], ta = in.ta[idx], qa = in.qa[idx], tsf = in.tsf[idx];
  const T us = sqrt(uw * uw + vw * vw);
  const T usm = fmax(us, (T)1.0);
  const T t0 = ta * ((T)1.0 + c.zvir * qa);
  const T ssq = ((T)3.797915 * exp(lath * ((T)7.93252e-6 - (T)2.166847e-3 / tsf))) / (T)1013.0;
  const T deltap = (ta - tsf) + c.gamma_blk * ht;
  const T delq = qa - ssq;
  const T rdn = c.karman / log(zref / zice);
  const T rhn = rdn, ren = rdn;
  T ustar = rdn * usm, tstar = rhn * deltap, qstar = ren * delq;
  T rd = rdn, rh = rhn, re = ren;
  const T pi_half = (T)M_PI * (T)0.5;
  for (int it = 0; it < 5; ++it) {
    T huol = (c.czol / (ustar * ustar)) * (tstar / t0 + qstar / ((T)1.0 / c.zvir + qa));
    huol = fmin(fabs(huol), (T)10.0) * sgn(huol);
    const T stable = (T)0.5 + (T)0.5 * sgn(huol);
    const T xsq = fmax(sqrt(fabs((T)1.0 - (T)16.0 * huol)), (T)1.0);
    const T x = sqrt(xsq);
    const T st5 = ((T)(-5.0) * huol) * stable;
    const T lx2 = (T)2.0 * log((T)0.5 * ((T)1.0 + xsq));
    const T psimh =
        st5 + ((T)1.0 - stable) *
                  ((((T)2.0 * log((T)0.5 * ((T)1.0 + x)) + lx2) - (T)2.0 * atan(x)) + pi_half);
    const T psixh = st5 + ((T)1.0 - stable) * lx2;
    rd = rdn / ((T)1.0 + (rdn * (aln - psimh)) / c.karman);
    rh = rhn / ((T)1.0 + (rhn * (aln - psixh)) / c.karman);
    re = rh;
    ustar = rd * usm;
    qstar = re * delq;
    tstar = rh * deltap;
  }
  const T csha = (((c.rhoAir * c.cpdair) * us) * rh) * rd;
  const T clha = (((c.rhoAir * lath) * us) * re) * rd;
  const T fsha = csha * deltap;
  const T flha = clha * delq;
  const T es = c.emis * c.stefBoltz;
  const T flwupa = es * pow(tsf, (T)4);
  const T dflwupdt = ((T)4.0 * c.emis) * c.stefBoltz * pow(tsf, (T)3);
  const T devdt = ((clha * ssq) * (T)2.166847e-3) / (tsf * tsf);
  const T dflhdt = (-lath) * devdt;
  const T dfshdt = -csha;
  const T cdn = ((T)2.7e-3 / usm + (T)0.142e-3) + (T)0.0764e-3 * usm;
  o.flwupa[idx] = flwupa;
  o.flha[idx] = flha;
  o.fsha[idx] = fsha;
  o.df0dT[idx] = ((-dflwupdt) + dfshdt) + dflhdt;
  o.ust[idx] = ((c.rhoAir * cdn) * us) * uw;
  o.vst[idx] = ((c.rhoAir * cdn) * us) * vw;
  o.evp[idx] = (-flha) / lath;
  o.ssq[idx] = ssq;
  o.dEvdT[idx] = devdt;
}

template <typename T>
int bulk_flux(const Geom& g, const vh_settings* s, double grav, const void* const in[5],
              const vh_bulk_flux_out* out, hipStream_t st) {
  VH_REQUIRE(out != nullptr, VH_EINVAL, "bulkf_formula_lanl: null output struct");
  void* const o[9] = {out->flwupa, out->flha, out->fsha, out->df0dT, out->ust,
                      out->vst,    out->evp,  out->ssq,  out->dEvdT};
  for (int q = 0; q < 5; ++q) VH_REQUIRE(in[q] != nullptr, VH_EINVAL, "bulkf_formula_lanl: null input");
  for (int q = 0; q < 9; ++q) {
    VH_REQUIRE(o[q] != nullptr, VH_EINVAL, "bulkf_formula_lanl: null output");
    for (int r = 0; r < q; ++r) VH_REQUIRE(o[q] != o[r], VH_EINVAL, "bulkf_formula_lanl: outputs alias");
    for (int r = 0; r < 5; ++r)
      VH_REQUIRE(o[q] != in[r], VH_EINVAL, "bulkf_formula_lanl: output aliases an input");
  }
  BulkIn<T> bi{(const T*)in[0], (const T*)in[1], (const T*)in[2], (const T*)in[3], (const T*)in[4]};
  BulkOut<T> bo{(T*)o[0], (T*)o[1], (T*)o[2], (T*)o[3], (T*)o[4],
                (T*)o[5], (T*)o[6], (T*)o[7], (T*)o[8]};
  BulkConst<T> c;
  c.karman = (T)s->karman;
  c.czol = (T)((10.0 * s->karman) * grav);  // zref * karman * grav
  c.latvap = (T)s->latvap;
  c.zvir = (T)s->zvir;
  c.gamma_blk = (T)s->gamma_blk;
  c.rhoAir = (T)s->rhoAir;
  c.cpdair = (T)s->cpdair;
  c.emis = (T)s->ocean_emissivity;
  c.stefBoltz = (T)s->stefBoltz;
  const size_t n = field_elems(g);
  hipLaunchKernelGGL(bulk_flux_kernel<T>, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, st,
                     n, bi, c, bo);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

}  // namespace
}  // namespace vh

extern "C" int vh_bulkf_formula_lanl(const vh_domain* d, const vh_settings* s, double grav,
                                     const void* uw, const void* vw, const void* ta,
                                     const void* qa, const void* tsf, const void* ocn_mask,
                                     const vh_bulk_flux_out* out, void* stream) {
  (void)ocn_mask;  // the reference uses it only for the shape of lath
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(s != nullptr, VH_EINVAL, "null settings");
  const void* const in[5] = {uw, vw, ta, qa, tsf};
  hipStream_t st = (hipStream_t)stream;
  return d->dtype == VH_F64 ? vh::bulk_flux<double>(g, s, grav, in, out, st)
                            : vh::bulk_flux<float>(g, s, grav, in, out, st);
}
