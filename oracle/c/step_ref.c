/*
 * Scalar-loop C restatement of the whole Veris sea-ice step -- TEST
 * INFRASTRUCTURE ONLY (a checker of the NumPy oracle, never the product).
 *
 * Follows the step of veris/setup/seaice_global_4deg/seaice_global_4deg.py:874-932:
 * SeaIceMass and AreaWS (veris/area_mass.py:5-28), WindForcingXY
 * (veris/dynsolver.py:7-84), SeaIceStrength (veris/dynamics_routines.py:5-18),
 * IceVelocities (veris/dynsolver.py:87-102 -> evp_ref.c), OceanStressUV
 * (veris/ocean_stress.py:8-40), Advection (veris/advection.py:15-175),
 * clean_up_advection and ridging (veris/clean_up.py:5-47), Growth
 * (veris/growth.py:8-397) with solve4temp (veris/solve4temp.py:5-174), and the
 * overlap fills (veris/fill_overlap.py:6-18).  Cell by cell with explicit
 * wrapped indices, every expression in the reference's association order, so
 * that an FP64 evaluation without contraction (oracle/c/Makefile) gives the
 * NumPy oracle's bits wherever only +, -, *, /, sqrt and comparisons enter.
 *
 * Fields and settings are passed by name; the fields are updated in place as
 * the oracle updates state.variables.  np.maximum / np.minimum are restated
 * with NumPy's results: NaN propagates, and on a tie (+0 against -0) the
 * second operand is returned; np.clip keeps the value on a tie.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "evp_ref.h"

typedef struct {
  int NX, NY;
  int nf, ns;
  const char* const* fnames;
  double* const* fields;
  const char* const* snames;
  const double* svals;
  int missing;
} Step;

static char g_missing[64];  /* name of the first missing field or setting */

const char* step_ref_missing(void) { return g_missing; }

static void note_missing(Step* S, const char* name) {
  if (!S->missing) {
    strncpy(g_missing, name, sizeof(g_missing) - 1);
    g_missing[sizeof(g_missing) - 1] = 0;
  }
  S->missing = 1;
}

static double* F_(Step* S, const char* name) {
  for (int k = 0; k < S->nf; ++k)
    if (strcmp(S->fnames[k], name) == 0) return S->fields[k];
  note_missing(S, name);
  return NULL;
}

static double P_(Step* S, const char* name) {
  for (int k = 0; k < S->ns; ++k)
    if (strcmp(S->snames[k], name) == 0) return S->svals[k];
  note_missing(S, name);
  return 0.0;
}

/* optional setting: `dflt` when absent */
static double Popt(Step* S, const char* name, double dflt) {
  for (int k = 0; k < S->ns; ++k)
    if (strcmp(S->snames[k], name) == 0) return S->svals[k];
  return dflt;
}

static int wr(int i, int n) { return ((i % n) + n) % n; }
#define NXY (S->NX * S->NY)
#define AT(x, y) ((size_t)wr((x), S->NX) * (size_t)S->NY + (size_t)wr((y), S->NY))
/* every EACH body writes only its own cell (see FOR_CELLS in evp_ref.c) */
#define EACH                                                                         \
  _Pragma("omp parallel for schedule(static) if (S->NX * S->NY >= 16384)")          \
  for (int i = 0; i < S->NX; ++i) for (int j = 0; j < S->NY; ++j)

static double npmax(double a, double b) { return a != a ? a : (b != b ? b : (a > b ? a : b)); }
static double npmin(double a, double b) { return a != a ? a : (b != b ? b : (a < b ? a : b)); }
static double npclip(double x, double lo, double hi) {
  if (x != x) return x;
  const double t = x >= lo ? x : lo;
  return t <= hi ? t : hi;
}
static double npsign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : (x == 0 ? 0.0 : x)); }

static double* tmp(size_t n) { return (double*)calloc(n, sizeof(double)); }

/* veris/fill_overlap.py:6-18 */
static void fill(Step* S, double* a) {
  const int NX = S->NX, NY = S->NY;
  if (P_(S, "veros_fill") != 0.0) {
    if (P_(S, "enable_cyclic_x") == 0.0) return;
    for (int j = 0; j < NY; ++j)
      for (int k = 0; k < 2; ++k) {
        a[(size_t)(NX - 2 + k) * NY + j] = a[(size_t)(2 + k) * NY + j];
        a[(size_t)k * NY + j] = a[(size_t)(NX - 4 + k) * NY + j];
      }
    return;
  }
  for (int j = 0; j < NY; ++j)
    for (int k = 0; k < 2; ++k) {
      a[(size_t)k * NY + j] = a[(size_t)(NX - 4 + k) * NY + j];
      a[(size_t)(NX - 2 + k) * NY + j] = a[(size_t)(2 + k) * NY + j];
    }
  for (int i = 0; i < NX; ++i)
    for (int k = 0; k < 2; ++k) {
      a[(size_t)i * NY + k] = a[(size_t)i * NY + NY - 4 + k];
      a[(size_t)i * NY + NY - 2 + k] = a[(size_t)i * NY + 2 + k];
    }
}

/* veris/area_mass.py:5-28 */
static void mass_area(Step* S) {
  const double *h = F_(S, "hIceMean"), *hs = F_(S, "hSnowMean"), *A = F_(S, "Area");
  double *mC = F_(S, "SeaIceMassC"), *mU = F_(S, "SeaIceMassU"), *mV = F_(S, "SeaIceMassV");
  double *AW = F_(S, "AreaW"), *AS = F_(S, "AreaS");
  const double rI = P_(S, "rhoIce"), rS = P_(S, "rhoSnow");
  if (S->missing) return;
  EACH mC[AT(i, j)] = rI * h[AT(i, j)] + rS * hs[AT(i, j)];
  EACH {
    mU[AT(i, j)] = 0.5 * (mC[AT(i, j)] + mC[AT(i - 1, j)]);
    mV[AT(i, j)] = 0.5 * (mC[AT(i, j)] + mC[AT(i, j - 1)]);
    AW[AT(i, j)] = 0.5 * (A[AT(i, j)] + A[AT(i - 1, j)]);
    AS[AT(i, j)] = 0.5 * (A[AT(i, j)] + A[AT(i, j - 1)]);
  }
}

/* veris/dynsolver.py:7-84 */
static void wind_forcing(Step* S) {
  const double *uW = F_(S, "uWind"), *vW = F_(S, "vWind"), *u = F_(S, "uIce"), *v = F_(S, "vIce");
  const double *fC = F_(S, "fCori"), *MU = F_(S, "iceMaskU"), *MV = F_(S, "iceMaskV");
  const double *AW = F_(S, "AreaW"), *AS = F_(S, "AreaS"), *ssh = F_(S, "ssh_an");
  const double *sp = F_(S, "surfPress"), *load = F_(S, "SeaIceLoad");
  const double *mU = F_(S, "SeaIceMassU"), *mV = F_(S, "SeaIceMassV");
  const double *rdxC = F_(S, "recip_dxC"), *rdyC = F_(S, "recip_dyC");
  double *WX = F_(S, "WindForcingX"), *WY = F_(S, "WindForcingY");
  const double caN = P_(S, "airIceDrag"), caS = P_(S, "airIceDrag_south");
  const double ang = P_(S, "airTurnAngle"), rel = P_(S, "useRelativeWind");
  const double wmin = P_(S, "wSpeedMin"), rhoA = P_(S, "rhoAir"), grav = P_(S, "gravity");
  const double rfw = P_(S, "useRealFreshWaterFlux"), lf = P_(S, "seaIceLoadFac");
  const double rrs = P_(S, "recip_rhoSea");
  if (S->missing) return;
  const size_t n = (size_t)NXY;
  double *tx = tmp(n), *ty = tmp(n), *phi = tmp(n);
  const double rad = ang * (M_PI / 180.0);
  const double sinW = sin(rad), cosW = cos(rad);
  EACH {
    const size_t c = AT(i, j);
    double ur = uW[c], vr = vW[c];
    if (rel != 0.0) {
      ur = uW[c] - 0.5 * (u[c] + u[AT(i + 1, j)]);
      vr = vW[c] - 0.5 * (v[c] + v[AT(i, j + 1)]);
    }
    const double wsq = ur * ur + vr * vr;
    const double ws = wsq < wmin * wmin ? wmin : sqrt(wsq);
    const double cda = ((fC[c] < 0 ? caS : caN) * rhoA) * ws;
    const double sg = npsign(fC[c]);
    tx[c] = cda * (cosW * ur - (sg * sinW) * vr);
    ty[c] = cda * (cosW * vr + (sg * sinW) * ur);
    double p = grav * ssh[c];
    if (rfw != 0.0) p = p + (sp[c] + (load[c] * grav) * lf) * rrs;
    else p = p + sp[c] * rrs;
    phi[c] = p;
  }
  EACH {
    const size_t c = AT(i, j);
    const double txu = (0.5 * (tx[c] + tx[AT(i - 1, j)])) * MU[c];
    const double tyv = (0.5 * (ty[c] + ty[AT(i, j - 1)])) * MV[c];
    WX[c] = txu * AW[c] - (mU[c] * rdxC[c]) * (phi[c] - phi[AT(i - 1, j)]);
    WY[c] = tyv * AS[c] - (mV[c] * rdyC[c]) * (phi[c] - phi[AT(i, j - 1)]);
  }
  free(tx);
  free(ty);
  free(phi);
}

/* veris/dynamics_routines.py:5-18 */
static void strength(Step* S) {
  const double *h = F_(S, "hIceMean"), *A = F_(S, "Area"), *M = F_(S, "iceMask");
  double* P = F_(S, "SeaIceStrength");
  const double ps = P_(S, "pStar"), cs = P_(S, "cStar");
  if (S->missing) return;
  /* keep_strength: the caller supplies P (so that exp of another library
     does not enter a bit-for-bit comparison of the rest of the step) */
  if (Popt(S, "keep_strength", 0.0) != 0.0) return;
  EACH {
    const size_t c = AT(i, j);
    P[c] = ((ps * h[c]) * exp(-cs * (1 - A[c]))) * M[c];
  }
}

/* EVP subcycle or free drift through evp_ref.c (veris/dynsolver.py:87-102) */
static const char* const kSetNames[ES_COUNT] = {
    "NX", "NY", "noSlip", "secondOrderBC", "PlasDefCoeff", "deltaMin", "tensileStrFac",
    "pressReplFac", "waterIceDrag", "waterIceDrag_south", "rhoSea", "cDragMin", "basalDragK1",
    "basalDragK2", "basalDragU0", "cBasalStar", "use_coastline", "sideDragCoeff", "sideDragU0",
    "veros_fill", "enable_cyclic_x", "useAdaptiveEVP", "deltatDyn", "aEvpCoeff", "evpAlpha",
    "aEVPalphaMin", "evpBeta", "recip_deltatDyn", "cosWat", "sinWat", "rhoIce"};
static const char* const kFieldNames[EF_COUNT] = {
    "uIce", "vIce", "sigma1", "sigma2", "sigma12", "iceMask", "iceMaskU", "iceMaskV",
    "maskInC", "maskInU", "maskInV", "recip_dxU", "recip_dyU", "recip_dxV", "recip_dyV", "dxU",
    "dyU", "dxV", "dyV", "recip_rA", "rAz", "recip_rAu", "recip_rAv", "k1AtC", "k2AtC", "k1AtZ",
    "k2AtZ", "SeaIceStrength", "fCori", "uOcean", "vOcean", "Area", "hIceMean", "R_low", "AreaW",
    "AreaS", "SeaIceMassC", "SeaIceMassU", "SeaIceMassV", "WindForcingX", "WindForcingY", "Fu",
    "Fv"};

static int ice_velocities(Step* S, int nsub) {
  double es[ES_COUNT];
  const double* ef[EF_COUNT];
  for (int k = 0; k < ES_COUNT; ++k)
    es[k] = k == ES_NX ? S->NX : (k == ES_NY ? S->NY : P_(S, kSetNames[k]));
  for (int k = 0; k < EF_COUNT; ++k) ef[k] = F_(S, kFieldNames[k]);
  const double useEVP = P_(S, "useEVP"), useFD = P_(S, "useFreedrift");
  if (S->missing) return -2;
  const size_t n = (size_t)NXY;
  double *u = tmp(n), *v = tmp(n), *s1 = tmp(n), *s2 = tmp(n), *s12 = tmp(n);
  int rc = 0;
  if (useEVP != 0.0) {
    rc = evp_solver_ref(es, ef, nsub, u, v, s1, s2, s12);
    if (rc == 0) {
      memcpy(F_(S, "sigma1"), s1, n * sizeof(double));
      memcpy(F_(S, "sigma2"), s2, n * sizeof(double));
      memcpy(F_(S, "sigma12"), s12, n * sizeof(double));
    }
  } else if (useFD != 0.0) {
    rc = freedrift_ref(es, ef, u, v);
  } else {
    rc = -3;
  }
  if (rc == 0) {
    memcpy(F_(S, "uIce"), u, n * sizeof(double));
    memcpy(F_(S, "vIce"), v, n * sizeof(double));
  }
  free(u);
  free(v);
  free(s1);
  free(s2);
  free(s12);
  return rc;
}

/* veris/ocean_stress.py:8-40 with the drag of veris/dynamics_routines.py:21-52 */
static void ocean_stress(Step* S) {
  const double *u = F_(S, "uIce"), *v = F_(S, "vIce"), *uo = F_(S, "uOcean"), *vo = F_(S, "vOcean");
  const double *fC = F_(S, "fCori"), *mIU = F_(S, "maskInU"), *mIV = F_(S, "maskInV");
  const double* M = F_(S, "iceMask");
  double *tu = F_(S, "OceanStressU"), *tv = F_(S, "OceanStressV");
  const double cwN = P_(S, "waterIceDrag"), cwS = P_(S, "waterIceDrag_south");
  const double rho = P_(S, "rhoSea"), cmin = P_(S, "cDragMin"), ang = P_(S, "waterTurnAngle");
  if (S->missing) return;
  const size_t n = (size_t)NXY;
  double *cd = tmp(n), *du = tmp(n), *dv = tmp(n), *cdv = tmp(n), *cdu = tmp(n);
  EACH {
    const size_t c = AT(i, j);
    const double D = (fC[c] < 0 ? cwS : cwN) * rho;
    const double a = (u[c] - uo[c]) * mIU[c], ap = (u[AT(i + 1, j)] - uo[AT(i + 1, j)]) * mIU[AT(i + 1, j)];
    const double b = (v[c] - vo[c]) * mIV[c], bp = (v[AT(i, j + 1)] - vo[AT(i, j + 1)]) * mIV[AT(i, j + 1)];
    const double w = 0.25 * ((a + ap) * (a + ap) + (b + bp) * (b + bp));
    cd[c] = ((D * D) * w > cmin * cmin ? D * sqrt(w) : cmin) * M[c];
    du[c] = u[c] - uo[c];
    dv[c] = v[c] - vo[c];
  }
  EACH {
    const size_t c = AT(i, j);
    cdv[c] = cd[c] * (0.5 * (dv[c] + dv[AT(i, j + 1)]));
    cdu[c] = cd[c] * (0.5 * (du[c] + du[AT(i + 1, j)]));
  }
  const double rad = ang * (M_PI / 180.0);
  const double sinW = sin(rad), cosW = cos(rad);
  EACH {
    const size_t c = AT(i, j);
    const double sg = npsign(fC[c]);
    tu[c] = ((0.5 * (cd[c] + cd[AT(i - 1, j)])) * cosW) * du[c] -
            ((sg * sinW) * 0.5) * (cdv[c] + cdv[AT(i, j - 1)]);
    tv[c] = ((0.5 * (cd[c] + cd[AT(i, j - 1)])) * cosW) * dv[c] +
            ((sg * sinW) * 0.5) * (cdu[c] + cdu[AT(i - 1, j)]);
  }
  fill(S, tu);
  fill(S, tv);
  free(cd);
  free(du);
  free(dv);
  free(cdv);
  free(cdu);
}

/* veris/advection.py:170-175 and the slope ratio of :94-129 */
static double superbee(double cr) { return npmax(0, npmax(npmin(1, 2 * cr), npmin(2, cr))); }

static double limited(double rjm, double rj, double rjp, double trans, double crmax) {
  double cr = trans > 0 ? rjm : rjp;
  cr = fabs(rj) * crmax > fabs(cr) ? cr / rj : (npsign(cr) * crmax) * npsign(rj);
  return superbee(cr);
}

/* veris/advection.py:28-129 for one extensive field */
static void advect_field(Step* S, double* f0, const double* uT, const double* vT) {
  const double *u = F_(S, "uIce"), *v = F_(S, "vIce"), *MU = F_(S, "iceMaskU");
  const double *MV = F_(S, "iceMaskV"), *mIU = F_(S, "maskInU"), *mIV = F_(S, "maskInV");
  const double *rdxC = F_(S, "recip_dxC"), *rdyC = F_(S, "recip_dyC"), *mIC = F_(S, "maskInC");
  const double *rrA = F_(S, "recip_rA"), *M = F_(S, "iceMask");
  const double dt = P_(S, "deltatTherm"), crmax = P_(S, "CrMax");
  const int NX = S->NX, NY = S->NY;
  const size_t n = (size_t)NXY;
  double *f = tmp(n), *Fx = tmp(n);
  memcpy(f, f0, n * sizeof(double));
  /* zonal flux on i in [2, NX-1), zero elsewhere, then the overlap fill */
  for (int i = 2; i < NX - 1; ++i)
    for (int j = 0; j < NY; ++j) {
      const size_t c = (size_t)i * NY + j, cm = c - NY, cmm = c - 2 * (size_t)NY, cp = c + NY;
      const double rjp = (f[cp] - f[c]) * (MU[cp] * mIU[cp]);
      const double rj = (f[c] - f[cm]) * (MU[c] * mIU[c]);
      const double rjm = (f[cm] - f[cmm]) * (MU[cm] * mIU[cm]);
      const double cr = limited(rjm, rj, rjp, uT[c], crmax);
      const double cfl = fabs((u[c] * dt) * rdxC[c]);
      Fx[c] = (uT[c] * (f[c] + f[cm])) * 0.5 - ((fabs(uT[c]) * ((1 - cr) + cfl * cr)) * rj) * 0.5;
    }
  fill(S, Fx);
  EACH {
    const size_t c = AT(i, j);
    f[c] = f[c] - ((dt * mIC[c]) * rrA[c]) * (Fx[AT(i + 1, j)] - Fx[c]);
  }
  double* Fy = tmp(n);
  for (int i = 0; i < NX; ++i)
    for (int j = 2; j < NY - 1; ++j) {
      const size_t c = (size_t)i * NY + j;
      const double rjp = (f[c + 1] - f[c]) * (MV[c + 1] * mIV[c + 1]);
      const double rj = (f[c] - f[c - 1]) * (MV[c] * mIV[c]);
      const double rjm = (f[c - 1] - f[c - 2]) * (MV[c - 1] * mIV[c - 1]);
      const double cr = limited(rjm, rj, rjp, vT[c], crmax);
      const double cfl = fabs((v[c] * dt) * rdyC[c]);
      Fy[c] = (vT[c] * (f[c] + f[c - 1])) * 0.5 - ((fabs(vT[c]) * ((1 - cr) + cfl * cr)) * rj) * 0.5;
    }
  fill(S, Fy);
  EACH {
    const size_t c = AT(i, j);
    f[c] = f[c] - ((dt * mIC[c]) * rrA[c]) * (Fy[AT(i, j + 1)] - Fy[c]);
  }
  EACH f0[AT(i, j)] = f[AT(i, j)] * M[AT(i, j)];
  free(f);
  free(Fx);
  free(Fy);
}

static void advection(Step* S) {
  const double *u = F_(S, "uIce"), *v = F_(S, "vIce"), *dyG = F_(S, "dyG"), *dxG = F_(S, "dxG");
  const double *MU = F_(S, "iceMaskU"), *MV = F_(S, "iceMaskV");
  double *h = F_(S, "hIceMean"), *hs = F_(S, "hSnowMean"), *A = F_(S, "Area");
  if (S->missing) return;
  const size_t n = (size_t)NXY;
  double *uT = tmp(n), *vT = tmp(n);
  EACH {
    const size_t c = AT(i, j);
    uT[c] = u[c] * (dyG[c] * MU[c]);
    vT[c] = v[c] * (dxG[c] * MV[c]);
  }
  advect_field(S, h, uT, vT);
  advect_field(S, hs, uT, vT);
  advect_field(S, A, uT, vT);
  free(uT);
  free(vT);
}

/* veris/clean_up.py:5-47 */
static void clean_up_and_ridging(Step* S) {
  double *h = F_(S, "hIceMean"), *hs = F_(S, "hSnowMean"), *A = F_(S, "Area");
  double *T = F_(S, "TSurf"), *osh = F_(S, "os_hIceMean"), *oshs = F_(S, "os_hSnowMean");
  const double hmin = P_(S, "hIce_min"), c2k = P_(S, "celsius2K"), amin = P_(S, "Area_min");
  if (S->missing) return;
  EACH {
    const size_t c = AT(i, j);
    osh[c] = npmax(-h[c], 0);
    oshs[c] = npmax(-hs[c], 0);
    double hh = npmax(h[c], 0), hhs = npmax(hs[c], 0), a = npmax(A[c], 0);
    const int thin = hh <= hmin;
    hh = hh * (double)!thin;
    hhs = hhs * (double)!thin;
    T[c] = thin ? c2k : T[c];
    a = (hh == 0 && hhs == 0) ? 0 : a;
    a = (hh > 0 || hhs > 0) ? npmax(a, amin) : a;
    h[c] = hh;
    hs[c] = hhs;
    A[c] = npmin(a, 1);
  }
}

/* veris/solve4temp.py:5-174 at one cell */
typedef struct {
  double d1, d1i, Tmelt, Tmeltw, minLW, minTAir, c2k, wmin, snowE, iceE, sb, dIA, dSA, wIA,
      wSA, dIAs, dSAs, wIAs, wSAs, hCut, shortwave, iceK, snowK, lhSub, minTIce, AA1, AA2, BB1,
      PPA, CC1, CC2;
  int n_newton;
} S4T;

static void s4t_fluxes(const S4T* q, double t1, int ice, double kEff, double Tfrz, double ug,
                       double aqh, double d3, double ta, double lw, double absSW, double* Fc,
                       double* Flh, double* Fia, double* dF) {
  const double t2 = t1 * t1, t3 = t2 * t1, t4 = t2 * t2;
  if (!ice) {
    *Fc = 0;
    *Flh = 0;
    *Fia = 0;
    *dF = 0;
    return;
  }
  const double svp = pow(10.0, -q->AA1 / t1 + q->AA2);
  const double qs = (q->BB1 * svp) / (q->PPA - (1 - q->BB1) * svp);
  const double c3 = pow(10.0, q->AA1 / t1);
  const double den = q->CC2 - c3 * q->PPA;
  const double dq = (q->CC1 * c3) / ((den * den) * t2);
  *Fc = kEff * (Tfrz - t1);
  *Flh = (q->d1i * ug) * (qs - aqh);
  const double Flw = t4 * d3;
  const double Fs = (q->d1 * ug) * (t1 - ta);
  *Fia = (((-lw - absSW) + Flw) + Fs) + *Flh;
  *dF = ((4 * d3) * t3 + q->d1 * ug) + (q->d1i * ug) * dq;
}

static void solve4temp_cell(const S4T* q, double hI, double hS, double Tin, double Tfrz,
                            double LWdown, double ATemp, double wSpeed, double fC, double SWdown,
                            double aqh, double* Tout, double* Fio, double* FiaNet, double* penSW) {
  double T = Tin;
  const double lwc = npmax(q->minLW, LWdown);
  const double ta = npmax(q->c2k + q->minTAir, ATemp);
  const double ug = npmax(q->wmin, wSpeed);
  const int ice = hI > 0, snow = hS > 0;
  const double d3 = (snow ? q->snowE : q->iceE) * q->sb;
  const double lw = (snow ? q->snowE : q->iceE) * lwc;
  double aI = ice ? q->dIA : 0, aS = ice ? q->dSA : 0;
  if (ice && T >= q->Tmeltw) {
    aI = q->wIA;
    aS = q->wSA;
  }
  if (ice && fC < 0) {
    aI = q->dIAs;
    aS = q->dSAs;
  }
  if (ice && fC < 0 && T >= q->Tmeltw) {
    aI = q->wIAs;
    aS = q->wSAs;
  }
  double alb = ice ? aI + (hS / q->hCut) * (aS - aI) : 0;
  alb = hS > q->hCut ? aS : alb;
  alb = hS == 0 ? aI : alb;
  double pen = ice ? q->shortwave * exp(-1.5 * hI) : 0;
  pen = snow ? 0 : pen;
  *penSW = ice ? ((-(1 - alb)) * pen) * SWdown : 0;
  const double absSW = ice ? ((1 - alb) * (1 - pen)) * SWdown : 0;
  const double kEff = ice ? (q->iceK * q->snowK) / (q->snowK * hI + q->iceK * hS) : 0;
  double Fc, Flh, Fia, dF;
  for (int it = 0; it < q->n_newton; ++it) {
    s4t_fluxes(q, T, ice, kEff, Tfrz, ug, aqh, d3, ta, lw, absSW, &Fc, &Flh, &Fia, &dF);
    T = ice ? T + (Fc - Fia) / (kEff + dF) : 0;
    T = npmin(T, q->Tmelt);
    T = npmax(T, q->c2k + q->minTIce);
  }
  s4t_fluxes(q, T, ice, kEff, Tfrz, ug, aqh, d3, ta, lw, absSW, &Fc, &Flh, &Fia, &dF);
  const int up = Fc > 0;
  *Fio = up ? Fc : 0;
  *FiaNet = up ? 0 : Fia;
  *Tout = ice ? T : Tin;
}

/* veris/growth.py:8-397 */
static void growth(Step* S) {
  double *h = F_(S, "hIceMean"), *hs = F_(S, "hSnowMean"), *A = F_(S, "Area");
  double *TS = F_(S, "TSurf"), *EmPmR = F_(S, "EmPmR"), *fss = F_(S, "forc_salt_surface_ice");
  double *Qsw = F_(S, "Qsw"), *Qnet = F_(S, "Qnet"), *load = F_(S, "SeaIceLoad");
  double *penO = F_(S, "IcePenetSW"), *rh = F_(S, "recip_hIceMean");
  const double *ocS = F_(S, "ocSalt"), *LW = F_(S, "LWdown"), *AT_ = F_(S, "ATemp");
  const double *wS = F_(S, "wSpeed"), *fC = F_(S, "fCori"), *SW = F_(S, "SWdown");
  const double *aqh = F_(S, "aqh"), *snowfall = F_(S, "snowfall"), *precip = F_(S, "precip");
  const double *theta = F_(S, "theta"), *M = F_(S, "iceMask"), *osh = F_(S, "os_hIceMean");
  const double *oshs = F_(S, "os_hSnowMean"), *evap = F_(S, "evap"), *runoff = F_(S, "runoff");
  S4T q;
  q.d1 = (P_(S, "dalton") * P_(S, "cpAir")) * P_(S, "rhoAir");
  q.d1i = (P_(S, "dalton") * P_(S, "lhSublim")) * P_(S, "rhoAir");
  q.c2k = P_(S, "celsius2K");
  q.Tmelt = q.c2k;
  q.Tmeltw = q.Tmelt + P_(S, "wetAlbTemp");
  q.minLW = P_(S, "minLWdown");
  q.minTAir = P_(S, "minTAir");
  q.wmin = P_(S, "wSpeedMin");
  q.snowE = P_(S, "snowEmiss");
  q.iceE = P_(S, "iceEmiss");
  q.sb = P_(S, "stefBoltz");
  q.dIA = P_(S, "dryIceAlb");
  q.dSA = P_(S, "drySnowAlb");
  q.wIA = P_(S, "wetIceAlb");
  q.wSA = P_(S, "wetSnowAlb");
  q.dIAs = P_(S, "dryIceAlb_south");
  q.dSAs = P_(S, "drySnowAlb_south");
  q.wIAs = P_(S, "wetIceAlb_south");
  q.wSAs = P_(S, "wetSnowAlb_south");
  q.hCut = P_(S, "hCut");
  q.shortwave = P_(S, "shortwave");
  q.iceK = P_(S, "iceConduct");
  q.snowK = P_(S, "snowConduct");
  q.lhSub = P_(S, "lhSublim");
  q.minTIce = P_(S, "minTIce");
  q.AA1 = P_(S, "AA1");
  q.AA2 = P_(S, "AA2");
  q.BB1 = P_(S, "BB1");
  q.PPA = P_(S, "PPASCALS");
  q.CC1 = P_(S, "CC1");
  q.CC2 = P_(S, "CC2");
  q.n_newton = 6;
  const double rI = P_(S, "rhoIce"), rS = P_(S, "rhoSnow"), lhF = P_(S, "lhFusion");
  const double Areg = P_(S, "Area_reg"), hreg = P_(S, "hIce_reg");
  const double tF = P_(S, "tempFrz"), dtF = P_(S, "dtempFrz_dS");
  const int nITC = (int)P_(S, "nITC");
  const double rn = P_(S, "recip_nITC"), r2s = P_(S, "rhoFresh2rhoSnow");
  const double dt = P_(S, "deltatTherm"), rdt = P_(S, "recip_deltatTherm");
  const double st = P_(S, "stantonNr"), uSB = P_(S, "uStarBase"), rhoSea = P_(S, "rhoSea");
  const double cpW = P_(S, "cpWater"), McP = P_(S, "McPheeTaperFac");
  const double rh0 = P_(S, "recip_h0"), rh0s = P_(S, "recip_h0_south");
  const double rrs = P_(S, "recip_rhoSea"), rI2rS = P_(S, "rhoIce2rhoSnow");
  const double sIr = P_(S, "saltIce_ref"), rI2rF = P_(S, "rhoIce2rhoFresh");
  const double rhoF = P_(S, "rhoFresh"), t0 = P_(S, "McPheeT0"), t1 = P_(S, "McPheeT1");
  if (S->missing) return;
  const double qi = 1 / (rI * lhF), qs = 1 / (rS * lhF);
  const double t2 = ((st * uSB) * rhoSea) * cpW;
  EACH {
    const size_t c = AT(i, j);
    const int noIce0 = h[c] == 0 || A[c] == 0;
    const double hP = h[c] * (double)!noIce0, hsP = hs[c] * (double)!noIce0;
    const double AP = A[c] * (double)!noIce0;
    const int isIce = hP > 0;
    const double regA = sqrt(AP * AP + Areg), rregA = 1 / regA;
    double hIA = isIce ? hP * rregA : 0;
    const double rhIA = AP / sqrt(hP * hP + hreg);
    const double hSA = isIce ? hsP * rregA : 0;
    hIA = npmax(hIA, 0.05);
    const double Tfrz = (tF + dtF * ocS[c]) + q.c2k;
    double Tsum = 0, Fio = 0, Fia = 0, pen = 0;
    for (int l = 0; l < nITC; ++l) {
      const double p = (double)(2 * (l + 1) - 1) * rn;
      double To, Fo, Fa, ps;
      solve4temp_cell(&q, hIA * p, hSA * p, TS[c], Tfrz, LW[c], AT_[c], wS[c], fC[c], SW[c],
                      aqh[c], &To, &Fo, &Fa, &ps);
      if (l == 0) {
        Tsum = To;
        Fio = Fo;
        Fia = Fa;
        pen = ps;
      } else {
        Tsum = Tsum + To;
        Fio = Fio + Fo;
        Fia = Fia + Fa;
        pen = pen + ps;
      }
    }
    const double TSurf = Tsum * rn;
    Fio = (Fio * rn) * AP;
    Fia = (Fia * rn) * AP;
    pen = (pen * rn) * AP;
    const int cold = AP > 0 && Tsum / nITC < q.c2k;
    const double snowRate = (cold ? snowfall[c] + precip[c] : snowfall[c]) * r2s;
    const double rainToSea = cold ? 0 : precip[c];
    const double snowAcc = (snowRate * AP) * dt;
    double gUnder = Fio * qi;
    gUnder = AP == 0 ? 0 : gUnder;
    const double potMeltRate = -Fia * qs;
    const double potMelt = potMeltRate * dt;
    const int allMelt = potMelt >= hsP;
    const double melt = allMelt ? hsP : potMelt;
    const double meltRate = allMelt ? melt * rdt : potMeltRate;
    const double toMelt = allMelt ? (-hsP * rdt) / qs : Fia;
    Fia = Fia - toMelt;
    const double gSurf = Fia * qi;
    const double gNet = (gUnder + gSurf) * rregA;
    const double sth = npmax(theta[c], Tfrz);
    const double mltf = 1 + (McP - 1) / (1 + exp((AP - t0) * t1));
    const double Foi = (-t2 * (sth - Tfrz)) * mltf;
    const double gML = Foi * qi, gOW = qi * Qnet[c];
    const double dh = (gNet * AP + gOW * (1 - AP)) + gML;
    const double dhs = (snowRate - meltRate) * AP;
    const double half = 0.5 * rhIA;
    double dAoa = (half * gOW) * (1 - AP);
    const int grow = gOW > 0 && (AP > 0 || dh > 0);
    dAoa = grow ? gOW * (1 - AP) : dAoa;
    dAoa = (grow && fC[c] < 0) ? dAoa * rh0s : dAoa * rh0;
    const double dAoi = gML <= 0 ? half * gML : 0;
    const double dAia = (gNet <= 0 && hP > 0) ? (half * gNet) * AP : 0;
    const double dA = (dAoa + dAoi) + dAia;
    double a = AP + (dA * M[c]) * dt;
    double hh = hP + (dh * M[c]) * dt;
    double hhs = hsP + (dhs * M[c]) * dt;
    a = npclip(a, 0, 1);
    hh = npmax(hh, 0);
    hhs = npmax(hhs, 0);
    const int noIce = hh == 0 || a == 0;
    a = a * (double)!noIce;
    hhs = hhs * (double)!noIce;
    hh = hh * (double)!noIce;
    const double hsub = (hhs * rS + hh * rI) * rrs;
    const double flood = npmax(0, hsub - hh);
    hh = hh + flood;
    hhs = hhs - flood * rI2rS;
    const double rhh = 1 / sqrt(hh * hh + hreg);
    const double qsw = pen * AP + Qsw[c] * (1 - AP);
    const double dH = hh - hP;
    const double meltedSnow = (hsP + snowAcc) - hhs;
    const double eIce = dH / qi;
    const double eOut = (AP * ((Fia + Fio) + pen) + (1 - AP) * Qnet[c]) * dt;
    const double qnet = (eOut - eIce) * rdt;
    const double sal = (ocS[c] > 0 && ocS[c] > sIr) ? (1 - sIr / ocS[c]) : 1;
    const double fwIce = (-dH * rI2rF) * sal;
    const double t0s = npmin(sIr, ocS[c]);
    const double saltflux = ((((dH + osh[c]) * t0s) * M[c]) * rI) * rdt;
    const double fwSnow = meltedSnow / r2s;
    const double e = (M[c] * (((((-evap[c] - precip[c]) * (1 - AP)) - rainToSea * AP) - runoff[c]) -
                              (fwIce + fwSnow) / dt)) *
                         rhoF +
                     (M[c] * (osh[c] * rI + oshs[c] * rS)) * rdt;
    h[c] = hh;
    hs[c] = hhs;
    A[c] = a;
    TS[c] = TSurf;
    EmPmR[c] = e;
    fss[c] = (e * ocS[c]) / rhoF + saltflux;
    Qsw[c] = qsw;
    Qnet[c] = qnet;
    load[c] = hh * rI + hhs * rS;
    penO[c] = pen;
    rh[c] = rhh;
  }
}

/* The whole step (seaice_global_4deg.py:874-932, after the forcing copies). */
int seaice_step_ref(int NX, int NY, int nf, const char* const* fnames, double* const* fields,
                    int ns, const char* const* snames, const double* svals, int nsub) {
  Step st = {NX, NY, nf, ns, fnames, fields, snames, svals, 0};
  g_missing[0] = 0;
  Step* S = &st;
  mass_area(S);
  wind_forcing(S);
  strength(S);
  if (S->missing) return -2;
  const int rc = ice_velocities(S, nsub);
  if (rc) return rc;
  ocean_stress(S);
  advection(S);
  clean_up_and_ridging(S);
  growth(S);
  if (S->missing) return -2;
  const char* fills[] = {"hIceMean", "hSnowMean", "Area", "forc_salt_surface_ice", "Qnet"};
  for (int k = 0; k < 5; ++k) fill(S, F_(S, fills[k]));
  return S->missing ? -2 : 0;
}

/* veris/heat_flux_MITgcm.py:5-123 (bulkf_formula_lanl), cell by cell.  The
 * scalars aln = log(ht/zref) and rdn = karman/log(zref/zice) come from the
 * caller so that NumPy's and libm's log agree on them by construction;
 * out[] = flwupa, flha, fsha, df0dT, ust, vst, evp, ssq, dEvdT. */
int bulkf_ref(int n, const double* uw, const double* vw, const double* ta, const double* qa,
              const double* tsf, double grav, double karman, double zvir, double latvap,
              double gamma_blk, double rhoAir, double cpdair, double emis, double stefBoltz,
              double aln, double rdn, double* const out[9]) {
  const double ht = 2.0, zref = 10.0;
  const double czol = zref * karman * grav;
  for (int c = 0; c < n; ++c) {
    const double lath = latvap;
    const double us = sqrt(uw[c] * uw[c] + vw[c] * vw[c]);
    const double usm = npmax(us, 1.0);
    const double t0 = ta[c] * (1.0 + zvir * qa[c]);
    const double ssq = 3.797915 * exp(lath * (7.93252e-6 - 2.166847e-3 / tsf[c])) / 1013.0;
    const double deltap = ta[c] - tsf[c] + gamma_blk * ht;
    const double delq = qa[c] - ssq;
    const double rhn = rdn, ren = rdn;
    double ustar = rdn * usm, tstar = rhn * deltap, qstar = ren * delq;
    double rd = rdn, rh = rhn, re = ren;
    for (int it = 0; it < 5; ++it) {
      double huol = czol / (ustar * ustar) * (tstar / t0 + qstar / (1.0 / zvir + qa[c]));
      huol = npmin(fabs(huol), 10.0) * npsign(huol);
      const double stable = 0.5 + 0.5 * npsign(huol);
      const double xsq = npmax(sqrt(fabs(1.0 - 16.0 * huol)), 1.0);
      const double x = sqrt(xsq);
      const double psimh = -5.0 * huol * stable +
                           (1.0 - stable) * (2.0 * log(0.5 * (1.0 + x)) + 2.0 * log(0.5 * (1.0 + xsq)) -
                                             2.0 * atan(x) + M_PI * 0.5);
      const double psixh = -5.0 * huol * stable + (1.0 - stable) * (2.0 * log(0.5 * (1.0 + xsq)));
      rd = rdn / (1.0 + rdn * (aln - psimh) / karman);
      rh = rhn / (1.0 + rhn * (aln - psixh) / karman);
      re = rh;
      ustar = rd * usm;
      qstar = re * delq;
      tstar = rh * deltap;
    }
    const double csha = rhoAir * cpdair * us * rh * rd;
    const double clha = rhoAir * lath * us * re * rd;
    const double fsha = csha * deltap;
    const double flha = clha * delq;
    const double t = tsf[c];
    const double devdt = clha * ssq * 2.166847e-3 / (t * t);
    const double dflwupdt = 4.0 * emis * stefBoltz * pow(t, 3.0);
    const double cdn = 2.7e-3 / usm + 0.142e-3 + 0.0764e-3 * usm;
    out[0][c] = emis * stefBoltz * pow(t, 4.0);
    out[1][c] = flha;
    out[2][c] = fsha;
    out[3][c] = -dflwupdt + -csha + -lath * devdt;
    out[4][c] = rhoAir * cdn * us * uw[c];
    out[5][c] = rhoAir * cdn * us * vw[c];
    out[6][c] = -flha / lath;
    out[7][c] = ssq;
    out[8][c] = devdt;
  }
  return 0;
}
