/*
 * Scalar-loop C restatement of the Veris EVP subcycle -- TEST INFRASTRUCTURE ONLY.
 *
 * A second, independent CPU restatement of veris/evp_solver.py:22-319 and the
 * routines it calls (veris/dynamics_routines.py:21-302, veris/averaging.py:4-19,
 * veris/fill_overlap.py:6-23), written cell by cell with explicit wrapped
 * indices instead of array rolls.  Only the tests load it, to check the NumPy
 * oracle (oracle/veris_np.py) bit for bit: both evaluate every expression in
 * the reference's association order in IEEE FP64, so any transcription
 * mistake in either shows up as a difference.  Built with -ffp-contract=off
 * (no fused multiply-add) by oracle/c/Makefile.
 *
 * Arrays are padded (NX, NY) in C order (i = axis 0, j = axis 1 contiguous);
 * neighbour (i-1) of i = 0 is NX-1, as numpy.roll does.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "evp_ref.h"

#ifdef _OPENMP
#include <omp.h>
#endif

int evp_ref_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

typedef struct {
  int NX, NY;
  const double* s;
  const double* const* f;
} Ctx;

static int wrap(int i, int n) { return ((i % n) + n) % n; }
#define AT(x, y) ((size_t)wrap((x), C->NX) * (size_t)C->NY + (size_t)wrap((y), C->NY))
#define F(k, x, y) (C->f[k][AT(x, y)])
#define S(k) (C->s[k])
/* every FOR_CELLS body writes only its own cell: rows are split over OpenMP
 * threads on grids large enough to pay for it (bits do not depend on it) */
#define FOR_CELLS(C)                                                                 \
  _Pragma("omp parallel for schedule(static) if ((C)->NX * (C)->NY >= 16384)")       \
  for (int i = 0; i < (C)->NX; ++i) for (int j = 0; j < (C)->NY; ++j)

static double sgn(double x) { return (double)((x > 0) - (x < 0)); }

/* veris/dynamics_routines.py:143-217 */
static void strainrates(const Ctx* C, const double* u, const double* v, double* e11,
                        double* e22, double* e12) {
  FOR_CELLS(C) {
    const size_t c = AT(i, j);
    const double dudx = (u[AT(i + 1, j)] - u[c]) * F(EF_RDXU, i, j);
    const double uave = (u[AT(i + 1, j)] + u[c]) * 0.5;
    const double dvdy = (v[AT(i, j + 1)] - v[c]) * F(EF_RDYV, i, j);
    const double vave = (v[AT(i, j + 1)] + v[c]) * 0.5;
    e11[c] = (dudx + vave * F(EF_K2C, i, j)) * F(EF_MINC, i, j);
    e22[c] = (dvdy + uave * F(EF_K1C, i, j)) * F(EF_MINC, i, j);
    const double dudyz = (u[c] - u[AT(i, j - 1)]) * F(EF_RDYU, i, j);
    const double uavez = (u[c] + u[AT(i, j - 1)]) * 0.5;
    const double dvdxz = (v[c] - v[AT(i - 1, j)]) * F(EF_RDXV, i, j);
    const double vavez = (v[c] + v[AT(i - 1, j)]) * 0.5;
    const double mskZ = (F(EF_M, i, j) * F(EF_M, i - 1, j)) * (F(EF_M, i, j - 1) * F(EF_M, i - 1, j - 1));
    double z = (0.5 * (((dudyz + dvdxz) - F(EF_K1Z, i, j) * vavez) - F(EF_K2Z, i, j) * uavez)) * mskZ;
    if (S(ES_NOSLIP) != 0.0) {
      const double hFacU = F(EF_MU, i, j) - F(EF_MU, i, j - 1);
      const double hFacV = F(EF_MV, i, j) - F(EF_MV, i - 1, j);
      z = z + ((((2.0 * uavez) * F(EF_RDYU, i, j)) * hFacU) + (((2.0 * vavez) * F(EF_RDXV, i, j)) * hFacV));
    }
    if (S(ES_NOSLIP) != 0.0 && S(ES_SOBC) != 0.0) {
      double hU = (F(EF_MU, i, j) - F(EF_MU, i - 1, j)) / 3.0;
      double hV = (F(EF_MV, i, j) - F(EF_MV, i, j - 1)) / 3.0;
      hU = hU * (F(EF_MU, i - 2, j) * F(EF_MU, i - 1, j) + F(EF_MU, i + 1, j) * F(EF_MU, i, j));
      hV = hV * (F(EF_MV, i, j - 2) * F(EF_MV, i, j - 1) + F(EF_MV, i, j + 1) * F(EF_MV, i, j));
      const double tu = (F(EF_RDYU, i, j) *
                         (((6.0 * uavez) - u[AT(i - 2, j)] * F(EF_MU, i - 1, j)) - u[AT(i + 1, j)] * F(EF_MU, i, j))) *
                        hU;
      const double tv = (F(EF_RDXV, i, j) *
                         (((6.0 * vavez) - v[AT(i, j - 2)] * F(EF_MV, i, j - 1)) - v[AT(i, j + 1)] * F(EF_MV, i, j))) *
                        hV;
      z = z + 0.5 * (tu + tv);
    }
    e12[c] = z;
  }
}

/* veris/dynamics_routines.py:220-268 -> zeta, press (eta is not used downstream) */
static void viscosities(const Ctx* C, const double* e11, const double* e22, const double* e12,
                        double* zeta, double* press, double* q) {
  const double rpd = 1.0 / (S(ES_PLAS) * S(ES_PLAS));
  FOR_CELLS(C) q[AT(i, j)] = F(EF_RAZ, i, j) * (e12[AT(i, j)] * e12[AT(i, j)]);
  FOR_CELLS(C) {
    const size_t c = AT(i, j);
    const double qa = q[c] + q[AT(i + 1, j)];
    const double qb = q[AT(i, j + 1)] + q[AT(i + 1, j + 1)];
    const double e12Csq = (0.25 * F(EF_RRA, i, j)) * (qa + qb);
    const double ep = e11[c] + e22[c], em = e11[c] - e22[c];
    const double deltaSq = ep * ep + rpd * (em * em + 4.0 * e12Csq);
    const double deltaC = sqrt(deltaSq);
    const double dreg = deltaC + S(ES_DELTAMIN);
    const double P = F(EF_P, i, j);
    const double t = S(ES_TENS), r = S(ES_PRF);
    zeta[c] = (0.5 * (P * (1 + t))) / dreg;
    press[c] = (1 * ((P * (1 - r)) + (((2.0 * zeta[c]) * deltaC) * r) / (1 + t))) * (1 - t);
  }
}

/* veris/averaging.py:4-19 (noSlip normaliser, the only one callers use) */
static double c_to_z(const Ctx* C, const double* a, int i, int j) {
  const double n = (F(EF_M, i, j) + F(EF_M, i - 1, j)) + (F(EF_M, i, j - 1) + F(EF_M, i - 1, j - 1));
  const double norm = n > 0 ? 1.0 / n : 0.0;
  return norm * ((a[AT(i, j)] + a[AT(i - 1, j)]) + (a[AT(i, j - 1)] + a[AT(i - 1, j - 1)]));
}

/* veris/dynamics_routines.py:21-52 */
static void ocean_drag(const Ctx* C, const double* u, const double* v, double* cd) {
  FOR_CELLS(C) {
    const double D = (F(EF_FC, i, j) < 0 ? S(ES_CWS) : S(ES_CWN)) * S(ES_RHOSEA);
    const double du = (u[AT(i, j)] - F(EF_UOC, i, j)) * F(EF_MINU, i, j);
    const double dup = (u[AT(i + 1, j)] - F(EF_UOC, i + 1, j)) * F(EF_MINU, i + 1, j);
    const double dv = (v[AT(i, j)] - F(EF_VOC, i, j)) * F(EF_MINV, i, j);
    const double dvp = (v[AT(i, j + 1)] - F(EF_VOC, i, j + 1)) * F(EF_MINV, i, j + 1);
    const double w = 0.25 * ((du + dup) * (du + dup) + (dv + dvp) * (dv + dvp));
    const double c = (D * D) * w > S(ES_CDRAGMIN) * S(ES_CDRAGMIN) ? D * sqrt(w) : S(ES_CDRAGMIN);
    cd[AT(i, j)] = c * F(EF_M, i, j);
  }
}

/* veris/dynamics_routines.py:55-96 */
static void basal_drag(const Ctx* C, const double* u, const double* v, double* cb) {
  FOR_CELLS(C) {
    const double um = u[AT(i, j)] * F(EF_MINU, i, j), ump = u[AT(i + 1, j)] * F(EF_MINU, i + 1, j);
    const double vm = v[AT(i, j)] * F(EF_MINV, i, j), vmp = v[AT(i, j + 1)] * F(EF_MINV, i, j + 1);
    double t = 0.25 * (((um * um + ump * ump) + vm * vm) + vmp * vmp);
    t = S(ES_K2) / sqrt(t + S(ES_U0) * S(ES_U0));
    const double A = F(EF_AREA, i, j);
    const double hc = (fabs(F(EF_RLOW, i, j)) * A) / S(ES_K1);
    const double val = ((t * log(exp(10.0 * (F(EF_H, i, j) - hc)) + 1.0)) * (1.0 / 10.0)) *
                       exp(-S(ES_CBSTAR) * (1.0 - A));
    cb[AT(i, j)] = A > 0.01 ? val : 0.0;
  }
}

/* veris/dynamics_routines.py:98-140 at one u-point / v-point */
static void side_drag(const Ctx* C, const double* u, const double* v, const double* spd, int i,
                      int j, double* sdU, double* sdV) {
  const double spU = F(EF_AW, i, j) > 0 ? 0.5 * (spd[AT(i, j)] + spd[AT(i - 1, j)]) : 0;
  const double spV = F(EF_AS, i, j) > 0 ? 0.5 * (spd[AT(i, j)] + spd[AT(i, j - 1)]) : 0;
  double mU = ((2 - F(EF_MU, i - 1, j)) - F(EF_MU, i + 1, j)) * F(EF_MU, i, j);
  double mV = ((2 - F(EF_MV, i, j - 1)) - F(EF_MV, i, j + 1)) * F(EF_MV, i, j);
  if (S(ES_COAST) != 0.0) {
    mU = F(EF_FU, i, j);
    mV = F(EF_FV, i, j);
  }
  (void)u;
  (void)v;
  *sdU = ((F(EF_MASSU, i, j) * S(ES_SDCOEFF)) * mU) / (spU + S(ES_SDU0));
  *sdV = ((F(EF_MASSV, i, j) * S(ES_SDCOEFF)) * mV) / (spV + S(ES_SDU0));
}

/* veris/fill_overlap.py:6-18 (Veros single-process fill, or the plugin's own) */
static void fill(const Ctx* C, double* a) {
  const int NX = C->NX, NY = C->NY;
  if (S(ES_VEROS_FILL) != 0.0) {
    if (S(ES_CYCLIC) == 0.0) return;
    for (int j = 0; j < NY; ++j) {
      for (int k = 0; k < 2; ++k) {
        a[(size_t)(NX - 2 + k) * NY + j] = a[(size_t)(2 + k) * NY + j];
        a[(size_t)k * NY + j] = a[(size_t)(NX - 4 + k) * NY + j];
      }
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

/* One substep, veris/evp_solver.py:22-254 (residual branch aside). */
static void substep(const Ctx* C, double* u, double* v, double* s1, double* s2, double* s12,
                    double** w) {
  const int NX = C->NX, NY = C->NY;
  const size_t n = (size_t)NX * NY;
  double *e11 = w[0], *e22 = w[1], *e12 = w[2], *zeta = w[3], *press = w[4], *q = w[5];
  double *aC = w[6], *aZ = w[7], *cd = w[8], *cb = w[9], *fv = w[10], *fu = w[11];
  double *dx = w[12], *dy = w[13], *un = w[14], *vn = w[15], *t11 = w[16], *t12 = w[17];
  double *t22 = w[18], *t21 = w[19], *spd = w[20];
  strainrates(C, u, v, e11, e22, e12);
  viscosities(C, e11, e22, e12, zeta, press, q);
  const int adaptive = S(ES_ADAPTIVE) != 0.0;
  const double cfac = adaptive ? S(ES_DELTATDYN) * 4 * ((M_PI * S(ES_AEVPCOEFF)) * (M_PI * S(ES_AEVPCOEFF))) : 0;
  FOR_CELLS(C) {
    const size_t c = AT(i, j);
    double a = S(ES_EVPALPHA), den1 = 1 / S(ES_EVPALPHA);
    if (adaptive) {
      a = sqrt(((zeta[c] * cfac) / fmax(F(EF_MASSC, i, j), 1e-4)) * F(EF_RRA, i, j)) * F(EF_M, i, j);
      a = fmax(a, S(ES_AEVPALPHAMIN));
      den1 = 1.0 / a;
    }
    aC[c] = a;
    const double ep = e11[c] + e22[c], em = e11[c] - e22[c];
    const double divergence = ((2 * zeta[c]) * ep) - press[c];
    const double tension = (2 * zeta[c]) * em;
    s1[c] = (((s1[c] * (a - 1)) + divergence) * den1) * F(EF_M, i, j);
    s2[c] = (((s2[c] * (a - 1)) + tension * 0.25) * den1) * F(EF_M, i, j);
  }
  FOR_CELLS(C) {
    const size_t c = AT(i, j);
    double az = S(ES_EVPALPHA), den2 = 1 / S(ES_EVPALPHA);
    if (adaptive) {
      az = 0.5 * (0.5 * (aC[c] + aC[AT(i, j - 1)]) + 0.5 * (aC[AT(i - 1, j)] + aC[AT(i - 1, j - 1)]));
      den2 = 1.0 / az;
    }
    aZ[c] = az;
    const double shear = (2 * c_to_z(C, zeta, i, j)) * e12[c];
    s12[c] = ((s12[c] * (az - 1)) + shear * 0.25) * den2;
  }
  /* stress divergence: sig11 = 0.5 (s1 + s2), sig22 = 0.5 (s1 - s2) */
  for (size_t c = 0; c < n; ++c) {
    const size_t i = c / NY, j = c % NY;
    t11[c] = (0.5 * (s1[c] + s2[c])) * F(EF_DYV, (int)i, (int)j);
    t22[c] = (0.5 * (s1[c] - s2[c])) * F(EF_DXU, (int)i, (int)j);
    t12[c] = s12[c] * F(EF_DXV, (int)i, (int)j);
    t21[c] = s12[c] * F(EF_DYU, (int)i, (int)j);
  }
  FOR_CELLS(C) {
    const size_t c = AT(i, j);
    dx[c] = (((t11[c] - t11[AT(i - 1, j)]) - t12[c]) + t12[AT(i, j + 1)]) * F(EF_RRAU, i, j);
    dy[c] = (((t22[c] - t22[AT(i, j - 1)]) - t21[c]) + t21[AT(i + 1, j)]) * F(EF_RRAV, i, j);
  }
  ocean_drag(C, u, v, cd);
  basal_drag(C, u, v, cb);
  FOR_CELLS(C) {
    const size_t c = AT(i, j);
    const double mf = (F(EF_MASSC, i, j) * F(EF_FC, i, j)) * 0.5;
    fv[c] = mf * (v[c] + v[AT(i, j + 1)]);
    fu[c] = mf * (u[c] + u[AT(i + 1, j)]);
    const double su = u[c] + u[AT(i + 1, j)], sv = v[c] + v[AT(i, j + 1)];
    spd[c] = 0.5 * sqrt(su * su + sv * sv);
  }
  const double rdt = S(ES_RDT), cosW = S(ES_COSWAT), sinW = S(ES_SINWAT);
  FOR_CELLS(C) {
    const size_t c = AT(i, j);
    const double lmU = F(EF_MASSU, i, j) != 0 ? 1 : F(EF_MASSU, i, j);
    const double lmV = F(EF_MASSV, i, j) != 0 ? 1 : F(EF_MASSV, i, j);
    const double sg = sgn(F(EF_FC, i, j));
    /* uOcean - uIce and its shifted copies (the turning term) */
#define DU(x, y) (F(EF_UOC, x, y) - u[AT(x, y)])
#define DV(x, y) (F(EF_VOC, x, y) - v[AT(x, y)])
    const double X = (cd[c] * 0.5) * (DU(i, j) + DU(i, j + 1)) +
                     (cd[AT(i - 1, j)] * 0.5) * (DU(i - 1, j) + DU(i - 1, j + 1));
    const double Y = (cd[c] * 0.5) * (DV(i, j) + DV(i + 1, j)) +
                     (cd[AT(i, j - 1)] * 0.5) * (DV(i, j - 1) + DV(i + 1, j - 1));
#undef DU
#undef DV
    double fx = F(EF_WFX, i, j) +
                ((((0.5 * (cd[c] + cd[AT(i - 1, j)])) * cosW) * F(EF_UOC, i, j)) - (((sg * sinW) * 0.5) * X) * lmU) *
                    F(EF_AW, i, j);
    double fy = F(EF_WFY, i, j) +
                ((((0.5 * (cd[c] + cd[AT(i, j - 1)])) * cosW) * F(EF_VOC, i, j)) + (((sg * sinW) * 0.5) * Y) * lmV) *
                    F(EF_AS, i, j);
    fx = fx + 0.5 * (fv[c] + fv[AT(i - 1, j)]);
    fy = fy - 0.5 * (fu[c] + fu[AT(i, j - 1)]);
    double bU = S(ES_EVPBETA), bV = S(ES_EVPBETA);
    if (adaptive) {
      bU = 0.5 * (aC[c] + aC[AT(i - 1, j)]);
      bV = 0.5 * (aC[c] + aC[AT(i, j - 1)]);
    }
    const double bfU = bU * rdt, bfV = bV * rdt;
    const double bp1U = bfU + rdt, bp1V = bfV + rdt;
    double denU = F(EF_MASSU, i, j) * bp1U +
                  F(EF_AW, i, j) * ((0.5 * (cd[c] + cd[AT(i - 1, j)])) * cosW + 0.5 * (cb[c] + cb[AT(i - 1, j)]));
    double denV = F(EF_MASSV, i, j) * bp1V +
                  F(EF_AS, i, j) * ((0.5 * (cd[c] + cd[AT(i, j - 1)])) * cosW + 0.5 * (cb[c] + cb[AT(i, j - 1)]));
    denU = denU == 0 ? 1 : denU;
    denV = denV == 0 ? 1 : denV;
    if (S(ES_NOSLIP) == 0.0) {
      double sdU, sdV;
      side_drag(C, u, v, spd, i, j, &sdU, &sdV);
      denU = denU + sdU;
      denV = denV + sdV;
    }
    un[c] = (F(EF_MU, i, j) * ((((bfU * F(EF_MASSU, i, j)) * u[c] + (F(EF_MASSU, i, j) * rdt) * F(EF_UICE, i, j)) + fx) + dx[c])) / denU;
    vn[c] = (F(EF_MV, i, j) * ((((bfV * F(EF_MASSV, i, j)) * v[c] + (F(EF_MASSV, i, j) * rdt) * F(EF_VICE, i, j)) + fy) + dy[c])) / denV;
  }
  fill(C, un);
  fill(C, vn);
  memcpy(u, un, n * sizeof(double));
  memcpy(v, vn, n * sizeof(double));
  (void)aZ;
}

int evp_solver_ref(const double* settings, const double* const* fields, int nsub, double* u,
                   double* v, double* s1, double* s2, double* s12) {
  Ctx ctx = {(int)settings[ES_NX], (int)settings[ES_NY], settings, fields};
  const Ctx* C = &ctx;
  const size_t n = (size_t)C->NX * C->NY;
  memcpy(u, fields[EF_UICE], n * sizeof(double));
  memcpy(v, fields[EF_VICE], n * sizeof(double));
  memcpy(s1, fields[EF_S1], n * sizeof(double));
  memcpy(s2, fields[EF_S2], n * sizeof(double));
  memcpy(s12, fields[EF_S12], n * sizeof(double));
  double* w[21];
  for (int k = 0; k < 21; ++k) {
    w[k] = (double*)calloc(n, sizeof(double));
    if (!w[k]) {
      for (int m = 0; m < k; ++m) free(w[m]);
      return -1;
    }
  }
  for (int k = 0; k < nsub; ++k) substep(C, u, v, s1, s2, s12, w);
  for (int k = 0; k < 21; ++k) free(w[k]);
  return 0;
}

/* Residual norms of one substep over the interior, veris/evp_solver.py:192-230
 * (the previous iterate's (sigma1 +- sigma2)/2 as the previous sig11/sig22).
 * Sums in index order. */
static void residuals(const Ctx* C, double* const old[5], const double* u, const double* v,
                      const double* s1, const double* s2, const double* s12, const double* aC,
                      const double* aZ, double* rsig, double* rvel) {
  const int adaptive = S(ES_ADAPTIVE) != 0.0;
  double sig = 0.0, vel = 0.0;
  for (int i = 2; i < C->NX - 2; ++i)
    for (int j = 2; j < C->NY - 2; ++j) {
      const size_t c = AT(i, j);
      const double r11 = (((0.5 * (s1[c] + s2[c])) - (0.5 * (old[2][c] + old[3][c]))) * aC[c]) * F(EF_M, i, j);
      const double r22 = (((0.5 * (s1[c] - s2[c])) - (0.5 * (old[2][c] - old[3][c]))) * aC[c]) * F(EF_M, i, j);
      const double r12 = (s12[c] - old[4][c]) * aZ[c];
      const double bU = adaptive ? 0.5 * (aC[c] + aC[AT(i - 1, j)]) : S(ES_EVPBETA);
      const double bV = adaptive ? 0.5 * (aC[c] + aC[AT(i, j - 1)]) : S(ES_EVPBETA);
      const double ru = (F(EF_MU, i, j) * (u[c] - old[0][c])) * bU;
      const double rv = (F(EF_MV, i, j) * (v[c] - old[1][c])) * bV;
      sig += (r11 * r11 + r22 * r22) + r12 * r12;
      vel += ru * ru + rv * rv;
    }
  *rsig = sig;
  *rvel = vel;
}

/* The subcycle with per-substep residual norms (computeEvpResidual). */
int evp_solver_residuals_ref(const double* settings, const double* const* fields, int nsub,
                             double* u, double* v, double* s1, double* s2, double* s12,
                             double* resU, double* resSig) {
  Ctx ctx = {(int)settings[ES_NX], (int)settings[ES_NY], settings, fields};
  const Ctx* C = &ctx;
  const size_t n = (size_t)C->NX * C->NY;
  memcpy(u, fields[EF_UICE], n * sizeof(double));
  memcpy(v, fields[EF_VICE], n * sizeof(double));
  memcpy(s1, fields[EF_S1], n * sizeof(double));
  memcpy(s2, fields[EF_S2], n * sizeof(double));
  memcpy(s12, fields[EF_S12], n * sizeof(double));
  double* w[26];
  for (int k = 0; k < 26; ++k) {
    w[k] = (double*)calloc(n, sizeof(double));
    if (!w[k]) {
      for (int m = 0; m < k; ++m) free(w[m]);
      return -1;
    }
  }
  double* const old[5] = {w[21], w[22], w[23], w[24], w[25]};
  double* const cur[5] = {u, v, s1, s2, s12};
  for (int k = 0; k < nsub; ++k) {
    for (int q = 0; q < 5; ++q) memcpy(old[q], cur[q], n * sizeof(double));
    substep(C, u, v, s1, s2, s12, w);
    residuals(C, old, u, v, s1, s2, s12, w[6], w[7], &resSig[k], &resU[k]);
  }
  for (int k = 0; k < 26; ++k) free(w[k]);
  return 0;
}

/* veris/freedrift_solver.py:5-62 */
int freedrift_ref(const double* settings, const double* const* fields, double* u, double* v) {
  Ctx ctx = {(int)settings[ES_NX], (int)settings[ES_NY], settings, fields};
  const Ctx* C = &ctx;
  const size_t n = (size_t)C->NX * C->NY;
  double* uc = (double*)malloc(n * sizeof(double));
  double* vc = (double*)malloc(n * sizeof(double));
  if (!uc || !vc) {
    free(uc);
    free(vc);
    return -1;
  }
  FOR_CELLS(C) {
    const double tx = 0.5 * (F(EF_WFX, i, j) + F(EF_WFX, i + 1, j));
    const double ty = 0.5 * (F(EF_WFY, i, j) + F(EF_WFY, i, j + 1));
    const double f = F(EF_FC, i, j);
    const double mf = (S(ES_RHOICE) * F(EF_H, i, j)) * f;
    const double uoc = 0.5 * (F(EF_UOC, i, j) + F(EF_UOC, i + 1, j));
    const double voc = 0.5 * (F(EF_VOC, i, j) + F(EF_VOC, i, j + 1));
    const double rx = -tx - mf * voc, ry = -ty + mf * uoc;
    const double r2 = rx * rx + ry * ry;
    const double rN = r2 > 0 ? sqrt(rx * rx + ry * ry) : 0;
    const double rA = r2 > 0 ? atan2(ry, rx) : 0;
    double t1 = 1 / ((f < 0 ? S(ES_CWS) : S(ES_CWN)) * S(ES_RHOSEA));
    double t2 = (t1 * t1) * (mf * mf);
    double t3 = (t1 * t1) * (rN * rN);
    double t4 = t2 * t2 + 4 * t3;
    const double sN = t3 > 0 ? sqrt(0.5 * (sqrt(t4) - t2)) : 0;
    t1 = 1 / t1;
    t2 = t1 * (sN * sN);
    t3 = mf * sN;
    t4 = t2 * t2 + t3 * t3;
    const double ang = t4 > 0 ? rA - atan2(t3, t2) : 0;
    uc[AT(i, j)] = uoc - sN * cos(ang);
    vc[AT(i, j)] = voc - sN * sin(ang);
  }
  FOR_CELLS(C) {
    u[AT(i, j)] = (0.5 * (uc[AT(i - 1, j)] + uc[AT(i, j)])) * F(EF_MU, i, j);
    v[AT(i, j)] = (0.5 * (vc[AT(i, j - 1)] + vc[AT(i, j)])) * F(EF_MV, i, j);
  }
  free(uc);
  free(vc);
  return 0;
}
