/*
 * Scalar-loop C restatement of the Veris EVP subcycle and free-drift solver --
 * TEST INFRASTRUCTURE ONLY (a checker of the NumPy oracle, never the product).
 *
 * Settings are one double array indexed by ES_*, fields one array of pointers
 * indexed by EF_*, each a padded (NX, NY) FP64 array in C order.  Fields that a
 * configuration does not use (Fu/Fv without the coastline form factors, the
 * metric terms when they are off) must still point at zero-filled arrays.
 */
#ifndef VERIS_EVP_REF_H
#define VERIS_EVP_REF_H

enum evp_ref_setting {
  ES_NX, ES_NY, ES_NOSLIP, ES_SOBC, ES_PLAS, ES_DELTAMIN, ES_TENS, ES_PRF,
  ES_CWN, ES_CWS, ES_RHOSEA, ES_CDRAGMIN, ES_K1, ES_K2, ES_U0, ES_CBSTAR,
  ES_COAST, ES_SDCOEFF, ES_SDU0, ES_VEROS_FILL, ES_CYCLIC, ES_ADAPTIVE,
  ES_DELTATDYN, ES_AEVPCOEFF, ES_EVPALPHA, ES_AEVPALPHAMIN, ES_EVPBETA, ES_RDT,
  ES_COSWAT, ES_SINWAT, ES_RHOICE,
  ES_COUNT
};

enum evp_ref_field {
  EF_UICE, EF_VICE, EF_S1, EF_S2, EF_S12,
  EF_M, EF_MU, EF_MV, EF_MINC, EF_MINU, EF_MINV,
  EF_RDXU, EF_RDYU, EF_RDXV, EF_RDYV, EF_DXU, EF_DYU, EF_DXV, EF_DYV,
  EF_RRA, EF_RAZ, EF_RRAU, EF_RRAV,
  EF_K1C, EF_K2C, EF_K1Z, EF_K2Z,
  EF_P, EF_FC, EF_UOC, EF_VOC, EF_AREA, EF_H, EF_RLOW, EF_AW, EF_AS,
  EF_MASSC, EF_MASSU, EF_MASSV, EF_WFX, EF_WFY, EF_FU, EF_FV,
  EF_COUNT
};

/* nsub EVP substeps from the state in fields; returns 0, or -1 when out of memory */
int evp_solver_ref(const double* settings, const double* const* fields, int nsub, double* u,
                   double* v, double* s1, double* s2, double* s12);

/* the same with the residual norms of every substep (resU[k], resSig[k]) */
int evp_solver_residuals_ref(const double* settings, const double* const* fields, int nsub,
                             double* u, double* v, double* s1, double* s2, double* s12,
                             double* resU, double* resSig);

/* closed-form free drift; returns 0, or -1 when out of memory */
int freedrift_ref(const double* settings, const double* const* fields, double* u, double* v);

/* The whole sea-ice step on fields and settings passed by name (step_ref.c);
 * the fields are updated in place.  Returns 0, -1 when out of memory, -2 when
 * a field or setting is missing, -3 when neither solver is selected. */
int seaice_step_ref(int NX, int NY, int nf, const char* const* fnames, double* const* fields,
                    int ns, const char* const* snames, const double* svals, int nsub);

/* name of the field or setting a failed seaice_step_ref call did not find */
const char* step_ref_missing(void);

/* The open-ocean bulk fluxes on n cells (step_ref.c); out[9] = flwupa, flha,
 * fsha, df0dT, ust, vst, evp, ssq, dEvdT. */
int bulkf_ref(int n, const double* uw, const double* vw, const double* ta, const double* qa,
              const double* tsf, double grav, double karman, double zvir, double latvap,
              double gamma_blk, double rhoAir, double cpdair, double emis, double stefBoltz,
              double aln, double rdn, double* const out[9]);

/* OpenMP threads of the cell loops (n > 0 sets them); returns the count in
 * effect (1 in a build without OpenMP) */
int evp_ref_set_threads(int n);

#endif
