/*
 * veris_hip.h -- C ABI of the MI355X (gfx950) sea-ice step.
 *
 * Each entry point replaces one kernel of the reference plugin (cited per
 * declaration as file:line of team-ocean/veris).  The reference kernels are
 * pure functions of a Veros `state`; here the state is passed as plain
 * structs of device pointers (vh_fields) and scalars (vh_settings) and the
 * results are written into caller-owned device buffers.
 *
 * Conventions for every entry point:
 *  - every field is a device array of shape (nx+2*olx, ny+2*oly) in C order
 *    (axis 0 = x / i, axis 1 = y / j contiguous) with the element type
 *    vh_domain.dtype (VH_F64 or VH_F32);
 *  - neighbour access wraps over the padded array exactly as numpy.roll does in
 *    the reference, so halo cells hold the same values the reference computes;
 *  - outputs must not alias inputs; calls are stream-ordered and asynchronous
 *    on `stream` (a hipStream_t, NULL = legacy default stream);
 *  - no entry point allocates device memory: scratch comes from the caller's
 *    `workspace` (size from the matching *_workspace_bytes call);
 *  - the return value is VH_OK (0) or a negative VH_E* code, never an
 *    exception; vh_last_error() gives the message of the calling thread's
 *    last failure.
 */
#ifndef VERIS_HIP_H
#define VERIS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VH_ABI_VERSION 1

enum {
  VH_OK = 0,
  VH_EINVAL = -1,      /* null pointer, aliasing or bad argument value */
  VH_ESHAPE = -2,      /* unsupported nx/ny/olx/oly */
  VH_EDTYPE = -3,      /* unsupported dtype for this entry point */
  VH_EUNSUPPORTED = -4,/* setting combination not implemented */
  VH_EHIP = -5,        /* HIP runtime error (launch, copy) */
  VH_EWORKSPACE = -6   /* workspace missing or too small */
};

enum { VH_F64 = 0, VH_F32 = 1 };

/* Domain geometry.  fill_mode 0 reproduces Veros' single-process
 * enforce_boundaries (x-periodic copy when cyclic_x, nothing otherwise;
 * veris/fill_overlap.py:10-11); fill_mode 1 is the plugin's own doubly
 * periodic copy (veris/fill_overlap.py:12-17, setting veros_fill=False). */
typedef struct vh_domain {
  int32_t nx, ny;   /* interior cells */
  int32_t olx, oly; /* overlap widths; must both be 2 */
  int32_t cyclic_x; /* host grid is x-cyclic (Veros enable_cyclic_x) */
  int32_t fill_mode;
  int32_t dtype;    /* VH_F64 | VH_F32 */
  int32_t tile;     /* EVP tile shape index (1..vh_evp_tile_count()), 0 = library default */
} vh_domain;

/* Every scalar setting of veris/settings.py:4-176, by name and in order, as
 * double (ints and flags converted). */
typedef struct vh_settings {
  double deltatTherm;
  double recip_deltatTherm;
  double deltatDyn;
  double recip_deltatDyn;
  double olx;
  double oly;
  double gridcellWidth;
  double nITC;
  double recip_nITC;
  double noSlip;
  double useRelativeWind;
  double secondOrderBC;
  double extensiveFld;
  double useRealFreshWaterFlux;
  double useFreedrift;
  double useEVP;
  double evpAlpha;
  double evpBeta;
  double useAdaptiveEVP;
  double aEVPalphaMin;
  double aEvpCoeff;
  double explicitDrag;
  double nEVPsteps;
  double computeEvpResidual;
  double veros_fill;
  double use_coastline;
  double rhoIce;
  double rhoFresh;
  double rhoSea;
  double rhoAir;
  double rhoSnow;
  double recip_rhoFresh;
  double recip_rhoSea;
  double rhoIce2rhoSnow;
  double rhoIce2rhoFresh;
  double rhoFresh2rhoSnow;
  double dryIceAlb;
  double dryIceAlb_south;
  double wetIceAlb;
  double wetIceAlb_south;
  double drySnowAlb;
  double drySnowAlb_south;
  double wetSnowAlb;
  double wetSnowAlb_south;
  double wetAlbTemp;
  double lhFusion;
  double lhEvap;
  double lhSublim;
  double cpAir;
  double cpWater;
  double stefBoltz;
  double iceEmiss;
  double snowEmiss;
  double iceConduct;
  double snowConduct;
  double hCut;
  double shortwave;
  double tempFrz;
  double dtempFrz_dS;
  double saltIce_ref;
  double saltOcn_ref;
  double minLWdown;
  double maxTIce;
  double minTIce;
  double minTAir;
  double dalton;
  double Area_reg;
  double hIce_reg;
  double celsius2K;
  double stantonNr;
  double uStarBase;
  double McPheeTaperFac;
  double h0;
  double recip_h0;
  double h0_south;
  double recip_h0_south;
  double airTurnAngle;
  double waterTurnAngle;
  double sinWat;
  double cosWat;
  double wSpeedMin;
  double hIce_min;
  double Area_min;
  double airIceDrag;
  double airIceDrag_south;
  double waterIceDrag;
  double waterIceDrag_south;
  double cDragMin;
  double seaIceLoadFac;
  double gravity;
  double PlasDefCoeff;
  double deltaMin;
  double pressReplFac;
  double pStar;
  double cStar;
  double basalDragU0;
  double basalDragK1;
  double basalDragK2;
  double cBasalStar;
  double tensileStrFac;
  double CrMax;
  double sideDragCoeff;
  double sideDragU0;
  double umin_o;
  double umin_i;
  double zref;
  double ztref;
  double bolzc;
  double avogad;
  double rgas;
  double mwdair;
  double mwwv;
  double rdair;
  double rwv;
  double zvir;
  double cpdair;
  double cpwv;
  double cpvir;
  double karman;
  double latvap;
  double p0;
  double cappa;
  double zzsice;
  double ch;
  double ce;
  double eps2;
  double emissivity;
  double ocean_emissivity;
  double snow_emissivity;
  double ice_emissivity;
  double tf0kel;
  double gamma_blk;
  double ocean_albedo;
  double ice_albedo;
} vh_settings;

/* One device pointer per variable of veris/variables.py:5-106, by name and in
 * order.  Entry points read only the fields they need; a NULL k1AtC, k2AtC,
 * k1AtZ or k2AtZ is read as an all-zero field. */
typedef struct vh_fields {
  const void* hIceMean;
  const void* hSnowMean;
  const void* Area;
  const void* TSurf;
  const void* SeaIceMassC;
  const void* SeaIceMassU;
  const void* SeaIceMassV;
  const void* SeaIceStrength;
  const void* os_hIceMean;
  const void* os_hSnowMean;
  const void* AreaW;
  const void* AreaS;
  const void* uIce;
  const void* vIce;
  const void* sigma1;
  const void* sigma2;
  const void* sigma12;
  const void* WindForcingX;
  const void* WindForcingY;
  const void* recip_hIceMean;
  const void* SeaIceLoad;
  const void* IcePenetSW;
  const void* uOcean;
  const void* vOcean;
  const void* theta;
  const void* ocSalt;
  const void* Qnet;
  const void* OceanStressU;
  const void* OceanStressV;
  const void* saltflux;
  const void* R_low;
  const void* ssh_an;
  const void* Qsw;
  const void* uWind;
  const void* vWind;
  const void* wSpeed;
  const void* surfPress;
  const void* SWdown;
  const void* LWdown;
  const void* ATemp;
  const void* aqh;
  const void* precip;
  const void* snowfall;
  const void* evap;
  const void* runoff;
  const void* EmPmR;
  const void* maskInC;
  const void* maskInU;
  const void* maskInV;
  const void* iceMask;
  const void* iceMaskU;
  const void* iceMaskV;
  const void* k1AtC;
  const void* k2AtC;
  const void* k1AtZ;
  const void* k2AtZ;
  const void* Fu;
  const void* Fv;
  const void* fCori;
  const void* dxC;
  const void* dyC;
  const void* dxG;
  const void* dyG;
  const void* dxU;
  const void* dyU;
  const void* dxV;
  const void* dyV;
  const void* recip_dxC;
  const void* recip_dyC;
  const void* recip_dxG;
  const void* recip_dyG;
  const void* recip_dxU;
  const void* recip_dyU;
  const void* recip_dxV;
  const void* recip_dyV;
  const void* rA;
  const void* rAu;
  const void* rAv;
  const void* rAz;
  const void* recip_rA;
  const void* recip_rAu;
  const void* recip_rAv;
  const void* recip_rAz;
} vh_fields;

/* Outputs of vh_growth, the 11-tuple of veris/growth.py:385-397. */
typedef struct vh_growth_out {
  void* hIceMean;
  void* hSnowMean;
  void* Area;
  void* TSurf;
  void* EmPmR;
  void* forc_salt_surface;
  void* Qsw;
  void* Qnet;
  void* SeaIceLoad;
  void* IcePenetSW;
  void* recip_hIceMean;
} vh_growth_out;

/* ABI version (VH_ABI_VERSION) of the loaded library. */
int vh_version(void);

/* Copies the calling thread's last error message into buf (NUL-terminated);
 * returns its full length. */
int vh_last_error(char* buf, size_t len);

/* SeaIceMass (veris/area_mass.py:17-28): massC = rhoIce*h + rhoSnow*hs and its
 * two-point averages to u- and v-points. */
int vh_seaice_mass(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                   void* SeaIceMassC, void* SeaIceMassU, void* SeaIceMassV, void* stream);

/* AreaWS (veris/area_mass.py:5-14). */
int vh_area_ws(const vh_domain* d, const vh_settings* s, const vh_fields* f,
               void* AreaW, void* AreaS, void* stream);

/* SeaIceStrength (veris/dynamics_routines.py:5-18).  pStarField may be NULL
 * (scalar setting pStar) or a per-cell P* field of the domain's dtype. */
int vh_seaice_strength(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                       const void* pStarField, void* SeaIceStrength, void* stream);

/* WindForcingXY with tauXY (veris/dynsolver.py:7-84). */
int vh_wind_forcing_xy(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                       void* WindForcingX, void* WindForcingY, void* stream);

/* Number of EVP tile shapes (valid vh_domain.tile values are 1..count). */
int vh_evp_tile_count(void);

/* Launch shape of EVP tile index `tile`: ri x rj threads, one per cell of a
 * (ri-4k) x (rj-4k) tile plus a 2k-cell ring, running k substeps per launch.
 * k = 1 reproduces the reference on every cell.  k > 1 reads each filled
 * overlap through the interior cells the fill copies from: on inputs for
 * which vh_evp_overlaps_periodic gives 1 (every domain without a fill), it
 * keeps the velocities on every cell and the stresses on every cell except
 * the filled overlap columns (rows), where it gives periodic values while the
 * reference's are numpy.roll wrap-around artefacts (they need the post-fill
 * velocities of both edges within one substep, and nothing else reads them). */
int vh_evp_tile_shape(int32_t tile, int32_t* ri, int32_t* rj, int32_t* k);

/* Whether multi-substep tile shapes (k > 1) give the single-substep results
 * on these inputs: writes 1 to the device int `flag` (on `stream`, no host
 * synchronisation) when every field such a launch reads at neighbouring
 * cells holds, in each filled overlap, the value of the cell the fill copies
 * from (bit for bit): uIce, vIce, uOcean, vOcean, the six masks, the ten
 * strain/stress metrics, SeaIceStrength, fCori, SeaIceMassC, the metric terms
 * when set, and with basal drag R_low, Area and hIceMean; and sigma1, sigma2
 * in overlap column (row) 1 and sigma12 in column (row) NX-2 (NY-2), the only
 * overlap stresses the interior reads.  Otherwise writes 0.  A domain
 * without a filled overlap always gets 1.  No counterpart in the reference:
 * it guards the multi-substep launches of this library. */
int vh_evp_overlaps_periodic(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                             int32_t* flag, void* stream);

/* Workspace bytes vh_evp_solver needs for this domain. */
int vh_evp_workspace_bytes(const vh_domain* d, size_t* bytes);

/* evp_solver / IceVelocities (veris/evp_solver.py:257-319, body :22-254,
 * dispatch veris/dynsolver.py:87-102): n_substeps fused EVP substeps starting
 * from f->uIce, f->vIce, f->sigma1, f->sigma2, f->sigma12 (f->uIce, f->vIce are
 * also the previous-step velocities uIceNm1, vIceNm1); the final state is
 * written to the five outputs.  Fixed or adaptive relaxation (useAdaptiveEVP),
 * no-slip with or without secondOrderBC (single-substep tile shapes), or free
 * slip with lateral drag (noSlip = 0; f->Fu, f->Fv when use_coastline). */
int vh_evp_solver(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                  int32_t n_substeps, void* uIce, void* vIce, void* sigma1, void* sigma2,
                  void* sigma12, void* workspace, size_t workspace_bytes, void* stream);

/* Largest scratch vh_evp_solver_noalloc holds per GPU: 48 MiB, enough for the
 * 0.25-degree grid (1440 x 720) in FP64 (about 43 MB). */
#define VH_EVP_NOALLOC_MAX_BYTES (48u * 1024u * 1024u)

/* vh_evp_solver without a workspace argument: the same subcycle (the same
 * kernels and launch shapes, the same outputs bit for bit), with its scratch
 * in a device array of the library (one per GPU, VH_EVP_NOALLOC_MAX_BYTES,
 * held until the process ends) and every substep launched directly on
 * `stream` (no graph is captured or instantiated).  It allocates no device
 * memory, so callers that cannot (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc) can still run the subcycle; the
 * first call on a device looks up that device's array once.  Domains whose
 * vh_evp_workspace_bytes exceeds the maximum return VH_EWORKSPACE.  All calls
 * on one GPU share its array: calls that overlap on one GPU (other streams,
 * other threads) race and must be serialised by the caller; nothing in the
 * array needs to start at zero.  No counterpart in the reference: it exists
 * for the callers named above, which cannot allocate. */
int vh_evp_solver_noalloc(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                          int32_t n_substeps, void* uIce, void* vIce, void* sigma1, void* sigma2,
                          void* sigma12, void* stream);

/* Workspace bytes vh_evp_solver_residuals needs for this domain and count. */
int vh_evp_residuals_workspace_bytes(const vh_domain* d, int32_t n_substeps, size_t* bytes);

/* The subcycle of vh_evp_solver (same outputs, bit for bit) with the residual
 * norms of computeEvpResidual (veris/evp_solver.py:192-230): per substep k,
 * resSig[k] = sum over the interior of (alphaC d sig11)^2 + (alphaC d sig22)^2
 * + (alphaZ d sigma12)^2 and resU[k] = sum of (betaU du)^2 + (betaV dv)^2, where
 * d is the change over the substep, sig11/sig22 = (sigma1 +- sigma2)/2 (the
 * reference reads the previous ones from undefined names) and alpha/beta are
 * the fixed or the substep's adaptive relaxation parameters.  resU and resSig
 * are device arrays of n_substeps FP64 values; the sums are FP64. */
int vh_evp_solver_residuals(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                            int32_t n_substeps, void* uIce, void* vIce, void* sigma1,
                            void* sigma2, void* sigma12, double* resU, double* resSig,
                            void* workspace, size_t workspace_bytes, void* stream);

/* One member of a batched ensemble: a padded (nx+4, ny+4) box starting at
 * element `offset` of every packed field. */
typedef struct vh_member {
  int32_t nx, ny;     /* interior cells (>= 4 each) */
  int32_t cyclic_x;   /* x-periodic overlap fill (Veros enable_cyclic_x) */
  int32_t fill_mode;  /* as vh_domain.fill_mode */
  int64_t offset;     /* element offset of the member's box; boxes must not overlap */
} vh_member;

/* A batched ensemble of members of uneven sizes packed into shared arrays. */
typedef struct vh_batch {
  int32_t n_members;
  int32_t dtype;            /* VH_F64 | VH_F32 */
  int32_t tile;             /* EVP tile index, 0 = library default */
  int32_t arith;            /* 0: arithmetic in dtype (FP32: Delta, zeta, press in FP64);
                               1 (VH_F32 only): FP64 arithmetic on FP32-stored fields;
                               2 (VH_F32 only): FP32-stored step constants, FP64 state
                                 (uIce, vIce, sigma1, sigma2, sigma12 fields and the five
                                 outputs are FP64) and FP64 arithmetic */
  const vh_member* members; /* host array of n_members */
  int64_t total_elems;      /* elements of every packed field */
} vh_batch;

/* Workspace bytes vh_evp_solver_batched needs for this batch (any tile). */
int vh_evp_batched_workspace_bytes(const vh_batch* b, size_t* bytes);

/* evp_solver (veris/evp_solver.py:257-319) of every member of a batch in one
 * launch sequence: `f` holds packed fields (one pointer per variable, each
 * member's box at its offset) and the five packed outputs receive the final
 * state.  Members run the same substep kernel as vh_evp_solver; with
 * dtype VH_F32 the bulk is FP32 and Delta, zeta and the replacement pressure
 * are FP64 (arith 0), or every substep computes in FP64 on the FP32 fields
 * (arith 1), or the state is FP64 and only the step constants are FP32
 * (arith 2).  No extension of the reference (which has no batching). */
int vh_evp_solver_batched(const vh_batch* b, const vh_settings* s, const vh_fields* f,
                          int32_t n_substeps, void* uIce, void* vIce, void* sigma1,
                          void* sigma2, void* sigma12, void* workspace, size_t workspace_bytes,
                          void* stream);

/* Adjoints of the EVP subcycle's inputs (outputs of vh_evp_adjoint).  Every
 * pointer is a caller-owned device field of the domain's shape and dtype and
 * must be non-NULL.  adj_uNm1/adj_vNm1 are the adjoints of the previous-step
 * velocities in the momentum equation (the fields' uIce/vIce), separate from
 * adj_uIce/adj_vIce, the adjoints of the subcycle's initial state (the same
 * arrays in the reference: add both for the total).  The water drag adjoint
 * is per cell, split by hemisphere (north: fCori >= 0); its sum over the
 * whole padded field is d(objective)/d(waterIceDrag) and d/d(waterIceDrag_south). */
typedef struct vh_evp_adjoint_out {
  void *adj_uIce, *adj_vIce, *adj_sigma1, *adj_sigma2, *adj_sigma12;
  void *adj_uNm1, *adj_vNm1;
  void *adj_SeaIceStrength, *adj_SeaIceMassC, *adj_SeaIceMassU, *adj_SeaIceMassV;
  void *adj_AreaW, *adj_AreaS, *adj_WindForcingX, *adj_WindForcingY;
  void *adj_waterDrag_north, *adj_waterDrag_south;
  /* thickness adjoint through the basal drag coefficient (accumulated when
   * basalDragK2 != 0, may be NULL otherwise); vh_pre_evp_adjoint adds it to
   * its adj_hIceMean when non-NULL */
  void* adj_hIceMean;
} vh_evp_adjoint_out;

/* Adjoints of the pre-EVP inputs (outputs of vh_pre_evp_adjoint): per-cell
 * strength parameter P*, thickness, snow thickness, and the air drag per cell
 * split by hemisphere (sums over the whole padded field give d/d(airIceDrag)
 * and d/d(airIceDrag_south)). */
typedef struct vh_pre_evp_adjoint_out {
  void *adj_pStar, *adj_hIceMean, *adj_hSnowMean, *adj_airDrag_north, *adj_airDrag_south;
} vh_pre_evp_adjoint_out;

/* Workspace bytes of vh_evp_forward_checkpointed / vh_evp_adjoint for a
 * subcycle of n_substeps with a checkpoint every ckpt_every substeps (the
 * two calls share the workspace: the adjoint reads the checkpoints). */
int vh_evp_adjoint_workspace_bytes(const vh_domain* d, int32_t n_substeps, int32_t ckpt_every,
                                   size_t* bytes);

/* evp_solver (veris/evp_solver.py:257-319) that also keeps the state every
 * ckpt_every substeps in `workspace` for a later vh_evp_adjoint with the same
 * inputs.  Results equal vh_evp_solver's.  Every setting of the subcycle is
 * supported: fixed or adaptive relaxation; no-slip boundaries (with or
 * without the second-order term) or free slip with lateral drag; with or
 * without basal drag, the water-stress turning angle (sinWat, cosWat) and the
 * metric terms (k1AtC, k2AtC, k1AtZ, k2AtZ). */
int vh_evp_forward_checkpointed(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                                int32_t n_substeps, int32_t ckpt_every, void* uIce, void* vIce,
                                void* sigma1, void* sigma2, void* sigma12, void* workspace,
                                size_t workspace_bytes, void* stream);

/* Reverse mode of the subcycle (the derivative jax.grad takes through
 * veris/evp_solver.py:22-319 in the reference): given the adjoints of the
 * final u, v, sigma1, sigma2, sigma12 (any may be NULL = zero), writes the
 * adjoints of every subcycle input into `out`.  Recomputes each segment from
 * the checkpoints of the preceding vh_evp_forward_checkpointed call (same d,
 * s, f, n_substeps, ckpt_every, workspace).  d sqrt(x)/dx is taken as 0 at
 * x = 0 and the denominator guard contributes no derivative. */
int vh_evp_adjoint(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                   int32_t n_substeps, int32_t ckpt_every, const void* adj_uIce,
                   const void* adj_vIce, const void* adj_sigma1, const void* adj_sigma2,
                   const void* adj_sigma12, const vh_evp_adjoint_out* out, void* workspace,
                   size_t workspace_bytes, void* stream);

/* Reverse mode of SeaIceStrength (veris/dynamics_routines.py:5-18, with
 * pStarField per cell or NULL for the pStar setting), SeaIceMass
 * (veris/area_mass.py:17-28) and WindForcingXY (veris/dynsolver.py:7-84):
 * maps the strength, mass and wind-forcing adjoints of `evp_adj` to P*,
 * hIceMean, hSnowMean and the air drag.  With useRealFreshWaterFlux the
 * surface tilt holds the SeaIceLoad field, a state input here as in the
 * reference's kernel. */
int vh_pre_evp_adjoint(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                       const void* pStarField, const vh_evp_adjoint_out* evp_adj,
                       const vh_pre_evp_adjoint_out* out, void* stream);

/* Batched adjoint of n_members equal members: every field of `f`, every
 * output and every adjoint holds n_members padded boxes of the geometry `d`
 * back to back (member m at element m * (nx+4) * (ny+4)), and one launch
 * sequence runs all members (launch grids carry the member index).  Each
 * member's results equal the single-domain entry point's on that member bit
 * for bit; the settings (drag coefficients among them) are shared.  The
 * reference has no batching: these are vh_evp_adjoint_workspace_bytes,
 * vh_evp_forward_checkpointed, vh_evp_adjoint and vh_pre_evp_adjoint (the
 * reverse of veris/evp_solver.py:22-319 and of the pre-EVP forcing) for
 * ensembles of equal regional domains. */
int vh_evp_adjoint_batched_workspace_bytes(const vh_domain* d, int32_t n_members,
                                           int32_t n_substeps, int32_t ckpt_every, size_t* bytes);
int vh_evp_forward_checkpointed_batched(const vh_domain* d, int32_t n_members,
                                        const vh_settings* s, const vh_fields* f,
                                        int32_t n_substeps, int32_t ckpt_every, void* uIce,
                                        void* vIce, void* sigma1, void* sigma2, void* sigma12,
                                        void* workspace, size_t workspace_bytes, void* stream);
int vh_evp_adjoint_batched(const vh_domain* d, int32_t n_members, const vh_settings* s,
                           const vh_fields* f, int32_t n_substeps, int32_t ckpt_every,
                           const void* adj_uIce, const void* adj_vIce, const void* adj_sigma1,
                           const void* adj_sigma2, const void* adj_sigma12,
                           const vh_evp_adjoint_out* out, void* workspace,
                           size_t workspace_bytes, void* stream);
int vh_pre_evp_adjoint_batched(const vh_domain* d, int32_t n_members, const vh_settings* s,
                               const vh_fields* f, const void* pStarField,
                               const vh_evp_adjoint_out* evp_adj,
                               const vh_pre_evp_adjoint_out* out, void* stream);

/* Outputs of vh_pre_evp (caller-owned device fields). */
typedef struct vh_pre_evp_out {
  void *SeaIceMassC, *SeaIceMassU, *SeaIceMassV, *AreaW, *AreaS, *WindForcingX, *WindForcingY,
      *SeaIceStrength;
} vh_pre_evp_out;

/* SeaIceMass (veris/area_mass.py:17-28), AreaWS (:5-14), WindForcingXY
 * (veris/dynsolver.py:7-84) and SeaIceStrength (veris/dynamics_routines.py:
 * 5-18, pStarField per cell or NULL for the pStar setting) in one sweep, with
 * the same results as the four separate entry points. */
int vh_pre_evp(const vh_domain* d, const vh_settings* s, const vh_fields* f,
               const void* pStarField, const vh_pre_evp_out* out, void* stream);

/* vh_pre_evp of every member of a batch (vh_batch, packed fields as for
 * vh_evp_solver_batched; pStarField packed or NULL) in one launch per 64
 * members: each member's outputs equal vh_pre_evp's on that member bit for
 * bit.  No counterpart in the reference, which has no batching: it runs
 * SeaIceMass, AreaWS, WindForcingXY and SeaIceStrength per domain
 * (veris/area_mass.py:5-28, veris/dynsolver.py:7-84,
 * veris/dynamics_routines.py:5-18). */
int vh_pre_evp_batched(const vh_batch* b, const vh_settings* s, const vh_fields* f,
                       const void* pStarField, const vh_pre_evp_out* out, void* stream);

/* FP64 budgets of a batch in one launch per 64 members: for each of the
 * n_fields (1..8) packed fields (host array of device pointers) and each
 * member m, the FP64 sum over m's interior cells of field times `weight`
 * (packed, or NULL for 1), written to result[q * n_members + m] (device
 * memory, n_fields * n_members doubles).  The batched form of global_sum
 * (veris/global_sum.py:1-4, a stub in the reference) for the mass, snow,
 * area and heat budgets (weights rA). */
int vh_global_sums_batched(const vh_batch* b, int32_t n_fields, const void* const* fields,
                           const void* weight, double* result, void* stream);

/* Outputs of vh_ocean_to_ice. */
typedef struct vh_ocean_to_ice_out {
  void *uOcean, *vOcean, *theta, *ocSalt, *Qnet, *Qsw, *wSpeed;
} vh_ocean_to_ice_out;

/* The host ocean's surface fields handed to the sea ice before a step, as the
 * reference setup does inline (veris/setup/seaice_global_4deg/
 * seaice_global_4deg.py:847,868-876): ocean surface velocities and salinity
 * copied, theta = surface temperature + celsius2K, Qnet = -forc_temp_surface
 * * cpWater * rhoSea, Qsw = -SWdown (f->SWdown), wSpeed = |(uWind, vWind)|.
 * The surface fields are contiguous padded 2-D fields (the caller takes the
 * top level and current time slice). */
int vh_ocean_to_ice(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                    const void* u_surface, const void* v_surface, const void* temp_surface,
                    const void* salt_surface, const void* forc_temp_surface,
                    const vh_ocean_to_ice_out* out, void* stream);

/* Outputs of vh_ice_to_ocean. */
typedef struct vh_ice_to_ocean_out {
  void *surface_taux, *surface_tauy, *forc_temp_surface, *forc_salt_surface;
} vh_ice_to_ocean_out;

/* The fluxes handed back to the ocean after a step
 * (veris/setup/seaice_global_4deg/seaice_global_4deg.py:933-940): wind stress
 * blended with f->OceanStressU/V by f->AreaW/AreaS, forc_temp_surface =
 * -f->Qnet / (cpWater * rhoSea), forc_salt_surface = ice salt flux + the
 * restoring term. */
int vh_ice_to_ocean(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                    const void* surface_taux, const void* surface_tauy,
                    const void* forc_salt_surface_ice, const void* forc_salt_surface_res,
                    const vh_ice_to_ocean_out* out, void* stream);

/* Outputs of vh_bulkf_formula_lanl, in the reference's return order. */
typedef struct vh_bulk_flux_out {
  void *flwupa, *flha, *fsha, *df0dT, *ust, *vst, *evp, *ssq, *dEvdT;
} vh_bulk_flux_out;

/* bulkf_formula_lanl (veris/heat_flux_MITgcm.py:5-123): open-ocean bulk
 * fluxes (upward long wave, latent, sensible, their surface-temperature
 * derivative, wind stress at c-points, evaporation, surface humidity and its
 * derivative) from the wind uw/vw, air temperature ta [K], humidity qa and
 * surface temperature tsf [K], with transfer coefficients from five
 * Monin-Obukhov iterations.  `grav` is the host model's gravity (Veros'
 * settings.grav).  Pointwise over the padded arrays; ocn_mask is not read
 * (the reference uses it only for an array shape). */
int vh_bulkf_formula_lanl(const vh_domain* d, const vh_settings* s, double grav, const void* uw,
                          const void* vw, const void* ta, const void* qa, const void* tsf,
                          const void* ocn_mask, const vh_bulk_flux_out* out, void* stream);

/* freedrift_solver (veris/freedrift_solver.py:5-62): steady free-drift
 * velocities from WindForcingX/Y, hIceMean, fCori, uOcean/vOcean and the
 * u/v masks, on every padded cell (the reference does not fill overlaps
 * here). Device atan2/sin/cos: equal to the reference to rounding. */
int vh_freedrift_solver(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                        void* uIce, void* vIce, void* stream);

/* OceanStressUV (veris/ocean_stress.py:8-40), overlap-filled. */
int vh_ocean_stress_uv(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                       void* OceanStressU, void* OceanStressV, void* stream);

/* Workspace bytes vh_advection needs for this domain. */
int vh_advection_workspace_bytes(const vh_domain* d, size_t* bytes);

/* Advection (veris/advection.py:15-91, fluxes :94-175) of hIceMean, hSnowMean
 * and Area with the superbee-limited, dimensionally split scheme. */
int vh_advection(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                 void* hIceMean, void* hSnowMean, void* Area, void* workspace,
                 size_t workspace_bytes, void* stream);

/* Workspace bytes vh_advection_adjoint needs for this domain. */
int vh_advection_adjoint_workspace_bytes(const vh_domain* d, size_t* bytes);

/* Reverse mode of vh_advection (the derivative jax.grad takes through
 * veris/advection.py:15-175 in the reference): given the adjoints of the
 * advected hIceMean, hSnowMean and Area (any may be NULL = zero), writes the
 * adjoints of the three input fields and of the ice velocities uIce, vIce
 * (f->uIce, f->vIce are the advecting velocities, f->hIceMean, hSnowMean,
 * Area the advected fields).  Each flux's derivative follows the branch its
 * forward evaluation takes (upwind side, ratio or clipped slope ratio,
 * superbee segment); d|x|/dx = sgn(x), 0 at 0. */
int vh_advection_adjoint(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                         const void* adj_hIceMean, const void* adj_hSnowMean,
                         const void* adj_Area, void* adj_hIceMean_in, void* adj_hSnowMean_in,
                         void* adj_Area_in, void* adj_uIce, void* adj_vIce, void* workspace,
                         size_t workspace_bytes, void* stream);

/* clean_up_advection (veris/clean_up.py:5-39). */
int vh_clean_up_advection(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                          void* hIceMean, void* hSnowMean, void* Area, void* TSurf,
                          void* os_hIceMean, void* os_hSnowMean, void* stream);

/* ridging (veris/clean_up.py:42-47): Area = min(f->Area, 1). */
int vh_ridging(const vh_domain* d, const vh_settings* s, const vh_fields* f, void* Area,
               void* stream);

/* Growth with solve4temp (veris/growth.py:8-397, veris/solve4temp.py:5-174). */
int vh_growth(const vh_domain* d, const vh_settings* s, const vh_fields* f,
              const vh_growth_out* out, void* stream);

/* fill_overlap (veris/fill_overlap.py:6-18) of one field, in place. */
int vh_fill_overlap(const vh_domain* d, void* field, void* stream);

/* fill_overlap of n distinct fields of one grid in place, in one launch
   (n <= 8): the halo fills that close a model step
   (veris/setup/seaice_global_4deg/seaice_global_4deg.py:928-932 fills them one call each). */
int vh_fill_overlaps(const vh_domain* d, int32_t n, void* const* fields, void* stream);

/* Workspace bytes vh_global_sum_f64 needs for this domain. */
int vh_global_sum_workspace_bytes(const vh_domain* d, size_t* bytes);

/* global_sum (veris/global_sum.py:1-4, a stub in the reference): FP64 sum over
 * the interior cells of field (times weight if not NULL), accumulated in FP64
 * whatever the domain dtype; the result is written to *result (device
 * memory, one double). */
int vh_global_sum_f64(const vh_domain* d, const void* field, const void* weight,
                      double* result, void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VERIS_HIP_H */
