// Fused EVP substep kernel template (gfx950): K substeps of
// veris/evp_solver.py:22-254 per launch (fixed or adaptive relaxation, no-slip
// with or without the second-order boundary term, or free-slip with lateral
// drag), shared by the
// single-domain drop-in (evp.hip, FP64/FP32 in the reference's operation
// order) and the batched ensemble path (evp_batched_fast.hip).
//
// Work decomposition.  A workgroup of RI x RJ threads owns a TI x TJ tile
// (TI = RI - 4K, TJ = RJ - 4K, j contiguous) plus a ring of 2K cells, ONE
// THREAD PER CELL.  Every thread issues all global loads its cell needs at
// once (at 1 degree the grid is under two waves per SIMD, so memory latency,
// not bandwidth, sets the time) and keeps its velocity and stresses in
// registers across the K substeps.  On grids whose masks are 0/1 and whose
// metrics depend on j alone, the masks come as bits of one byte per cell and
// the metrics as rows along j in LDS (evp_pack_kernel).  Each substep runs in LDS-exchange stages separated
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

#include <atomic>
#include <type_traits>

#include "vh_common.h"

namespace vh {

// S is the storage type of the state (u, v, sigma1, sigma2, sigma12 and the
// previous-step velocities); T that of every other field (S = T except in the
// batched mode with FP32 step constants around an FP64 state).
template <typename T, typename S = T>
struct EvpConst {
  const S *uNm1, *vNm1;
  const T *P, *massC, *massU, *massV, *AreaW, *AreaS, *WFX, *WFY, *uOc, *vOc, *fC;
  const T *M, *MU, *MV, *mInC, *mInU, *mInV;
  const T *rdxU, *rdyV, *rdyU, *rdxV, *dxU, *dyU, *dxV, *dyV, *rrAu, *rrAv, *rAz, *rrA;
  const T *k1C, *k2C, *k1Z, *k2Z;  // metric terms (read when kterms)
  const T *Rlow, *Area, *hIce;     // basal drag inputs (read when basal)
  const T *Fu, *Fv;                // coastline form factors (read when sidedrag and coast)
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
  double evpcfac, alphaMin;  // adaptive EVP: deltatDyn * 4 * (pi * aEvpCoeff)^2, aEVPalphaMin
  double sdCoeff, sdU0;      // sideDragCoeff, sideDragU0
  double alpha, beta;        // evpAlpha, evpBeta (residual weights of the fixed scheme)
  int kterms, turn, basal, adaptive;
  int sidedrag;  // noSlip = False: free-slip e12 and lateral drag in the denominators
  int coast;     // use_coastline: lateral drag form factors from Fu/Fv
  int sobc;      // noSlip and secondOrderBC: second-order boundary shear (K = 1 shapes only)
  int tens0;     // tensileStrFac == 0: the factors 1 + t and 1 - t are exactly 1
};

// Per-member geometry of a batched launch: element offset of the member's
// padded box inside every packed field, its extents, and the index of its
// first tile in the main-shape and remainder-shape launches (tiles of a
// member are consecutive; tb[.] of member m+1 = tb[.] of m + its tile count).
struct MemberDesc {
  long long off;
  Geom g;
  int tb[2];
  int roff;  // first element of the member's metric rows (packed constants)
  int pad;
};

// Placement of one tile of a batched launch (its member's offset, extents and
// row offset, and the tile's origin in the member), written once per call so
// that a workgroup finds its tile with one burst of scalar loads instead of a
// search over the member table.
struct TileDesc {
  long long off;
  int NX, NY, nx, ny, fill_x, fill_y;
  int i0, j0, roff, pad;
};

// Up to kMembersPerWrite member descriptors, passed by value to the kernel
// that stores them (kernel arguments are copied when the launch is enqueued).
constexpr int kMembersPerWrite = 48;
struct MemberChunk {
  int first, count;
  MemberDesc m[kMembersPerWrite];
};

static __global__ void evp_write_members_kernel(MemberDesc* dst, MemberChunk chunk) {
  const int q = threadIdx.x;
  if (q < chunk.count) dst[chunk.first + q] = chunk.m[q];
}

// Everything a launch reads, in device memory (written once per solver call,
// so a captured graph of launches stays valid when the caller's buffers move).
template <typename T, typename S = T>
struct EvpArgBlock {
  EvpConst<T, S> c;
  const S* in[5];  // u, v, sigma1, sigma2, sigma12 at the start of the subcycle
  S* out[5];       // final state (written by the last launch)
  S* alt[5];       // ping-pong partner
  EvpScalars sc;
  Geom g;          // single-domain geometry (unused when batched)
  int nlaunch;     // launches in the subcycle
  int pad;
  const MemberDesc* members;  // batched only
  int n_members;              // batched only
  int pad2;
  double* res;  // residual partial sums [launch][tile][2] (residual launches only)
  // compact step constants (evp_pack_kernel), read when packed = 1: the six
  // masks as bits of one byte per cell, and the twelve metrics as rows along j
  const unsigned char* pk;
  T* rows;
  int rstride;  // elements from one metric's rows to the next
  int packed;
  const TileDesc* tiles[2];  // batched: per-tile placement of the main and remainder shapes (or null)
  // single-domain launches: elements from one member to the next of a batch
  // of equal members packed back to back (launch grid y = member; 0 or any
  // value for one domain, whose grid has y = 1)
  long long mstride;
  // two-substep launches: where the state after the first substep goes (every
  // cell a launch owns; null: not kept)
  S* mid[5];
};


template <typename T>
inline const T* evp_ptr(const void* p) {
  return static_cast<const T*>(p);
}

// Scalars of one subcycle from the reference settings (host side).
inline void fill_evp_scalars(const vh_settings* s, int kterms, int basal, EvpScalars* out) {
  EvpScalars& sc = *out;
  sc.rpd = 1.0 / (s->PlasDefCoeff * s->PlasDefCoeff);
  sc.deltaMin = s->deltaMin;
  sc.tens1p = 1 + s->tensileStrFac;
  sc.tens1m = 1 - s->tensileStrFac;
  sc.tens0 = s->tensileStrFac == 0.0;
  sc.pr = s->pressReplFac;
  sc.pr1m = 1 - s->pressReplFac;
  sc.alphaM1 = s->evpAlpha - 1;
  sc.den1 = 1 / s->evpAlpha;
  sc.den2 = sc.den1;
  sc.cosWat = s->cosWat;
  sc.sinWat = s->sinWat;
  sc.cwN = s->waterIceDrag;
  sc.cwS = s->waterIceDrag_south;
  sc.rhoSea = s->rhoSea;
  sc.cDragMin = s->cDragMin;
  sc.cDragMin2 = s->cDragMin * s->cDragMin;
  sc.rdt = s->recip_deltatDyn;
  sc.betaU = s->evpBeta * s->recip_deltatDyn;
  sc.betaP1 = sc.betaU + s->recip_deltatDyn;
  sc.K2 = s->basalDragK2;
  sc.U0sq = s->basalDragU0 * s->basalDragU0;
  sc.K1 = s->basalDragK1;
  sc.cBasalStar = s->cBasalStar;
  sc.kterms = kterms;
  sc.turn = s->sinWat != 0.0 ? 1 : 0;
  sc.basal = basal;
  sc.adaptive = s->useAdaptiveEVP != 0.0 ? 1 : 0;
  // veris/evp_solver.py:266-271 (same operation order)
  sc.evpcfac = sc.adaptive ? s->deltatDyn * 4 * ((M_PI * s->aEvpCoeff) * (M_PI * s->aEvpCoeff)) : 0.0;
  sc.alphaMin = s->aEVPalphaMin;
  sc.sidedrag = s->noSlip == 0.0 ? 1 : 0;
  sc.coast = s->use_coastline != 0.0 ? 1 : 0;
  sc.sobc = (s->noSlip != 0.0 && s->secondOrderBC != 0.0) ? 1 : 0;
  sc.alpha = s->evpAlpha;
  sc.beta = s->evpBeta;
  sc.sdCoeff = s->sideDragCoeff;
  sc.sdU0 = s->sideDragU0;
}

// Step-constant field pointers of one subcycle (host side).
template <typename T, typename S>
inline void fill_evp_const(const vh_fields* fl, EvpConst<T, S>* out) {
  EvpConst<T, S>& f = *out;
  f.uNm1 = evp_ptr<S>(fl->uIce);
  f.vNm1 = evp_ptr<S>(fl->vIce);
  f.P = evp_ptr<T>(fl->SeaIceStrength);
  f.massC = evp_ptr<T>(fl->SeaIceMassC);
  f.massU = evp_ptr<T>(fl->SeaIceMassU);
  f.massV = evp_ptr<T>(fl->SeaIceMassV);
  f.AreaW = evp_ptr<T>(fl->AreaW);
  f.AreaS = evp_ptr<T>(fl->AreaS);
  f.WFX = evp_ptr<T>(fl->WindForcingX);
  f.WFY = evp_ptr<T>(fl->WindForcingY);
  f.uOc = evp_ptr<T>(fl->uOcean);
  f.vOc = evp_ptr<T>(fl->vOcean);
  f.fC = evp_ptr<T>(fl->fCori);
  f.M = evp_ptr<T>(fl->iceMask);
  f.MU = evp_ptr<T>(fl->iceMaskU);
  f.MV = evp_ptr<T>(fl->iceMaskV);
  f.mInC = evp_ptr<T>(fl->maskInC);
  f.mInU = evp_ptr<T>(fl->maskInU);
  f.mInV = evp_ptr<T>(fl->maskInV);
  f.rdxU = evp_ptr<T>(fl->recip_dxU);
  f.rdyV = evp_ptr<T>(fl->recip_dyV);
  f.rdyU = evp_ptr<T>(fl->recip_dyU);
  f.rdxV = evp_ptr<T>(fl->recip_dxV);
  f.dxU = evp_ptr<T>(fl->dxU);
  f.dyU = evp_ptr<T>(fl->dyU);
  f.dxV = evp_ptr<T>(fl->dxV);
  f.dyV = evp_ptr<T>(fl->dyV);
  f.rrAu = evp_ptr<T>(fl->recip_rAu);
  f.rrAv = evp_ptr<T>(fl->recip_rAv);
  f.rAz = evp_ptr<T>(fl->rAz);
  f.rrA = evp_ptr<T>(fl->recip_rA);
  f.k1C = evp_ptr<T>(fl->k1AtC);
  f.k2C = evp_ptr<T>(fl->k2AtC);
  f.k1Z = evp_ptr<T>(fl->k1AtZ);
  f.k2Z = evp_ptr<T>(fl->k2AtZ);
  f.Rlow = evp_ptr<T>(fl->R_low);
  f.Area = evp_ptr<T>(fl->Area);
  f.hIce = evp_ptr<T>(fl->hIceMean);
  f.Fu = evp_ptr<T>(fl->Fu);
  f.Fv = evp_ptr<T>(fl->Fv);
}

// Fields a subcycle needs (all must be non-null).
inline int evp_required_fields(const vh_fields* fl, const void** req) {
  const void* r[] = {fl->uIce, fl->vIce, fl->sigma1, fl->sigma2, fl->sigma12,
                     fl->SeaIceStrength, fl->SeaIceMassC, fl->SeaIceMassU, fl->SeaIceMassV,
                     fl->AreaW, fl->AreaS, fl->WindForcingX, fl->WindForcingY, fl->uOcean,
                     fl->vOcean, fl->fCori, fl->iceMask, fl->iceMaskU, fl->iceMaskV,
                     fl->maskInC, fl->maskInU, fl->maskInV, fl->recip_dxU, fl->recip_dyV,
                     fl->recip_dyU, fl->recip_dxV, fl->dxU, fl->dyU, fl->dxV, fl->dyV,
                     fl->recip_rAu, fl->recip_rAv, fl->rAz, fl->recip_rA};
  const int n = (int)(sizeof(r) / sizeof(r[0]));
  for (int q = 0; q < n; ++q) req[q] = r[q];
  return n;
}

constexpr int kEvpMaxRequired = 34;

template <typename T, typename S = T>
__global__ void evp_write_args_kernel(EvpArgBlock<T, S>* dst, EvpArgBlock<T, S> src) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = src;
}

// Packed step constants.  On the grids the benchmark and the reference's
// setups use, the six masks hold only +0 and 1 and the twelve grid metrics
// depend on j alone, so a cell's 18 constant words (144 B in FP64) carry one
// byte of information plus a row of metrics shared along i.  Once per
// subcycle this kernel writes, per cell, the masks as bits of one byte and,
// per metric, its values along the reference row i = NX/2; a cell whose masks
// are not exactly +0 or 1, or whose metrics differ from the row in any bit,
// clears `packed` (set to 1 by the host), and the launches then read the full
// fields.  Either way the substep sees the same values bit for bit.
constexpr int kEvpRowFields = 12;  // rdyU rdxV rAz rdxU rdyV rrA dyV dxU dxV dyU rrAu rrAv
constexpr int kEvpMaskFields = 6;  // iceMask iceMaskU iceMaskV maskInC maskInU maskInV


// Square root and quotient of the FP64 island (Delta, zeta, the replacement
// pressure) of the kernels over FP32 step constants, whose results are not
// held to the FP64 drop-in's bits: the hardware FP64 reciprocal (square-root) estimate refined by one
// Goldschmidt / Newton-Raphson step and one residual correction, within an
// ulp or two of the correctly rounded FP64 value in the ranges they see
// (x >= 0; d >= deltaMin > 0), in half the dependent instructions.  The
// FP64-arithmetic kernels keep the correctly rounded sqrt and /.
__device__ inline double island_sqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);  // 1/sqrt(x)
  double s = x * y, h = 0.5 * y;
  const double r = fma(-s, h, 0.5);
  s = fma(s, r, s);
  h = fma(h, r, h);
  s = fma(fma(-s, s, x), h, s);
  // +0 (Delta of a motionless or masked cell) and values whose root is far
  // below deltaMin, where the estimate's range ends: 0; a NaN stays NaN
  return (x >= 0.0 && x <= 1e-200) ? 0.0 : s;
}
__device__ inline double island_div(double n, double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  const double q = n * r;
  return fma(fma(-d, q, n), r, q);
}

// Division and square root of the momentum update and the water drag.  FP64
// arithmetic on FP64 step constants keeps the correctly rounded forms (the
// drop-in's bit-for-bit parity); FP32 arithmetic takes the hardware
// reciprocal and square root (about 1 ulp each, a third of the instructions
// of the rounded sequences); FP64 arithmetic over FP32 step constants (the
// ensemble's mixed modes) takes the refined FP64 forms above.
template <typename A, typename T>
__device__ inline A evp_div(A a, A b) {
  if constexpr (sizeof(A) == 4) return a * __builtin_amdgcn_rcpf(b);
  else if constexpr (sizeof(T) == 4) return island_div(a, b);
  else return a / b;
}
template <typename A, typename T>
__device__ inline A evp_sqrt(A x) {
  if constexpr (sizeof(A) == 4) return __builtin_amdgcn_sqrtf(x);
  else if constexpr (sizeof(T) == 4) return island_sqrt(x);
  else return sqrt(x);
}

__device__ inline unsigned long long bits_of(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ inline unsigned long long bits_of(float x) { return __float_as_uint(x); }
template <typename T>
constexpr unsigned long long kOneBits = sizeof(T) == 8 ? 0x3FF0000000000000ull : 0x3F800000ull;

// Bytes of the packed constants of `elems` cells with `rows` metric rows.
template <typename T>
inline size_t evp_pack_bytes(size_t elems, size_t rows) {
  return align256(elems) + align256((size_t)kEvpRowFields * rows * sizeof(T));
}

template <typename T, typename S>
__global__ __launch_bounds__(256) void evp_pack_kernel(EvpArgBlock<T, S>* ab) {
  Geom g;
  long long off = 0;
  int roff = 0;
  if (ab->members) {
    const MemberDesc& md = ab->members[blockIdx.y];
    g = md.g;
    off = md.off;
    roff = md.roff;
  } else {  // one domain, or equal members packed back to back
    g = ab->g;
    off = (long long)blockIdx.y * ab->mstride;
    roff = (int)blockIdx.y * g.NY;
  }
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= g.NX * g.NY) return;
  const int i = e / g.NY, j = e - i * g.NY;
  const EvpConst<T, S>& f = ab->c;
  const T* const masks[kEvpMaskFields] = {f.M, f.MU, f.MV, f.mInC, f.mInU, f.mInV};
  const T* const mets[kEvpRowFields] = {f.rdyU, f.rdxV, f.rAz, f.rdxU, f.rdyV, f.rrA,
                                        f.dyV,  f.dxU,  f.dxV, f.dyU,  f.rrAu, f.rrAv};
  unsigned bits = 0;
  bool ok = true;
#pragma unroll
  for (int q = 0; q < kEvpMaskFields; ++q) {
    const unsigned long long b = bits_of(masks[q][off + e]);
    if (b == kOneBits<T>) bits |= 1u << q;
    else if (b != 0) ok = false;
  }
  const int iref = g.NX / 2;
  const long long eref = off + (long long)iref * g.NY + j;
#pragma unroll
  for (int q = 0; q < kEvpRowFields; ++q) {
    const T x = mets[q][off + e];
    if (bits_of(x) != bits_of(mets[q][eref])) ok = false;
    if (i == iref) ab->rows[(long long)q * ab->rstride + roff + j] = x;
  }
  ((unsigned char*)ab->pk)[off + e] = (unsigned char)bits;
  if (!ok) ab->packed = 0;
}

// Tile table of a batched launch shape (`which`: 0 main, 1 remainder) with
// tile extents TI x TJ: one thread per tile finds its member in the member
// table (the last member whose first tile is at or before it).
template <typename T, typename S = T>
__global__ __launch_bounds__(256) void evp_tile_table_kernel(EvpArgBlock<T, S>* ab, int which, int TI,
                                                             int TJ, int n_tiles) {
  const int tile = blockIdx.x * 256 + threadIdx.x;
  if (tile >= n_tiles) return;
  const MemberDesc* ms = ab->members;
  int lo = 0, hi = ab->n_members - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ms[mid].tb[which] <= tile) lo = mid;
    else hi = mid - 1;
  }
  const MemberDesc& md = ms[lo];
  const int local = tile - md.tb[which];
  const int tj_m = (md.g.NY + TJ - 1) / TJ;
  const int ti = local / tj_m;
  TileDesc d;
  d.off = md.off;
  d.NX = md.g.NX;
  d.NY = md.g.NY;
  d.nx = md.g.nx;
  d.ny = md.g.ny;
  d.fill_x = md.g.fill_x;
  d.fill_y = md.g.fill_y;
  d.i0 = ti * TI;
  d.j0 = (local - ti * tj_m) * TJ;
  d.roff = md.roff;
  d.pad = 0;
  ((TileDesc*)ab->tiles[which])[tile] = d;
}

// Points the host-side block at packed storage `buf` (evp_pack_bytes<T>(elems,
// rows) bytes) and marks it packed; evp_pack_launch then fills it on device.
template <typename T, typename S>
inline void evp_pack_attach(EvpArgBlock<T, S>* ab, void* buf, size_t elems, int rows) {
  ab->pk = (const unsigned char*)buf;
  ab->rows = (T*)((char*)buf + align256(elems));
  ab->rstride = rows;
  ab->packed = 1;
}

template <typename T, typename S>
inline void evp_pack_launch(EvpArgBlock<T, S>* ab_dev, long long max_cells, int n_members,
                            hipStream_t st) {
  hipLaunchKernelGGL((evp_pack_kernel<T, S>), dim3((unsigned)((max_cells + 255) / 256), n_members),
                     dim3(256), 0, st, ab_dev);
}

// Global-address-space views of field pointers: loads through them compile to
// global loads with a uniform (SGPR) base and a 32-bit per-lane offset.
template <typename T>
using gconst = const T __attribute__((address_space(1)))*;
template <typename T>
using gmut = T __attribute__((address_space(1)))*;
template <typename T>
__device__ inline gconst<T> G(const T* p) {
  return (gconst<T>)p;
}
template <typename T>
__device__ inline gmut<T> GW(T* p) {
  return (gmut<T>)p;
}

// Element `idx` of a field whose base is wavefront-uniform: the byte offset is
// formed in 32 bits so the load uses the scalar base + 32-bit lane offset form.
template <typename T>
__device__ inline T ldg(const T* base, unsigned idx) {
  return *(gconst<T>)((const char*)base + (size_t)(idx * (unsigned)sizeof(T)));
}
template <typename T>
__device__ inline void stg(T* base, unsigned idx, T v) {
  *(gmut<T>)((char*)base + (size_t)(idx * (unsigned)sizeof(T))) = v;
}

__device__ inline int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// x folded into [0, m) by whole periods.  Region coordinates lie within a
// ring of a few cells around the domain, so one period added or taken away
// suffices for every domain larger than a workgroup's region; only smaller
// ones (lanes left out of range) take the integer remainder.
__device__ inline int fold(int x, int m) {
  x = x < 0 ? x + m : (x >= m ? x - m : x);
  if ((unsigned)x >= (unsigned)m) x = ((x % m) + m) % m;
  return x;
}

// padded index of region coordinate p along an axis of padded length N,
// interior length n, overlap 2: roll wrap, or periodic interior if filled
template <int K>
__device__ inline int map_axis(int p, int N, int n, int filled) {
  if (K > 1 && filled) return 2 + fold(p - 2, n);
  return fold(p, N);
}

// map_axis<K>(p + 1, ...) from q = map_axis<K>(p, ...): the next index of the
// same axis, without the remainder (hoisted out of the substep loop, the
// remainder's rare path would otherwise run for every lane of every tile)
template <int K>
__device__ inline int next_axis(int q, int N, int n, int filled) {
  if (K > 1 && filled) return q + 1 == 2 + n ? 2 : q + 1;
  return q + 1 == N ? 0 : q + 1;
}

template <typename T, typename TD, int RI, int RJ, int K, bool BATCHED, typename A = T,
          bool SO = false, bool RES = false, typename S = T, bool PL = false>
__global__ __launch_bounds__(RI* RJ) void evp_substep_kernel(const EvpArgBlock<T, S>* __restrict__ ab,
                                                             int which, int launch, int n_tiles,
                                                             int tiles_j, int per_xcd) {
  constexpr int R = 2 * K;
  constexpr int TI = RI - 2 * R, TJ = RJ - 2 * R, NT = RI * RJ;
  static_assert((RJ & (RJ - 1)) == 0, "RJ must be a power of two");
  static_assert(TI >= 1 && TJ >= 1, "tile too small for K");
  __shared__ A sU[NT], sV[NT], sM[NT], sMU[NT], sMV[NT], sDU[NT], sDV[NT], sE12[NT], sQ[NT], sZ[NT];
  __shared__ A sS11[NT], sS22[NT], sCD[NT], sCB[NT], sFV[NT], sFU[NT], sAl[NT];

  // tiles of this workgroup: workgroups are dealt round-robin to the 8 XCDs,
  // so each XCD owns a contiguous run of per_xcd tiles (neighbours share L2
  // lines); the grid may hold fewer workgroups than tiles, each then loops
  // over its XCD's run with a stride of the XCD's workgroup count
  const int bid = blockIdx.x;
  const int xcd = bid & 7, k0 = bid >> 3;
  const bool packed = ab->packed != 0;  // uniform: step constants in packed form
  // placement of a tile: its member's geometry and element offset, the row
  // offset of its metric rows, its origin and its index
  struct Place {
    Geom g;
    long long off;
    int roff, i0, j0, tile;
  };
  auto place = [&](const int tile) -> Place {
    Geom g;
    long long off = 0;
    int roff = 0;
    int i0, j0;
    if (BATCHED && ab->tiles[which] != nullptr) {
      // the tile's placement from the per-call tile table
      const TileDesc* td = ab->tiles[which] + tile;
      g.NX = uniform(td->NX);
      g.NY = uniform(td->NY);
      g.nx = uniform(td->nx);
      g.ny = uniform(td->ny);
      g.ol = 2;
      g.fill_x = uniform(td->fill_x);
      g.fill_y = uniform(td->fill_y);
      const long long to = td->off;
      off = ((long long)(unsigned)uniform((int)(to >> 32)) << 32) | (unsigned)uniform((int)to);
      roff = uniform(td->roff);
      i0 = uniform(td->i0);
      j0 = uniform(td->j0);
    } else if (BATCHED) {
      // member owning this tile: last m with tb[which] <= tile (uniform search)
      const MemberDesc* ms = ab->members;
      int lo = 0, hi = uniform(ab->n_members) - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (uniform(ms[mid].tb[which]) <= tile) lo = mid;
        else hi = mid - 1;
      }
      lo = uniform(lo);
      const MemberDesc* md = ms + lo;
      g.NX = uniform(md->g.NX);
      g.NY = uniform(md->g.NY);
      g.nx = uniform(md->g.nx);
      g.ny = uniform(md->g.ny);
      g.ol = 2;
      g.fill_x = uniform(md->g.fill_x);
      g.fill_y = uniform(md->g.fill_y);
      const long long mo = md->off;
      off = ((long long)(unsigned)uniform((int)(mo >> 32)) << 32) | (unsigned)uniform((int)mo);
      roff = uniform(md->roff);
      const int local = tile - uniform(md->tb[which]);
      const int tj_m = (g.NY + TJ - 1) / TJ;
      const int ti = local / tj_m;
      i0 = ti * TI;
      j0 = (local - ti * tj_m) * TJ;
    } else {  // one domain, or member blockIdx.y of equal members packed back to back
      g = ab->g;
      off = (long long)blockIdx.y * ab->mstride;
      roff = (int)blockIdx.y * g.NY;
      const int ti = tile / tiles_j;
      i0 = ti * TI;
      j0 = (tile - ti * tiles_j) * TJ;
    }
    return Place{g, off, roff, i0, j0, tile};
  };
  constexpr int RPT = (kEvpRowFields * RJ + NT - 1) / NT;  // metric-row loads per thread
  // global loads of one cell of a tile with packed step constants, issued ahead
  // of the tile's computation by the pipelined loop below
  struct Pre {
    A rowv[RPT];
    A u, v, uoc, voc;
    unsigned pkb;
    A P, s1, s2, fc, mC, s12, mU, mV, AW, AS, wfx, wfy, uNm1, vNm1;
  };
  const int nl = ab->nlaunch;
  const S* const* src = (launch == 0) ? ab->in
                                      : (((nl - launch) % 2 == 0) ? (const S* const*)ab->out
                                                                  : (const S* const*)ab->alt);
  S* const* dst = ((nl - 1 - launch) % 2 == 0) ? ab->out : ab->alt;
  const int t = threadIdx.x;
  const int ar = t / RJ, br = t & (RJ - 1);  // region coordinates
  // padded position of this thread's cell in a tile, and its element offset
  // (member offset + folded index) into every packed field
  auto cell_of = [&](const Place& pl, int& i, int& j, int& gi, int& gj, unsigned& ce) {
    i = pl.i0 + ar - R;
    j = pl.j0 + br - R;
    gi = map_axis<K>(i, pl.g.NX, pl.g.nx, pl.g.fill_x);
    gj = map_axis<K>(j, pl.g.NY, pl.g.ny, pl.g.fill_y);
    ce = (unsigned)pl.off + (unsigned)(gi * pl.g.NY + gj);
  };
  // the loads of Pre for tile placement pl, in the order run_tile issues them
  auto preload = [&](const Place& pl) -> Pre {
    int i, j, gi, gj;
    unsigned ce;
    cell_of(pl, i, j, gi, gj, ce);
    const EvpConst<T, S>& f = ab->c;
    Pre q;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int e = t + r * NT;
      q.rowv[r] = (A)0;
      if (e < kEvpRowFields * RJ) {
        const int mq = e / RJ, bq = e & (RJ - 1);
        const int gjq = map_axis<K>(pl.j0 + bq - R, pl.g.NY, pl.g.ny, pl.g.fill_y);
        q.rowv[r] = (A)ldg(ab->rows + (long long)mq * ab->rstride + pl.roff, (unsigned)gjq);
      }
    }
    q.u = ldg(src[0], ce);
    q.v = ldg(src[1], ce);
    q.uoc = ldg(f.uOc, ce);
    q.voc = ldg(f.vOc, ce);
    q.pkb = ldg(ab->pk, ce);
    q.P = ldg(f.P, ce);
    q.s1 = ldg(src[2], ce);
    q.s2 = ldg(src[3], ce);
    q.fc = ldg(f.fC, ce);
    q.mC = ldg(f.massC, ce);
    q.s12 = ldg(src[4], ce);
    q.mU = ldg(f.massU, ce);
    q.mV = ldg(f.massV, ce);
    q.AW = ldg(f.AreaW, ce);
    q.AS = ldg(f.AreaS, ce);
    q.wfx = ldg(f.WFX, ce);
    q.wfy = ldg(f.WFY, ce);
    q.uNm1 = ldg(f.uNm1, ce);
    q.vNm1 = ldg(f.vNm1, ce);
    return q;
  };
  // the tile body is instantiated for packed step constants and for full
  // fields (pk: std::true_type / std::false_type), each with its own schedule
  // and registers, the packed flag of the argument block picks one; pf
  // (std::true_type) takes the cell's loads from `pre` instead of issuing them
  auto run_tile = [&](const Place& pl, auto pk, auto pf, const Pre& pre) {
    constexpr bool PF = decltype(pf)::value;
    const Geom& g = pl.g;
    const int roff = pl.roff, i0 = pl.i0, j0 = pl.j0;
    const long long off = pl.off;
    const EvpConst<T, S>& f = ab->c;
    const EvpScalars& sc = ab->sc;
    // settings branches; PL instantiates the default settings (none of them,
    // tensileStrFac = 0) with every branch compiled out
    const bool f_kterms = !PL && sc.kterms, f_turn = !PL && sc.turn, f_basal = !PL && sc.basal;
    const bool f_adaptive = !PL && sc.adaptive, f_sidedrag = !PL && sc.sidedrag;
    const bool f_coast = !PL && sc.coast, f_tens0 = PL || sc.tens0;

    // padded position; member-local element index gi * NY + gj (every region
    // cell is in bounds).  Element offsets into the packed fields fit 32 bits
    // (the host checks the batch size), so every load below uses the field's
    // uniform base pointer plus one per-lane 32-bit offset shared by all fields
    int i, j, gi, gj;
    unsigned ce;
    cell_of(pl, i, j, gi, gj, ce);
    const unsigned oo = (unsigned)off;

    // final output cell of this thread
    const bool own = (ar >= R) && (ar < R + TI) && (br >= R) && (br < R + TJ) && (i < g.NX) &&
                     (j < g.NY);
    const bool upd_any = (ar >= 2) && (ar < RI - 2) && (br >= 2) && (br < RJ - 2) &&
                         (K > 1 || is_computed(g, i, j));

    // ---- all global loads of this cell, issued together.  Packed constants
    //      (uniform flag): the masks come from one byte, the metrics from rows
    //      held in LDS (sRow, filled below); otherwise from the full fields.
    constexpr bool packed = decltype(pk)::value;
    // metric rows of this tile's columns: issued first, so that stage A, which
    // stores them to LDS before its barrier, waits for no later load
    A rowv[RPT];
    if constexpr (PF) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) rowv[r] = pre.rowv[r];
    } else if (packed) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int q = t + r * NT;
        rowv[r] = (A)0;
        if (q < kEvpRowFields * RJ) {
          const int mq = q / RJ, bq = q & (RJ - 1);
          const int gjq = map_axis<K>(j0 + bq - R, g.NY, g.ny, g.fill_y);
          rowv[r] = (A)ldg(ab->rows + (long long)mq * ab->rstride + roff, (unsigned)gjq);
        }
      }
    }
    A u = PF ? pre.u : (A)ldg(src[0], ce), v = PF ? pre.v : (A)ldg(src[1], ce);
    const A uoc = PF ? pre.uoc : (A)ldg(f.uOc, ce), voc = PF ? pre.voc : (A)ldg(f.vOc, ce);
    A Mc, MUc, MVc, mInC, mInUc, mInVc;
    if (packed) {
      const unsigned pkb = PF ? pre.pkb : (unsigned)ldg(ab->pk, ce);
      Mc = (A)(pkb & 1u);
      MUc = (A)((pkb >> 1) & 1u);
      MVc = (A)((pkb >> 2) & 1u);
      mInC = (A)((pkb >> 3) & 1u);
      mInUc = (A)((pkb >> 4) & 1u);
      mInVc = (A)((pkb >> 5) & 1u);
    } else {
      Mc = ldg(f.M , ce);
      MUc = ldg(f.MU , ce);
      MVc = ldg(f.MV , ce);
      mInUc = ldg(f.mInU , ce);
      mInVc = ldg(f.mInV , ce);
      mInC = ldg(f.mInC , ce);
    }
    // issue order = order of use: stage A waits only for the loads above, the
    // rest stream in while A and B run
    __builtin_amdgcn_sched_barrier(0);
    A k1Z = 0, k2Z = 0, k1C = 0, k2C = 0;  // metric terms
    if (f_kterms) {
      k1Z = ldg(f.k1Z , ce);
      k2Z = ldg(f.k2Z , ce);
    }
    __builtin_amdgcn_sched_barrier(0);
    const A P = PF ? pre.P : (A)ldg(f.P, ce);
    A s1 = PF ? pre.s1 : (A)ldg(src[2], ce), s2 = PF ? pre.s2 : (A)ldg(src[3], ce);
    const A fc = PF ? pre.fc : (A)ldg(f.fC, ce);
    const A mC = PF ? pre.mC : (A)ldg(f.massC, ce);
    A s12 = PF ? pre.s12 : (A)ldg(src[4], ce);
    const A mU = PF ? pre.mU : (A)ldg(f.massU, ce), mV = PF ? pre.mV : (A)ldg(f.massV, ce);
    const A AW = PF ? pre.AW : (A)ldg(f.AreaW, ce), AS = PF ? pre.AS : (A)ldg(f.AreaS, ce);
    const A wfx = PF ? pre.wfx : (A)ldg(f.WFX, ce), wfy = PF ? pre.wfy : (A)ldg(f.WFY, ce);
    const A uNm1 = PF ? pre.uNm1 : (A)ldg(f.uNm1, ce), vNm1 = PF ? pre.vNm1 : (A)ldg(f.vNm1, ce);
    if (f_kterms) {
      k1C = ldg(f.k1C , ce);
      k2C = ldg(f.k2C , ce);
    }
    A cbc = (A)0;  // basal drag coefficient inputs are step constants: load lazily
    const A sg = sgn(fc);
    // residual inputs (RES, one substep per launch): the iterate at the start of
    // the substep and the relaxation weights (fixed, or this substep's adaptive ones)
    const A r_s1o = s1, r_s2o = s2, r_s12o = s12, r_uo = u, r_vo = v;
    A r_aC = (A)sc.alpha, r_aZ = (A)sc.alpha, r_bU = (A)sc.beta, r_bV = (A)sc.beta;
    // lateral drag form factors at this u- and v-point (veris/dynamics_routines.py:119-130)
    A sdmU = (A)0, sdmV = (A)0;
    if (f_sidedrag && f_coast) {
      sdmU = ldg(f.Fu , ce);
      sdmV = ldg(f.Fv , ce);
    }
    // metric rows in LDS (stored in stage A of the first substep)
    __shared__ A sRow[kEvpRowFields * RJ];
    // metric q (the order of the pack kernel's rows) at this cell: from the
    // rows in LDS, or, on grids whose constants do not pack, loaded from its
    // field where it is used (no registers held for it across the stages)
    auto met = [&](int q, const T* full) -> A {
      return packed ? sRow[q * RJ + br] : (A)ldg(full, ce);
    };

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
      A spd = (A)0;  // ice speed at c (lateral drag only)

      // ---- A
      if (s > 0) __syncthreads();  // E of the previous substep still reads LDS
      if (packed && s == 0) {
#pragma unroll
        for (int r = 0; r < RPT; ++r)
          if (t + r * NT < kEvpRowFields * RJ) sRow[t + r * NT] = rowv[r];
      }
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
        const A rdyU = met(0, f.rdyU), rdxV = met(1, f.rdxV), rAz = met(2, f.rAz);
        const A ujm = sU[t - 1], vim = sV[t - RJ];
        const A dudy = (u - ujm) * rdyU;
        const A uavz = (u + ujm) * (A)0.5;
        const A dvdx = (v - vim) * rdxV;
        const A vavz = (v + vim) * (A)0.5;
        const A mz = (Mc * sM[t - RJ]) * (sM[t - 1] * sM[t - RJ - 1]);
        A inner = dudy + dvdx;
        if (f_kterms) inner = (inner - k1Z * vavz) - k2Z * uavz;
        A e12 = ((A)0.5 * inner) * mz;
        if (!f_sidedrag) {  // no-slip boundary term (veris/dynamics_routines.py:170-175)
          const A hu = MUc - sMU[t - 1];
          const A hv = MVc - sMV[t - RJ];
          e12 = e12 + ((((A)2.0 * uavz) * rdyU) * hu + (((A)2.0 * vavz) * rdxV) * hv);
        }
        if constexpr (SO && K == 1) {
          // second-order boundary shear (veris/dynamics_routines.py:177-215): it
          // reads u, v and the masks two cells away along the face normal, beyond
          // the LDS region at its edge; with one substep per launch the launch
          // input is the current state, so they are read from memory.  A separate
          // instantiation keeps these registers out of the other variants.
          const int cim2 = wrap(i - 2, g.NX) * g.NY + gj, cim1 = wrap(i - 1, g.NX) * g.NY + gj;
          const int cip1 = wrap(i + 1, g.NX) * g.NY + gj;
          const int cjm2 = gi * g.NY + wrap(j - 2, g.NY), cjm1 = gi * g.NY + wrap(j - 1, g.NY);
          const int cjp1 = gi * g.NY + wrap(j + 1, g.NY);
          const A uim2 = ldg(src[0] , oo + (unsigned)cim2), uip1 = ldg(src[0] , oo + (unsigned)cip1);
          const A MUim2 = ldg(f.MU , oo + (unsigned)cim2), MUim1 = ldg(f.MU , oo + (unsigned)cim1);
          const A MUip1 = ldg(f.MU , oo + (unsigned)cip1);
          const A vjm2 = ldg(src[1] , oo + (unsigned)cjm2), vjp1 = ldg(src[1] , oo + (unsigned)cjp1);
          const A MVjm2 = ldg(f.MV , oo + (unsigned)cjm2), MVjm1 = ldg(f.MV , oo + (unsigned)cjm1);
          const A MVjp1 = ldg(f.MV , oo + (unsigned)cjp1);
          A hu2 = (MUc - MUim1) / (A)3.0;
          A hv2 = (MVc - MVjm1) / (A)3.0;
          hu2 = hu2 * (MUim2 * MUim1 + MUip1 * MUc);
          hv2 = hv2 * (MVjm2 * MVjm1 + MVjp1 * MVc);
          const A tu = (rdyU * ((((A)6.0 * uavz) - uim2 * MUim1) - uip1 * MUc)) * hu2;
          const A tv = (rdxV * ((((A)6.0 * vavz) - vjm2 * MVjm1) - vjp1 * MVc)) * hv2;
          e12 = e12 + (A)0.5 * (tu + tv);
        }
        sE12[t] = e12;
        sQ[t] = rAz * (e12 * e12);
      }
      if (f_sidedrag && !f_coast && upd) {  // land neighbours along the face normal
        sdmU = (((A)2 - sMU[t - RJ]) - sMU[t + RJ]) * MUc;
        sdmV = (((A)2 - sMV[t - 1]) - sMV[t + 1]) * MVc;
      }
      __syncthreads();

      // ---- C: c-point rheology (veris/dynamics_routines.py:151-158, 220-268),
      //         stresses (veris/evp_solver.py:79-94), drags (:108-110), Coriolis (:142-143)
      if (cc) {
        const A rdxU = met(3, f.rdxU), rdyV = met(4, f.rdyV), rrA = met(5, f.rrA);
        const A uip = sU[t + RJ], vjp = sV[t + 1];
        const A dudx = (uip - u) * rdxU;
        const A uave = (uip + u) * (A)0.5;
        const A dvdy = (vjp - v) * rdyV;
        const A vave = (vjp + v) * (A)0.5;
        A e11 = dudx, e22 = dvdy;
        if (f_kterms) {
          e11 = dudx + vave * k2C;
          e22 = dvdy + uave * k1C;
        }
        e11 = e11 * mInC;
        e22 = e22 * mInC;
        const A qa = sQ[t] + sQ[t + RJ];
        const A qb = sQ[t + 1] + sQ[t + RJ + 1];
        const A e12c = ((A)0.25 * rrA) * (qa + qb);
        if (f_sidedrag) {  // ice speed at c (veris/dynamics_routines.py:109-111)
          const A su = u + uip, sv = v + vjp;
          spd = (A)0.5 * sqrt(su * su + sv * sv);
        }
        const A ep = e11 + e22, em = e11 - e22;
        const TD dsq = (TD)ep * (TD)ep + (TD)sc.rpd * ((TD)em * (TD)em + (TD)4.0 * (TD)e12c);
        TD dC;
        if constexpr (sizeof(T) == 4 && sizeof(TD) == 8) dC = island_sqrt(dsq);
        else dC = sqrt(dsq);
        const TD dreg = dC + (TD)sc.deltaMin;
        // with t = 0 the products and the quotient by 1 + t = 1 are exact
        // identities, so skipping them keeps every bit
        const TD Pt = f_tens0 ? (TD)P : (TD)P * (TD)sc.tens1p;
        TD zeta;
        if constexpr (sizeof(T) == 4 && sizeof(TD) == 8) zeta = island_div((TD)0.5 * Pt, dreg);
        else zeta = ((TD)0.5 * Pt) / dreg;
        sZ[t] = (A)zeta;
        // relaxation parameter at c (veris/evp_solver.py:65-73 when adaptive)
        A am1 = (A)sc.alphaM1, den1 = (A)sc.den1;
        if (f_adaptive) {
          A aC = sqrt((((A)zeta * (A)sc.evpcfac) / fmax(mC, (A)1e-4)) * rrA) * Mc;
          aC = fmax(aC, (A)sc.alphaMin);
          sAl[t] = aC;
          am1 = aC - (A)1;
          den1 = (A)1.0 / aC;
          if constexpr (RES) r_aC = aC;
        }
        if (cs) {
          TD prep = (((TD)2.0 * zeta) * dC) * (TD)sc.pr;
          if (!f_tens0) prep = prep / (TD)sc.tens1p;
          TD pressD = ((TD)P * (TD)sc.pr1m) + prep;
          if (!f_tens0) pressD = pressD * (TD)sc.tens1m;
          const A zt = (A)zeta, press = (A)pressD;
          const A divg = (((A)2 * zt) * ep) - press;
          const A tens = ((A)2 * zt) * em;
          s1 = (((s1 * am1) + divg) * den1) * Mc;
          s2 = (((s2 * am1) + tens * (A)0.25) * den1) * Mc;
          sS11[t] = ((A)0.5 * (s1 + s2)) * met(6, f.dyV);
          sS22[t] = ((A)0.5 * (s1 - s2)) * met(7, f.dxU);
          const A D = (A)(fc < (A)0 ? sc.cwS : sc.cwN) * (A)sc.rhoSea;
          const A su = sDU[t] + sDU[t + RJ], sv = sDV[t] + sDV[t + 1];
          const A w = (A)0.25 * (su * su + sv * sv);
          const A cd = ((D * D) * w > (A)sc.cDragMin2) ? D * evp_sqrt<A, T>(w) : (A)sc.cDragMin;
          sCD[t] = cd * Mc;
          if (f_basal) {
            const int cip = next_axis<K>(gi, g.NX, g.nx, g.fill_x) * g.NY + gj;
            const int cjp = gi * g.NY + next_axis<K>(gj, g.NY, g.ny, g.fill_y);
            const A um = u * mInUc, umip = uip * ldg(f.mInU , oo + (unsigned)cip);
            const A vm = v * mInVc, vmjp = vjp * ldg(f.mInV , oo + (unsigned)cjp);
            A tb = (A)0.25 * (((um * um + umip * umip) + vm * vm) + vmjp * vmjp);
            tb = (A)sc.K2 / sqrt(tb + (A)sc.U0sq);
            const A Ar = ldg(f.Area , ce);
            const A hc = (fabs(ldg(f.Rlow , ce)) * Ar) / (A)sc.K1;
            const A val = ((tb * log(exp((A)10.0 * (ldg(f.hIce , ce) - hc)) + (A)1.0)) * (A)0.1) *
                          exp((A)(-sc.cBasalStar) * ((A)1.0 - Ar));
            cbc = Ar > (A)0.01 ? val : (A)0;
          }
          if (!PL) sCB[t] = cbc;
          const A mf = (mC * fc) * (A)0.5;
          sFV[t] = mf * (v + vjp);
          sFU[t] = mf * (u + uip);
        }
      }
      if (f_turn) {  // ocean velocity for the turning term (sMU/sMV are free after B)
        sMU[t] = uoc;
        sMV[t] = voc;
      }
      __syncthreads();

      // ---- D: sigma12 on z-points (veris/averaging.py:4-19, veris/evp_solver.py:86,103)
      if (f_sidedrag && cc) sDV[t] = spd;  // dv is dead after C
      if (zs) {
        const A n = (Mc + sM[t - RJ]) + (sM[t - 1] + sM[t - RJ - 1]);
        // 1/n: 0/1 masks give n in {0, ..., 4}, whose reciprocals are exact
        // constants (1/3 folded at compile time, correctly rounded like the quotient)
        A nrm = (A)0.0;
        if (n == (A)4) nrm = (A)0.25;
        else if (n == (A)2) nrm = (A)0.5;
        else if (n == (A)1) nrm = (A)1;
        else if (n == (A)3) nrm = (A)1 / (A)3;
        else if (n > (A)0) nrm = (A)1.0 / n;
        const A zz = nrm * ((sZ[t] + sZ[t - RJ]) + (sZ[t - 1] + sZ[t - RJ - 1]));
        const A shear = ((A)2 * zz) * sE12[t];
        A am1z = (A)sc.alphaM1, den2 = (A)sc.den2;
        if (f_adaptive) {  // veris/evp_solver.py:97-100
          A aZ = (A)0.5 * (sAl[t] + sAl[t - 1]);
          aZ = (A)0.5 * (aZ + (A)0.5 * (sAl[t - RJ] + sAl[t - RJ - 1]));
          am1z = aZ - (A)1;
          den2 = (A)1.0 / aZ;
          if constexpr (RES) r_aZ = aZ;
        }
        s12 = ((s12 * am1z) + shear * (A)0.25) * den2;
        sQ[t] = s12 * met(8, f.dxV);   // sigma12*dxV (q is dead after C)
        sDU[t] = s12 * met(9, f.dyU);  // sigma12*dyU (du is dead after C)
      }
      __syncthreads();

      // ---- E: stress divergence (veris/dynamics_routines.py:284-302), forcing
      //         (veris/evp_solver.py:112-145) and momentum update (:152-187)
      if (upd) {
        const A divx = (((sS11[t] - sS11[t - RJ]) - sQ[t]) + sQ[t + 1]) * met(10, f.rrAu);
        const A divy = (((sS22[t] - sS22[t - 1]) - sDU[t]) + sDU[t + RJ]) * met(11, f.rrAv);
        const A cdc = sCD[t];
        const A cdU = (A)0.5 * (cdc + sCD[t - RJ]);
        const A cdV = (A)0.5 * (cdc + sCD[t - 1]);
        A inx = (cdU * (A)sc.cosWat) * uoc;
        A iny = (cdV * (A)sc.cosWat) * voc;
        if (f_turn) {
          const A lmU = mU != (A)0 ? (A)1 : mU;
          const A lmV = mV != (A)0 ? (A)1 : mV;
          const A X = (cdc * (A)0.5) * ((uoc - u) + (sMU[t + 1] - sU[t + 1])) +
                      (sCD[t - RJ] * (A)0.5) *
                          ((sMU[t - RJ] - sU[t - RJ]) + (sMU[t - RJ + 1] - sU[t - RJ + 1]));
          const A Y = (cdc * (A)0.5) * ((voc - v) + (sMV[t + RJ] - sV[t + RJ])) +
                      (sCD[t - 1] * (A)0.5) *
                          ((sMV[t - 1] - sV[t - 1]) + (sMV[t + RJ - 1] - sV[t + RJ - 1]));
          inx = inx - (((sg * (A)sc.sinWat) * (A)0.5) * X) * lmU;
          iny = iny + (((sg * (A)sc.sinWat) * (A)0.5) * Y) * lmV;
        }
        A fx = wfx + inx * AW;
        A fy = wfy + iny * AS;
        fx = fx + (A)0.5 * (sFV[t] + sFV[t - RJ]);
        fy = fy - (A)0.5 * (sFU[t] + sFU[t - 1]);
        A bU = (A)sc.betaU, bV = (A)sc.betaU, bP1U = (A)sc.betaP1, bP1V = (A)sc.betaP1;
        if (f_adaptive) {  // veris/evp_solver.py:148-155
          bU = ((A)0.5 * (sAl[t] + sAl[t - RJ])) * (A)sc.rdt;
          bV = ((A)0.5 * (sAl[t] + sAl[t - 1])) * (A)sc.rdt;
          bP1U = bU + (A)sc.rdt;
          bP1V = bV + (A)sc.rdt;
          if constexpr (RES) {
            r_bU = (A)0.5 * (sAl[t] + sAl[t - RJ]);
            r_bV = (A)0.5 * (sAl[t] + sAl[t - 1]);
          }
        }
        // basal drag at u and v (+0 without it, as the averages of the +0 coefficients)
        const A cbU = PL ? (A)0 : (A)0.5 * (sCB[t] + sCB[t - RJ]);
        const A cbV = PL ? (A)0 : (A)0.5 * (sCB[t] + sCB[t - 1]);
        A denU = mU * bP1U + AW * ((cdU * (A)sc.cosWat) + cbU);
        A denV = mV * bP1V + AS * ((cdV * (A)sc.cosWat) + cbV);
        denU = denU == (A)0 ? (A)1 : denU;
        denV = denV == (A)0 ? (A)1 : denV;
        if (f_sidedrag) {  // lateral drag (veris/dynamics_routines.py:113-138, evp_solver.py:169-175)
          const A spU = AW > (A)0 ? (A)0.5 * (sDV[t] + sDV[t - RJ]) : (A)0;
          const A spV = AS > (A)0 ? (A)0.5 * (sDV[t] + sDV[t - 1]) : (A)0;
          denU = denU + ((mU * (A)sc.sdCoeff) * sdmU) / (spU + (A)sc.sdU0);
          denV = denV + ((mV * (A)sc.sdCoeff) * sdmV) / (spV + (A)sc.sdU0);
        }
        const A un =
            evp_div<A, T>(MUc * (((bU * mU) * u + (mU * (A)sc.rdt) * uNm1 + fx) + divx), denU);
        const A vn =
            evp_div<A, T>(MVc * (((bV * mV) * v + (mV * (A)sc.rdt) * vNm1 + fy) + divy), denV);
        u = un;  // other threads read sU/sV, never these registers
        v = vn;
      }
      if constexpr (K == 2) {
        // the intermediate state, for callers that need every substep's state
        if (s == 0 && ab->mid[0] != nullptr && own) {
          const unsigned cp = oo + (unsigned)(i * g.NY + j);
          stg(ab->mid[0], cp, (S)u);
          stg(ab->mid[1], cp, (S)v);
          stg(ab->mid[2], cp, (S)s1);
          stg(ab->mid[3], cp, (S)s2);
          stg(ab->mid[4], cp, (S)s12);
        }
      }
    }

    // ---- residual norms of the substep (veris/evp_solver.py:192-230) over the
    //      interior: per-workgroup partial sums, reduced per substep afterwards
    if constexpr (RES && K == 1) {
      double rsig = 0.0, rvel = 0.0;
      if (own && i >= 2 && i < g.NX - 2 && j >= 2 && j < g.NY - 2) {
        const A r11 = ((((A)0.5 * (s1 + s2)) - ((A)0.5 * (r_s1o + r_s2o))) * r_aC) * Mc;
        const A r22 = ((((A)0.5 * (s1 - s2)) - ((A)0.5 * (r_s1o - r_s2o))) * r_aC) * Mc;
        const A r12 = (s12 - r_s12o) * r_aZ;
        const A ru = (MUc * (u - r_uo)) * r_bU;
        const A rv = (MVc * (v - r_vo)) * r_bV;
        rsig = ((double)r11 * (double)r11 + (double)r22 * (double)r22) + (double)r12 * (double)r12;
        rvel = (double)ru * (double)ru + (double)rv * (double)rv;
      }
      for (int o = 32; o > 0; o >>= 1) {
        rsig += __shfl_xor(rsig, o);
        rvel += __shfl_xor(rvel, o);
      }
      constexpr int NW = (NT + 63) / 64;
      __shared__ double sRed[2 * NW];
      if ((t & 63) == 0) {
        sRed[2 * (t >> 6)] = rsig;
        sRed[2 * (t >> 6) + 1] = rvel;
      }
      __syncthreads();
      if (t == 0) {
        double a0 = 0.0, a1 = 0.0;
        for (int q = 0; q < NW; ++q) {
          a0 += sRed[2 * q];
          a1 += sRed[2 * q + 1];
        }
        double* dp = ab->res + ((size_t)launch * n_tiles + pl.tile) * 2;
        dp[0] = a0;
        dp[1] = a1;
      }
    }

    // ---- results of the launch, at the cell's own padded position
    if (own) {
      const int cp = i * g.NY + j;
      if (K == 1) {
        if (upd_any) {
          store_filled(g, GW(dst[0] + off), i, j, u);
          store_filled(g, GW(dst[1] + off), i, j, v);
        }
      } else {
        stg(dst[0], oo + (unsigned)cp, (S)u);
        stg(dst[1], oo + (unsigned)cp, (S)v);
      }
      stg(dst[2], oo + (unsigned)cp, (S)s1);
      stg(dst[3], oo + (unsigned)cp, (S)s2);
      stg(dst[4], oo + (unsigned)cp, (S)s12);
    }
  };
  const Pre none{};
  if constexpr (BATCHED) {
    const int nwx = (int)(gridDim.x >> 3);
    // FP32 arithmetic with packed constants under the default settings: a
    // software pipeline whose next tile's loads are in flight while the current
    // tile computes (the Pre registers fit beside the body's at one
    // 1024-thread workgroup per compute unit; FP64 arithmetic would not)
    constexpr bool PIPE = PL && sizeof(A) == 4;
    if (PIPE && packed) {
      int tile = xcd * per_xcd + k0;
      if (k0 >= per_xcd || tile >= n_tiles) return;
      Place pl = place(tile);
      Pre cur = preload(pl);
      for (int k = k0;; k += nwx) {
        const int tn = xcd * per_xcd + k + nwx;
        const bool more = k + nwx < per_xcd && tn < n_tiles;  // uniform
        // the last tile prefetches itself again (in bounds, never used)
        const Place pn = place(more ? tn : tile);
        const Pre nxt = preload(pn);
        __builtin_amdgcn_sched_barrier(0);
        if (k != k0) __syncthreads();  // the previous tile's last stage still reads LDS
        run_tile(pl, std::true_type{}, std::true_type{}, cur);
        if (!more) break;
        tile = tn;
        pl = pn;
        cur = nxt;
      }
      return;
    }
    for (int k = k0; k < per_xcd; k += nwx) {
      const int tile = xcd * per_xcd + k;
      if (tile >= n_tiles) break;
      if (k != k0) __syncthreads();  // the previous tile's last stage still reads LDS
      if (packed) run_tile(place(tile), std::true_type{}, std::false_type{}, none);
      else run_tile(place(tile), std::false_type{}, std::false_type{}, none);
    }
  } else {
    const int tile = xcd * per_xcd + k0;
    if (tile < n_tiles) {
      if (packed) run_tile(place(tile), std::true_type{}, std::false_type{}, none);
      else run_tile(place(tile), std::false_type{}, std::false_type{}, none);
    }
  }
}

}  // namespace vh

namespace vh {

constexpr int kMaxDevices = 64;

// Launch shapes: (index, ri x rj threads, k substeps per launch); index 0 =
// default.  One list generates the table and every launch switch.  The 13 and
// 14-row shapes exist for grids whose 16 x 32 tiling leaves compute units
// idle: at 1 degree (364 x 164 padded) 13 x 32 gives 246 tiles for 256 CUs.
#define VH_EVP_SHAPES(X)                                                                      \
  X(1, 8, 64, 1) X(2, 12, 64, 1) X(3, 16, 64, 1) X(4, 8, 32, 1) X(5, 16, 32, 1)                \
  X(6, 32, 32, 1) X(7, 6, 64, 1) X(8, 16, 64, 2) X(9, 32, 32, 2) X(10, 12, 64, 2)              \
  X(11, 32, 32, 3) X(12, 16, 64, 3) X(13, 32, 32, 4) X(14, 16, 32, 2) X(15, 24, 32, 2)         \
  X(16, 13, 32, 1) X(17, 14, 32, 1) X(18, 20, 32, 2) X(19, 28, 32, 2) X(20, 28, 32, 3)

struct TileCfg {
  int ri, rj, k;
};
#define VH_ROW(id, ri, rj, k) {ri, rj, k},
constexpr TileCfg kEvpTiles[] = {{0, 0, 0}, VH_EVP_SHAPES(VH_ROW)};
#undef VH_ROW
#define VH_CHECK(id, ri_, rj_, k_)                                                   \
  static_assert(kEvpTiles[id].ri == ri_ && kEvpTiles[id].rj == rj_ && kEvpTiles[id].k == k_, \
                "shape list out of order");
VH_EVP_SHAPES(VH_CHECK)
#undef VH_CHECK
constexpr int kEvpNumTiles = sizeof(kEvpTiles) / sizeof(kEvpTiles[0]);

template <int RI_, int RJ_, int K_>
struct EvpShape {
  static constexpr int ri = RI_, rj = RJ_, k = K_;
};

// Calls f(EvpShape<ri, rj, k>{}) for shape `cfg` (callers pass a validated
// index; anything else gets 16 x 32 x 1).
template <typename F>
inline void evp_with_shape(int cfg, F&& f) {
  switch (cfg) {
#define VH_CASE(id, ri, rj, k) \
  case id: f(EvpShape<ri, rj, k>{}); break;
    VH_EVP_SHAPES(VH_CASE)
#undef VH_CASE
    default: f(EvpShape<16, 32, 1>{}); break;
  }
}
// single-substep shape used for the n % k remainder of a subcycle
constexpr int kEvpRemainderCfg = 5;

inline int evp_tile_extent_i(int cfg) { return kEvpTiles[cfg].ri - 4 * kEvpTiles[cfg].k; }
inline int evp_tile_extent_j(int cfg) { return kEvpTiles[cfg].rj - 4 * kEvpTiles[cfg].k; }
inline int evp_n_launches(int cfg, int nsub) {
  return nsub / kEvpTiles[cfg].k + nsub % kEvpTiles[cfg].k;
}

// Workgroups of a launch: one per tile, capped at what the device holds
// resident at once (compute units x resident workgroups of `kernel`, cached
// per device in `cache`), so a large batch runs as a grid that stays resident
// and loops over its tiles instead of waiting on workgroup dispatch.
template <typename F>
inline int evp_grid(F kernel, int threads, int n_tiles, std::atomic<int>* cache) {
  const int per_xcd = (n_tiles + 7) / 8;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 8 * per_xcd;
  int cap = cache[dev].load(std::memory_order_relaxed);
  if (cap == 0) {
    int ncu = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
        ncu <= 0 || per_cu <= 0)
      return 8 * per_xcd;
    cap = ncu * per_cu;
    cache[dev].store(cap, std::memory_order_relaxed);
  }
  const int cap_x = (cap + 7) / 8;
  return 8 * (per_xcd < cap_x ? per_xcd : cap_x);
}

template <typename T, typename TD, int RI, int RJ, int K, bool BATCHED, typename A = T,
          bool SO = false, typename S = T, bool PL = false>
inline void evp_launch_shape(const EvpArgBlock<T, S>* ab, int which, int launch, int n_tiles,
                             int tiles_j, hipStream_t st) {
  static std::atomic<int> cache[kMaxDevices];
  const int per_xcd = (n_tiles + 7) / 8;
  auto kernel = evp_substep_kernel<T, TD, RI, RJ, K, BATCHED, A, SO, false, S, PL>;
  const int grid = BATCHED ? evp_grid(kernel, RI * RJ, n_tiles, cache) : 8 * per_xcd;
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(RI * RJ), 0, st, ab, which, launch, n_tiles,
                     tiles_j, per_xcd);
}

// Single-substep launch with the second-order no-slip boundary term (cfg must
// be a k = 1 shape; the host checks).
template <typename T, typename TD, bool BATCHED, typename A = T, typename S = T>
inline void evp_launch_so(int cfg, const EvpArgBlock<T, S>* ab, int which, int launch, int n_tiles,
                          int tiles_j, hipStream_t st) {
  evp_with_shape(cfg, [&](auto sh) {
    using Sh = decltype(sh);
    constexpr int RI = Sh::k == 1 ? Sh::ri : 16, RJ = Sh::k == 1 ? Sh::rj : 32;
    evp_launch_shape<T, TD, RI, RJ, 1, BATCHED, A, true, S>(ab, which, launch, n_tiles, tiles_j,
                                                            st);
  });
}

// Single-domain single-substep launch that also writes the substep's residual
// partial sums (shape 16 x 32, the library's default single-substep shape).
template <typename T, bool SO>
inline void evp_launch_res(const EvpArgBlock<T>* ab, int launch, int n_tiles, int tiles_j,
                           hipStream_t st) {
  const int per_xcd = (n_tiles + 7) / 8;
  hipLaunchKernelGGL((evp_substep_kernel<T, double, 16, 32, 1, false, T, SO, true>),
                     dim3(per_xcd * 8), dim3(16 * 32), 0, st, ab, 0, launch, n_tiles, tiles_j,
                     per_xcd);
}
constexpr int kEvpResidualCfg = 5;

// True when the settings take none of the optional branches of the substep
// (metric terms, turning angle, basal drag, adaptive EVP, free slip, the
// second-order boundary term) and tensileStrFac = 0: the default settings,
// which launch the kernel instantiation with those branches compiled out.
inline bool evp_plain_settings(const EvpScalars& sc) {
  return !sc.kterms && !sc.turn && !sc.basal && !sc.adaptive && !sc.sidedrag && !sc.sobc &&
         sc.tens0;
}

// One launch of shape `cfg`: tiles_j is used by single-domain launches, `which`
// (0 main shape, 1 remainder shape) by batched ones.  T is the storage type
// of the fields, A the arithmetic type (default T), TD that of Delta/zeta;
// `so` selects the second-order boundary variant (k = 1 shapes only), `plain`
// the default-settings instantiation (evp_plain_settings of the launch's
// scalars; the same results bit for bit as the general one).
template <typename T, typename TD, bool BATCHED, typename A = T, typename S = T>
inline void evp_launch_cfg(int cfg, const EvpArgBlock<T, S>* ab, int which, int launch, int n_tiles,
                           int tiles_j, hipStream_t st, bool so = false, bool plain = false) {
  if (so) {
    evp_launch_so<T, TD, BATCHED, A, S>(cfg, ab, which, launch, n_tiles, tiles_j, st);
    return;
  }
  evp_with_shape(cfg, [&](auto sh) {
    using Sh = decltype(sh);
    if (plain)
      evp_launch_shape<T, TD, Sh::ri, Sh::rj, Sh::k, BATCHED, A, false, S, true>(
          ab, which, launch, n_tiles, tiles_j, st);
    else
      evp_launch_shape<T, TD, Sh::ri, Sh::rj, Sh::k, BATCHED, A, false, S>(ab, which, launch,
                                                                           n_tiles, tiles_j, st);
  });
}

// Single-domain launch of the default-settings instantiation of shape `cfg`
// (a shape with k = 2, or the 32 x 32 shape with k = 4; anything else
// launches nothing) reading its arguments from `ab` (nlaunch = 1: in -> out;
// with k = 2 the state after the first substep to ab->mid when set), for
// n_members equal members packed back to back.
template <typename T>
inline void evp_launch_pair(int cfg, const EvpArgBlock<T>* ab, int n_tiles, int tiles_j,
                            hipStream_t st, int n_members = 1) {
  const int per_xcd = (n_tiles + 7) / 8;
  evp_with_shape(cfg, [&](auto sh) {
    using Sh = decltype(sh);
    if constexpr (Sh::k == 2 || (Sh::k == 4 && Sh::ri == 32 && Sh::rj == 32))
      hipLaunchKernelGGL(
          (evp_substep_kernel<T, double, Sh::ri, Sh::rj, Sh::k, false, T, false, false, T, true>),
          dim3(per_xcd * 8, n_members), dim3(Sh::ri * Sh::rj), 0, st, ab, 0, 0, n_tiles, tiles_j,
          per_xcd);
  });
}

// The 32 x 32 four-substep shape.
inline int evp_quad_cfg() {
  for (int c = 1; c < kEvpNumTiles; ++c)
    if (kEvpTiles[c].ri == 32 && kEvpTiles[c].rj == 32 && kEvpTiles[c].k == 4) return c;
  return 0;
}

// Single-domain, single-substep launch of shape `cfg` (a shape with k = 1)
// reading its arguments from `ab` (nlaunch = 1: in -> out), for n_members
// equal members packed back to back (ab->mstride apart; 1 = one domain);
// `so` and `plain` select the instantiation as for evp_launch_cfg.
template <typename T, typename TD>
inline void evp_launch_single(int cfg, const EvpArgBlock<T>* ab, int n_tiles, int tiles_j,
                              hipStream_t st, int n_members = 1, bool so = false,
                              bool plain = false) {
  const int per_xcd = (n_tiles + 7) / 8;
  evp_with_shape(cfg, [&](auto sh) {
    using S = decltype(sh);
    constexpr int RI = S::k == 1 ? S::ri : 16, RJ = S::k == 1 ? S::rj : 32;
    if (so)  // the second-order boundary term's instantiation
      hipLaunchKernelGGL((evp_substep_kernel<T, TD, RI, RJ, 1, false, T, true>),
                         dim3(per_xcd * 8, n_members), dim3(RI * RJ), 0, st, ab, 0, 0, n_tiles,
                         tiles_j, per_xcd);
    else if (plain)  // the default settings' instantiation
      hipLaunchKernelGGL((evp_substep_kernel<T, TD, RI, RJ, 1, false, T, false, false, T, true>),
                         dim3(per_xcd * 8, n_members), dim3(RI * RJ), 0, st, ab, 0, 0, n_tiles,
                         tiles_j, per_xcd);
    else
      hipLaunchKernelGGL((evp_substep_kernel<T, TD, RI, RJ, 1, false>), dim3(per_xcd * 8, n_members),
                         dim3(RI * RJ), 0, st, ab, 0, 0, n_tiles, tiles_j, per_xcd);
  });
}

}  // namespace vh
