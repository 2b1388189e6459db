// Hand-written adjoint of the EVP subcycle (veris/evp_solver.py:22-319) with
// segment checkpointing, and of the pre-EVP forcing it depends on
// (veris/dynamics_routines.py:5-18, veris/area_mass.py:17-28,
// veris/dynsolver.py:7-84).
//
// Forward (vh_evp_forward_checkpointed): n single-substep launches of the
// fused substep kernel; the state after every `ckpt_every` substeps stays in
// the workspace.  Reverse (vh_evp_adjoint): for each segment, last first, the
// segment's states are recomputed from its checkpoint by the same kernel, then
// each substep is reversed by one tiled kernel (adj_substep_kernel) that
// recomputes the substep's intermediates in LDS and runs the adjoint stages
// in gather form -- each cell sums what its neighbours contribute, so no
// atomics are needed.  Every launch reads its pointers from argument blocks
// in the workspace that one small kernel writes per call, so the n forward
// launches and the whole reverse sweep are each replayed from one hipGraph.
//
// Non-smooth points: d sqrt(x)/dx is taken as 0 at x = 0 (Delta = 0 on
// motionless or masked cells, zero relative velocity); the denominator guard
// and the drag regime switch contribute no derivative.  Every setting of the
// subcycle is supported: fixed or adaptive relaxation, no-slip boundaries
// (with or without the second-order term) or free slip with lateral drag,
// with or without basal drag, the turning angle and the metric terms.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>

#include "evp_kernel.h"

namespace vh {
namespace {

constexpr int BLK = 256;

template <typename T>
struct AdjAcc {  // parameter adjoints, accumulated over the substeps
  T *P, *mC, *mU, *mV, *AW, *AS, *WFX, *WFY, *u0, *v0, *DN, *DS, *h;
};
constexpr int kAdjAccFields = 13;

// What every reverse launch reads besides workspace-relative pointers.
template <typename T>
struct AdjArgs {
  EvpConst<T> c;
  EvpScalars sc;
  Geom g;
  AdjAcc<T> acc;
  const T* lam_final[5];  // adjoints of the final state (NULL = 0)
  T* outs[5];             // adjoints of the initial state
};

template <typename T>
struct LamSet {  // state adjoints: of the substep's output (in) and input (out)
  const T* in[5];
  T* out[5];
};

// Where the states of a subcycle live (all pointers into the workspace
// except the initial state in0 and the final state outs).
template <typename T>
struct BlockPlan {
  const T* in0[5];
  T* outs[5];
  char* ws;
  size_t fb, off_ckpt, off_pp, off_seg, seg_stride;
  int n, ce, nck;
  int kstep;  // substeps per forward launch (1, or 2: ce and n even)
};

#define CELL_SETUP_G                                                                \
  const int n_ = g.NX * g.NY;                                                       \
  const int c = blockIdx.x * blockDim.x + threadIdx.x;                              \
  if (c >= n_) return;                                                              \
  const int i = c / g.NY, j = c - (c / g.NY) * g.NY;                                \
  const int im = i == 0 ? g.NX - 1 : i - 1, ip = i == g.NX - 1 ? 0 : i + 1;         \
  const int jm = j == 0 ? g.NY - 1 : j - 1, jp = j == g.NY - 1 ? 0 : j + 1;         \
  (void)im; (void)ip; (void)jm; (void)jp;
#define CELL_SETUP         \
  const Geom g = a->g;     \
  CELL_SETUP_G
#define AT(ii, jj) ((ii) * g.NY + (jj))

__device__ inline int wm(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ inline int wp(int i, int n) { return i == n - 1 ? 0 : i + 1; }

// ---- argument blocks
template <typename T>
__device__ inline T* slot(const BlockPlan<T>& p, size_t off, int idx) {
  return (T*)(p.ws + off + (size_t)idx * p.fb);
}

// Segment slots of segment sgm: the last segment's states, kept from the
// checkpointed forward for every reverse sweep, have a set of their own; the
// other segments alternate between two sets (one is recomputed while the
// other is reversed).
template <typename T>
__device__ inline size_t seg_set(const BlockPlan<T>& p, int sgm) {
  const int set = sgm != p.nck - 1 ? sgm % 2 : (p.nck == 1 ? 0 : 2);
  return p.off_seg + (size_t)set * p.seg_stride;
}

// forward location of field q of the state after k substeps: the states
// inside the last segment go to that segment's slots, where the reverse
// sweep reads them without recomputing them; between earlier checkpoints the
// states (k a multiple of p.kstep) alternate between two slots
template <typename T>
__device__ inline T* fwd_loc(const BlockPlan<T>& p, int k, int q) {
  if (k == 0) return (T*)p.in0[q];
  if (k == p.n) return p.outs[q];
  if (k % p.ce == 0) return slot(p, p.off_ckpt, 5 * (k / p.ce) + q);
  const int last = p.nck - 1;
  if (k / p.ce == last) return slot(p, seg_set(p, last), 5 * (k - last * p.ce) + q);
  return slot(p, p.off_pp, 5 * (((k - 1) / p.kstep) % 2) + q);
}

// mode 0: forward blocks (substep k: state k -> state k+1);
// mode 1: recompute blocks and the reverse arguments.  Substep k of segment
// sgm reads state k from the checkpoint (first substep) or a segment slot and
// names state k+1 as its output: a segment slot, or for the segment's last
// substep the next checkpoint (slot nck holds a copy of the final state),
// which the reverse sweep reads and never recomputes;
// mode 2: two-substep forward blocks (block m: state 2m -> state 2m+2);
// mode 4: forward blocks of four-substep launches before the last segment
// and two-substep launches in it (block m: launch m of that sequence);
// mode 3: two-substep recompute blocks, at every even substep k of a segment
// short of its last: state k -> state k+2 with state k+1 kept (mid), all in
// the segment's slots (state k+2 goes to the spare slot len when it is the
// next checkpoint, which the reverse sweep reads from its own slot).
template <typename T>
__global__ void adj_write_blocks_kernel(EvpArgBlock<T>* blocks, size_t stride, EvpArgBlock<T> proto,
                                        BlockPlan<T> p, int mode, AdjArgs<T>* aa, AdjArgs<T> av,
                                        const EvpArgBlock<T>* packed_from) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (mode == 1 && k == 0) *aa = av;
  if (k >= p.n) return;
  EvpArgBlock<T> b = proto;
  // packed step constants: usable when the pack kernel left its flag set
  b.packed = packed_from ? packed_from->packed : 0;
  if (mode == 0) {
    for (int q = 0; q < 5; ++q) {
      b.in[q] = fwd_loc(p, k, q);
      b.out[q] = fwd_loc(p, k + 1, q);
    }
  } else if (mode == 2) {
    if (2 * k >= p.n) return;
    // the intermediate state is kept where the last segment's states go
    const bool keep_mid = (2 * k + 1) / p.ce == p.nck - 1;
    for (int q = 0; q < 5; ++q) {
      b.in[q] = fwd_loc(p, 2 * k, q);
      b.mid[q] = keep_mid ? fwd_loc(p, 2 * k + 1, q) : nullptr;
      b.out[q] = fwd_loc(p, 2 * k + 2, q);
    }
  } else if (mode == 4) {
    // forward launches with four substeps: each segment before the last runs
    // ce/4 four-substep launches (and one two-substep launch when ce/2 is
    // odd) from its checkpoint to the next through the two ping-pong slots;
    // the last segment runs two-substep launches that keep every state
    const int q4 = p.ce / 4, lps = q4 + (p.ce % 4) / 2, head = (p.nck - 1) * lps;
    if (k < head) {
      const int sgm = k / lps, j = k % lps;
      for (int q = 0; q < 5; ++q) {
        b.in[q] = j == 0 ? (sgm == 0 ? p.in0[q] : slot(p, p.off_ckpt, 5 * sgm + q))
                         : slot(p, p.off_pp, 5 * ((j - 1) % 2) + q);
        b.mid[q] = nullptr;
        b.out[q] = j == lps - 1 ? slot(p, p.off_ckpt, 5 * (sgm + 1) + q)
                                : slot(p, p.off_pp, 5 * (j % 2) + q);
      }
    } else {
      const int k0 = (p.nck - 1) * p.ce + 2 * (k - head);
      if (k0 >= p.n) return;
      for (int q = 0; q < 5; ++q) {
        b.in[q] = fwd_loc(p, k0, q);
        b.mid[q] = fwd_loc(p, k0 + 1, q);
        b.out[q] = fwd_loc(p, k0 + 2, q);
      }
    }
  } else if (mode == 3) {
    const int jj = k % p.ce, sgm = k / p.ce;
    const int len = min(p.ce, p.n - sgm * p.ce);
    if (jj % 2 != 0 || jj + 1 > len - 1) return;
    const size_t seg = seg_set(p, sgm);
    for (int q = 0; q < 5; ++q) {
      b.in[q] = jj == 0 ? (k == 0 ? p.in0[q] : slot(p, p.off_ckpt, 5 * sgm + q))
                        : slot(p, seg, 5 * jj + q);
      b.mid[q] = slot(p, seg, 5 * (jj + 1) + q);
      b.out[q] = slot(p, seg, 5 * (jj + 2) + q);
    }
  } else {
    const int jj = k % p.ce, sgm = k / p.ce;
    const int len = min(p.ce, p.n - sgm * p.ce);
    const size_t seg = seg_set(p, sgm);
    for (int q = 0; q < 5; ++q) {
      b.in[q] = jj == 0 ? (k == 0 ? p.in0[q] : slot(p, p.off_ckpt, 5 * sgm + q))
                        : slot(p, seg, 5 * jj + q);
      b.out[q] = jj == len - 1 ? slot(p, p.off_ckpt, 5 * (sgm + 1) + q)
                               : slot(p, seg, 5 * (jj + 1) + q);
    }
  }
  *(EvpArgBlock<T>*)((char*)blocks + (size_t)k * stride) = b;
}

// copy of the final state (the forward's last outputs) into checkpoint slot nck
template <typename T>
__global__ __launch_bounds__(BLK) void adj_keep_final_kernel(const EvpArgBlock<T>* __restrict__ last,
                                                             int n_all, T* dst, size_t fb) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_all) return;
#pragma unroll
  for (int q = 0; q < 5; ++q)
    stg((T*)((char*)dst + (size_t)q * fb), (unsigned)c, ldg((const T*)last->out[q], (unsigned)c));
}

// ---- reverse sweep: set-up and result copy
// pointwise over every cell of every member
template <typename T>
__global__ __launch_bounds__(BLK) void adj_init_kernel(const AdjArgs<T>* __restrict__ a,
                                                       LamSet<T> L, int n_all) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_all) return;
  const AdjAcc<T>& acc = a->acc;
  T* const accp[kAdjAccFields] = {acc.P,   acc.mC,  acc.mU, acc.mV, acc.AW, acc.AS, acc.WFX,
                                  acc.WFY, acc.u0, acc.v0, acc.DN, acc.DS, acc.h};
#pragma unroll
  for (int q = 0; q < kAdjAccFields; ++q)
    if (accp[q]) accp[q][c] = (T)0;
#pragma unroll
  for (int q = 0; q < 5; ++q) L.out[q][c] = a->lam_final[q] ? a->lam_final[q][c] : (T)0;
}

template <typename T>
__global__ __launch_bounds__(BLK) void adj_finish_kernel(const AdjArgs<T>* __restrict__ a,
                                                         LamSet<T> L, int n_all) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_all) return;
#pragma unroll
  for (int q = 0; q < 5; ++q) a->outs[q][c] = L.in[q][c];
}

// Loads and stores below go through the global address space with 32-bit
// offsets (ldg/stg): every index is a valid cell of the padded field, and the
// kernels issue their loads before any data-dependent selection.

// sum of a u/v adjoint over a computed cell and the overlap cells the fill
// copies it to (the adjoint of the periodic fill); branch-free: absent
// partners read the cell itself with weight 0
template <typename T>
__device__ inline T gather_filled(const Geom& g, const T* a, int i, int j) {
  int ix = i, jy = j;
  bool hx = false, hy = false;
  if (g.fill_x) {
    hx = (i < 2 * g.ol) || (i >= g.nx);
    ix = i < 2 * g.ol ? i + g.nx : (i >= g.nx ? i - g.nx : i);
  }
  if (g.fill_y) {
    hy = (j < 2 * g.ol) || (j >= g.ny);
    jy = j < 2 * g.ol ? j + g.ny : (j >= g.ny ? j - g.ny : j);
  }
  const T a00 = ldg(a, AT(i, j)), a10 = ldg(a, AT(ix, j)), a01 = ldg(a, AT(i, jy)),
          a11 = ldg(a, AT(ix, jy));
  return ((a00 + (hx ? a10 : (T)0)) + (hy ? a01 : (T)0)) + ((hx && hy) ? a11 : (T)0);
}

// ---- the reverse of one substep, one workgroup per tile.
//
// Threads cover a region of RI x RJ cells around a TI x TJ tile (ring of 4).
// Stages, each on a shrinking part of the region (rows/columns counted from
// the region's low edge, "hi" from its high edge):
//   A  load the substep's input state and constants          all cells
//   B  e12 at z-points (forward, veris/dynamics_routines.py:160-175)      >= 1
//   C  zeta, Delta, e11 +- e22, water drag at c-points (forward)     >= 1, hi >= 1
//   S1 adjoint of the u/v momentum update + fill (veris/evp_solver.py:152-191)
//                                                                    >= 2, hi >= 1
//   S2 adjoint of the stress divergence, sigma relaxation, drag and Coriolis
//      (veris/dynamics_routines.py:284-302, veris/evp_solver.py:79-145)
//                                                                    >= 3, hi >= 2
//   S3 adjoint of zeta and Delta (veris/dynamics_routines.py:220-268) >= 3, hi >= 3
//   S4 adjoint of e12 and of e11, e22, drag, Coriolis w.r.t. the input
//      velocities                                                 tile cells
// The forward intermediates B, C are recomputed from the input state exactly
// as the forward kernel computes them; nothing but the two states and the
// state adjoints is read from or written to memory per substep, plus the
// parameter adjoints of the tile's own cells.
//
// XB selects the instantiation for the extra drag branches, each a runtime
// flag inside it: free slip with lateral drag (noSlip = False: no boundary
// term in e12, and the lateral drag of each u/v point in its momentum
// denominator, whose adjoint reaches the ice speed at c;
// veris/dynamics_routines.py:98-140, veris/evp_solver.py:169-176), and basal
// drag (the c-point coefficient in the denominators, whose adjoint reaches
// the velocities and the thickness; veris/dynamics_routines.py:55-96), and
// adaptive EVP (relaxation weights alpha at c and z and beta at u/v points
// from zeta and the mass; veris/evp_solver.py:65-73, 97-100, 148-155).
//
// MODE 0 is the default instantiation, 1 the extra branches, 2 the extra
// branches and the second-order no-slip boundary term
// (veris/dynamics_routines.py:177-215): its stencil reaches two cells along
// the face normal, so MODE 2 recomputes one more ring of cells (R = 5);
// MODE 3 is MODE 0 under the default settings (no turning angle, no metric
// terms, tensileStrFac = 0) with those runtime branches compiled out.
template <typename T, typename TD, int RI, int RJ, int MODE>
__global__ __launch_bounds__(RI* RJ) void adj_substep_kernel(const AdjArgs<T>* __restrict__ a,
                                                             const EvpArgBlock<T>* __restrict__ blk,
                                                             LamSet<T> L, int n_tiles, int tiles_j,
                                                             int per_xcd) {
  constexpr bool XB = MODE == 1 || MODE == 2, SO = MODE == 2, PLN = MODE == 3;
  constexpr int R = SO ? 5 : 4, TI = RI - 2 * R, TJ = RJ - 2 * R, NT = RI * RJ;
  static_assert((RJ & (RJ - 1)) == 0, "RJ must be a power of two");
  static_assert(TI >= 1 && TJ >= 1, "tile too small");
  __shared__ T sU[NT], sV[NT], sM[NT], sMU[NT], sMV[NT], sA[NT], sB[NT], sE12[NT], sQ[NT], sZ[NT];
  __shared__ T sCD[NT], sNV[NT], sCXU[NT], sCXV[NT], sAB[NT], sBB[NT], sFU[NT], sFV[NT], sEP[NT];
  // The water-stress turning term (sinWat != 0) borrows arrays outside their
  // own lifetimes: the ocean-minus-ice velocities sit in sFU/sFV (A to C),
  // their face sums in sAB/sBB (C to S2), the u/v-point factors
  // nb sgn(f) lm Area of its adjoint in sU/sV (S1 to S2) and the velocity
  // adjoint partials in sU/sV again (S3 to S4).
  // Lateral drag: the ice speed at c (C to S1) and the speed adjoints of the
  // u/v points (S1 to S2); the velocity adjoint partials of the speed go
  // through sZ/sCXV, free after S2.  Basal drag: the coefficient at c (C to
  // S1), the denominator adjoints times the area at the u/v points (S1 to
  // S2) and the adjoint of the coefficient's velocity argument at c (S3 to S4).
  // Adaptive EVP: alpha at c (C to S2), the beta adjoints of the u/v points
  // (S1 to S2) and the alpha adjoints of the z-points (S2 to S3).
  constexpr int NX_ = XB ? NT : 1;
  __shared__ T sSPD[NX_], sLSU[NX_], sLSV[NX_], sCBc[NX_], sCBU[NX_], sCBV[NX_], sLTB[NX_];
  __shared__ T sAl[NX_], sLBU[NX_], sLBV[NX_], sLAZ[NX_];

  const int bid = blockIdx.x;
  const int tile = (bid & 7) * per_xcd + (bid >> 3);
  if (tile >= n_tiles) return;
  const Geom g = a->g;
  const int ti = tile / tiles_j;
  const int i0 = ti * TI, j0 = (tile - ti * tiles_j) * TJ;
  const EvpConst<T>& k = a->c;
  const EvpScalars& sc = a->sc;
  const AdjAcc<T>& acc = a->acc;
  const bool f_turn = !PLN && sc.turn, f_kterms = !PLN && sc.kterms, f_tens0 = PLN || sc.tens0;

  const int t = threadIdx.x;
  const int ar = t / RJ, br = t & (RJ - 1);
  const int i = i0 + ar - R, j = j0 + br - R;
  const int gi = wrap(i, g.NX), gj = wrap(j, g.NY);
  const int gip = wp(gi, g.NX), gjp = wp(gj, g.NY), gim = wm(gi, g.NX), gjm = wm(gj, g.NY);
  // equal members packed back to back: member blockIdx.y starts at element mo
  const unsigned mo = blockIdx.y * (unsigned)(g.NX * g.NY);
  const unsigned c = mo + gi * g.NY + gj;
  const unsigned cip = mo + gip * g.NY + gj, cjp = mo + gi * g.NY + gjp;
  const unsigned cim = mo + gim * g.NY + gj, cjm = mo + gi * g.NY + gjm;
  const bool own = (ar >= R) && (ar < R + TI) && (br >= R) && (br < R + TJ) && (i < g.NX) &&
                   (j < g.NY);
  const bool comp = is_computed(g, gi, gj);

  // ---- A: every load of this cell, issued together.  With the packed step
  //      constants of the forward launches (uniform flag, set when every mask
  //      is 0/1 and every metric constant along i), the masks come from one
  //      byte and the metrics from rows in LDS, as in the forward kernel.
  // the body is instantiated twice, for packed step constants and for full
  // fields (pk: std::true_type / std::false_type), each with its own schedule
  // and registers; the packed flag of the forward launches picks one
  auto body = [&](auto pk) {
    constexpr bool packed = decltype(pk)::value;
    constexpr int RPT = (kEvpRowFields * RJ + NT - 1) / NT;  // rows per thread
    __shared__ T sRow[kEvpRowFields * RJ];
    T rowv[RPT];
    if (packed) {
      const int roff = (int)blockIdx.y * g.NY;  // rows of member blockIdx.y
  #pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int q = t + r * NT;
        rowv[r] = (T)0;
        if (q < kEvpRowFields * RJ) {
          const int mq = q / RJ, bq = q & (RJ - 1);
          rowv[r] = ldg(blk->rows + (long long)mq * blk->rstride + roff,
                        (unsigned)wrap(j0 + bq - R, g.NY));
        }
      }
    }
    const T* xu = blk->in[0];
    const T* xv = blk->in[1];
    const T u = ldg(xu, c), v = ldg(xv, c);
    T Mc, MUc, MVc, mInC, mInUc, mInVc;
    if (packed) {
      const unsigned pkb = ldg(blk->pk, c);
      Mc = (T)(pkb & 1u);
      MUc = (T)((pkb >> 1) & 1u);
      MVc = (T)((pkb >> 2) & 1u);
      mInC = (T)((pkb >> 3) & 1u);
      mInUc = (T)((pkb >> 4) & 1u);
      mInVc = (T)((pkb >> 5) & 1u);
    } else {
      Mc = ldg(k.M, c);
      MUc = ldg(k.MU, c);
      MVc = ldg(k.MV, c);
      mInC = ldg(k.mInC, c);
      mInUc = ldg(k.mInU, c);
      mInVc = ldg(k.mInV, c);
    }
    const T uoc = ldg(k.uOc, c), voc = ldg(k.vOc, c);
    // metrics from the full fields (not packed); met(q, .) picks the source
    T g_met[kEvpRowFields] = {};
    if (!packed) {
      const T* const mp[kEvpRowFields] = {k.rdyU, k.rdxV, k.rAz, k.rdxU, k.rdyV, k.rrA,
                                          k.dyV,  k.dxU,  k.dxV, k.dyU,  k.rrAu, k.rrAv};
  #pragma unroll
      for (int q = 0; q < kEvpRowFields; ++q) g_met[q] = ldg(mp[q], c);
    }
    auto met = [&](int q) -> T { return packed ? sRow[q * RJ + br] : g_met[q]; };
    const T P = ldg(k.P, c), fc = ldg(k.fC, c);
    const T un = ldg(blk->out[0], c), vn = ldg(blk->out[1], c);
    const T mU = ldg(k.massU, c), mV = ldg(k.massV, c), AW = ldg(k.AreaW, c), AS = ldg(k.AreaS, c);
    const T mC = ldg(k.massC, c);
    const bool sdr = XB && sc.sidedrag;  // uniform; false in the default instantiation
    const bool bas = XB && sc.basal;
    const bool ada = XB && sc.adaptive;
    T so1 = 0, so2 = 0, so12 = 0;  // input stresses (adaptive: the alpha derivatives)
    if (ada) {
      so1 = ldg(blk->in[2], c);
      so2 = ldg(blk->in[3], c);
      so12 = ldg(blk->in[4], c);
    }
    T FuC = 0, FvC = 0;  // coastline form factors
    if (sdr && sc.coast) {
      FuC = ldg(k.Fu, c);
      FvC = ldg(k.Fv, c);
    }
    T bAr = 0, bRl = 0, bh = 0, bmUip = 0, bmVjp = 0;  // basal drag inputs
    if (bas) {
      bAr = ldg(k.Area, c);
      bRl = ldg(k.Rlow, c);
      bh = ldg(k.hIce, c);
      bmUip = ldg(k.mInU, cip);
      bmVjp = ldg(k.mInV, cjp);
    }
    const T l1 = ldg(L.in[2], c), l2 = ldg(L.in[3], c), l12 = ldg(L.in[4], c);
    const T lu = gather_filled(g, L.in[0] + mo, gi, gj), lv = gather_filled(g, L.in[1] + mo, gi, gj);
    // water drag factor of either hemisphere, selected by value (a select of
    // the two setting addresses would become a per-lane load)
    const T Dn = (T)sc.cwN * (T)sc.rhoSea, Ds = (T)sc.cwS * (T)sc.rhoSea;
    const T D = fc < (T)0 ? Ds : Dn;

    if (packed) {
  #pragma unroll
      for (int r = 0; r < RPT; ++r)
        if (t + r * NT < kEvpRowFields * RJ) sRow[t + r * NT] = rowv[r];
    }
    sU[t] = u;
    sV[t] = v;
    sM[t] = Mc;
    sMU[t] = MUc;
    sMV[t] = MVc;
    sA[t] = (u - uoc) * mInUc;
    sB[t] = (v - voc) * mInVc;
    if (f_turn) {
      sFU[t] = uoc - u;
      sFV[t] = voc - v;
    }
    __syncthreads();

    // ---- B: e12 at z-points (the forward kernel's expression order)
    T e12r = 0;
    if (ar >= 1 && br >= 1) {
      const T rdyU = met(0), rdxV = met(1), rAz = met(2);
      const T ujm = sU[t - 1], vim = sV[t - RJ];
      const T dudy = (u - ujm) * rdyU;
      const T uavz = (u + ujm) * (T)0.5;
      const T dvdx = (v - vim) * rdxV;
      const T vavz = (v + vim) * (T)0.5;
      const T mz = (Mc * sM[t - RJ]) * (sM[t - 1] * sM[t - RJ - 1]);
      T inner = dudy + dvdx;
      if (f_kterms) inner = (inner - ldg(k.k1Z, c) * vavz) - ldg(k.k2Z, c) * uavz;
      T e12 = ((T)0.5 * inner) * mz;
      if (!sdr) {  // no-slip boundary term
        const T hu = MUc - sMU[t - 1];
        const T hv = MVc - sMV[t - RJ];
        e12 = e12 + ((((T)2.0 * uavz) * rdyU) * hu + (((T)2.0 * vavz) * rdxV) * hv);
      }
      if (SO && sc.sobc) {
        // second-order boundary shear, read from memory as the forward kernel
        // reads it (two cells along the face normal reach past this region's edge)
        const unsigned cim2 = mo + wrap(i - 2, g.NX) * g.NY + gj, cim1 = mo + wrap(i - 1, g.NX) * g.NY + gj;
        const unsigned cjm2 = mo + gi * g.NY + wrap(j - 2, g.NY), cjm1 = mo + gi * g.NY + wrap(j - 1, g.NY);
        const T uim2 = ldg(xu, cim2), uip1 = ldg(xu, cip);
        const T MUim2 = ldg(k.MU, cim2), MUim1 = ldg(k.MU, cim1), MUip1 = ldg(k.MU, cip);
        const T vjm2 = ldg(xv, cjm2), vjp1 = ldg(xv, cjp);
        const T MVjm2 = ldg(k.MV, cjm2), MVjm1 = ldg(k.MV, cjm1), MVjp1 = ldg(k.MV, cjp);
        T hu2 = (MUc - MUim1) / (T)3.0;
        T hv2 = (MVc - MVjm1) / (T)3.0;
        hu2 = hu2 * (MUim2 * MUim1 + MUip1 * MUc);
        hv2 = hv2 * (MVjm2 * MVjm1 + MVjp1 * MVc);
        const T tu = (rdyU * ((((T)6.0 * uavz) - uim2 * MUim1) - uip1 * MUc)) * hu2;
        const T tv = (rdxV * ((((T)6.0 * vavz) - vjm2 * MVjm1) - vjp1 * MVc)) * hv2;
        e12 = e12 + (T)0.5 * (tu + tv);
      }
      sE12[t] = e12;
      sQ[t] = rAz * (e12 * e12);
      e12r = e12;
    }
    __syncthreads();

    // ---- C: zeta, Delta, strain invariants, water drag at c-points
    T zeta = 0, dC = 0, ep = 0, em = 0, sa = 0, sb = 0, ww = 0, sw = 0;
    T su = 0, sv = 0, spd = 0;  // lateral drag: u + u(i+1), v + v(j+1) and the speed at c
    T cTB = 0, cH = 0;          // basal drag: its derivatives at c
    T aC = 0, cZa = 0, cMa = 0, prs = 0;  // adaptive: alpha at c, d alpha / d (zeta, mC), press
    if (ar >= 1 && br >= 1 && ar <= RI - 2 && br <= RJ - 2) {
      const T uip = sU[t + RJ], vjp = sV[t + 1];
      const T rdxU = met(3), rdyV = met(4);
      const T dudx = (uip - u) * rdxU, dvdy = (vjp - v) * rdyV;
      T e11 = dudx, e22 = dvdy;
      if (f_kterms) {
        e11 = dudx + ((vjp + v) * (T)0.5) * ldg(k.k2C, c);
        e22 = dvdy + ((uip + u) * (T)0.5) * ldg(k.k1C, c);
      }
      e11 = e11 * mInC;
      e22 = e22 * mInC;
      const T qa = sQ[t] + sQ[t + RJ];
      const T qb = sQ[t + 1] + sQ[t + RJ + 1];
      const T e12c = ((T)0.25 * met(5)) * (qa + qb);
      if (sdr) {  // ice speed at c (veris/dynamics_routines.py:109-111)
        su = u + uip;
        sv = v + vjp;
        spd = (T)0.5 * sqrt(su * su + sv * sv);
        sSPD[t] = spd;
      }
      if (bas) {  // basal drag coefficient at c, the forward kernel's expression order
        const T um = u * mInUc, umip = uip * bmUip;
        const T vm = v * mInVc, vmjp = vjp * bmVjp;
        const T tbs = (T)0.25 * (((um * um + umip * umip) + vm * vm) + vmjp * vmjp) + (T)sc.U0sq;
        const T tb = (T)sc.K2 / sqrt(tbs);
        const T hc = (fabs(bRl) * bAr) / (T)sc.K1;
        const T ex = exp((T)10.0 * (bh - hc));
        const T lg = log(ex + (T)1.0);
        const T eb = exp((T)(-sc.cBasalStar) * ((T)1.0 - bAr));
        const T val = ((tb * lg) * (T)0.1) * eb;
        const bool on = bAr > (T)0.01;
        sCBc[t] = on ? val : (T)0;
        // d cb / d tb-argument and d cb / d h (the soft maximum's logistic slope)
        cTB = on ? ((T)-0.5 * (((tb * lg) * (T)0.1) * eb)) / tbs : (T)0;
        cH = on ? (tb * eb) * (ex / (ex + (T)1.0)) : (T)0;
      }
      ep = e11 + e22;
      em = e11 - e22;
      const TD dsq = (TD)ep * (TD)ep + (TD)sc.rpd * ((TD)em * (TD)em + (TD)4.0 * (TD)e12c);
      // the forward kernel's FP64 island: its square root and quotient forms
      TD dCd, zd;
      if constexpr (sizeof(T) == 4 && sizeof(TD) == 8) {
        dCd = island_sqrt(dsq);
        zd = island_div((TD)0.5 * ((TD)P * (TD)sc.tens1p), dCd + (TD)sc.deltaMin);
      } else {
        dCd = sqrt(dsq);
        zd = ((TD)0.5 * ((TD)P * (TD)sc.tens1p)) / (dCd + (TD)sc.deltaMin);
      }
      zeta = (T)zd;
      dC = (T)dCd;
      sZ[t] = zeta;
      if (ada) {  // alpha at c (the forward kernel's expression order) and its derivatives
        const T mmax = fmax(mC, (T)1e-4);
        const T Y = ((zeta * (T)sc.evpcfac) / mmax) * met(5);
        const T sY = sqrt(Y);
        const T raw = sY * Mc;
        aC = fmax(raw, (T)sc.alphaMin);
        sAl[t] = aC;
        const bool act = raw > (T)sc.alphaMin && sY > (T)0;
        const T dY = act ? (Mc * (T)0.5) / sY : (T)0;  // d alpha / d Y
        cZa = dY * (((T)sc.evpcfac / mmax) * met(5));
        cMa = mC > (T)1e-4 ? -dY * (Y / mmax) : (T)0;
        // the replacement pressure, for the sigma1 relaxation's alpha derivative
        T prep = (((T)2.0 * zeta) * dC) * (T)sc.pr;
        if (!f_tens0) prep = prep / (T)sc.tens1p;
        prs = (P * (T)sc.pr1m) + prep;
        if (!f_tens0) prs = prs * (T)sc.tens1m;
      }
      sa = sA[t] + sA[t + RJ];
      sb = sB[t] + sB[t + 1];
      ww = (T)0.25 * (sa * sa + sb * sb);
      sw = sqrt(ww);
      const T cd = ((D * D) * ww > (T)sc.cDragMin2) ? D * sw : (T)sc.cDragMin;
      sCD[t] = cd * Mc;
      if (f_turn) {  // (uOcean - u) summed over j, j+1 and (vOcean - v) over i, i+1
        sAB[t] = sFU[t] + sFU[t + 1];
        sBB[t] = sFV[t] + sFV[t + RJ];
      }
    }
    __syncthreads();

    // parameter adjoints of this cell, read-modify-written at the end: issued
    // here, after the forward recompute has released its registers
    const T aP = ldg(acc.P, c), amC = ldg(acc.mC, c), amU = ldg(acc.mU, c), amV = ldg(acc.mV, c);
    const T aWX = ldg(acc.WFX, c), aWY = ldg(acc.WFY, c), aDN = ldg(acc.DN, c), aDS = ldg(acc.DS, c);
    const T uNm1 = ldg(k.uNm1, c), vNm1 = ldg(k.vNm1, c);
    // neighbours' constants used by S4
    // (the packed rows hold these at column br + 1 and, constant along i, br)
    T rAz_jp = 0, rdyU_jp = 0, rAz_ip = 0, rdxV_ip = 0;
    if (!packed) {
      rAz_jp = ldg(k.rAz, cjp);
      rdyU_jp = ldg(k.rdyU, cjp);
      rAz_ip = ldg(k.rAz, cip);
      rdxV_ip = ldg(k.rdxV, cip);
    }
    const T fmc_im = ldg(k.fC, cim) * ldg(k.massC, cim), fmc_jm = ldg(k.fC, cjm) * ldg(k.massC, cjm);

    // ---- S1: momentum update at the u- and v-point of this cell
    const T cw = (T)sc.cosWat, bU = (T)sc.betaU, bP1 = (T)sc.betaP1, rdt = (T)sc.rdt;
    T nbU = 0, dbU = 0, nbV = 0, dbV = 0, aAW = 0, aAS = 0, aMU = 0, aMV = 0;
    T bUu = bU, bVv = bU;  // beta / deltatDyn at this u- and v-point (adaptive: from alpha)
    if (ar >= 2 && br >= 2 && ar <= RI - 2 && br <= RJ - 2) {
      T bP1u = bP1, bP1v = bP1;
      if (ada) {  // veris/evp_solver.py:148-155
        bUu = ((T)0.5 * (sAl[t] + sAl[t - RJ])) * rdt;
        bVv = ((T)0.5 * (sAl[t] + sAl[t - 1])) * rdt;
        bP1u = bUu + rdt;
        bP1v = bVv + rdt;
      }
      const T cdAU = (T)0.5 * (sCD[t] + sCD[t - RJ]);
      const T cdAV = (T)0.5 * (sCD[t] + sCD[t - 1]);
      // basal drag at the u/v point (zero without it: the same sum as the forward's)
      const T cbU = bas ? (T)0.5 * (sCBc[t] + sCBc[t - RJ]) : (T)0;
      const T cbV = bas ? (T)0.5 * (sCBc[t] + sCBc[t - 1]) : (T)0;
      const T denU = mU * bP1u + AW * ((cdAU * cw) + cbU);
      const T denV = mV * bP1v + AS * ((cdAV * cw) + cbV);
      const bool gU = denU == (T)0, gV = denV == (T)0;
      T dU = gU ? (T)1 : denU, dV = gV ? (T)1 : denV;
      // lateral drag, added after the zero guard as the forward adds it
      T spU = 0, spV = 0, sdnU = 1, sdnV = 1, sdU = 0, sdV = 0, sdmU = 0, sdmV = 0;
      if (sdr) {
        spU = AW > (T)0 ? (T)0.5 * (sSPD[t] + sSPD[t - RJ]) : (T)0;
        spV = AS > (T)0 ? (T)0.5 * (sSPD[t] + sSPD[t - 1]) : (T)0;
        sdmU = sc.coast ? FuC : (((T)2 - sMU[t - RJ]) - sMU[t + RJ]) * MUc;
        sdmV = sc.coast ? FvC : (((T)2 - sMV[t - 1]) - sMV[t + 1]) * MVc;
        sdnU = spU + (T)sc.sdU0;
        sdnV = spV + (T)sc.sdU0;
        sdU = ((mU * (T)sc.sdCoeff) * sdmU) / sdnU;
        sdV = ((mV * (T)sc.sdCoeff) * sdmV) / sdnV;
        dU = dU + sdU;
        dV = dV + sdV;
      }
      const T rU = (T)1 / dU, rV = (T)1 / dV;
      const T lcu = comp ? lu : (T)0, lcv = comp ? lv : (T)0;
      nbU = lcu * MUc * rU;
      const T ddU = -lcu * un * rU, ddV = -lcv * vn * rV;  // adjoints of the denominators
      dbU = gU ? (T)0 : ddU;
      nbV = lcv * MVc * rV;
      dbV = gV ? (T)0 : ddV;
      sA[t] = nbU * met(10);  // X: adjoint of the u-point stress divergence (rrAu)
      sB[t] = nbV * met(11);  // Y: of the v-point one (rrAv)
      sQ[t] = nbU;
      sNV[t] = nbV;
      sCXU[t] = (nbU * uoc + dbU) * (AW * cw);
      sCXV[t] = (nbV * voc + dbV) * (AS * cw);
      T turnU = 0, turnV = 0;
      if (f_turn) {  // turning term of the forcing (veris/evp_solver.py:119-139)
        const T sgs = sgn(fc) * (T)sc.sinWat;
        const T lmU = mU != (T)0 ? (T)1 : mU, lmV = mV != (T)0 ? (T)1 : mV;
        const T Xt = (sCD[t] * (T)0.5) * sAB[t] + (sCD[t - RJ] * (T)0.5) * sAB[t - RJ];
        const T Yt = (sCD[t] * (T)0.5) * sBB[t] + (sCD[t - 1] * (T)0.5) * sBB[t - 1];
        turnU = -((sgs * (T)0.5) * Xt) * lmU;  // d ForcingX / d AreaW, turning part
        turnV = ((sgs * (T)0.5) * Yt) * lmV;
        sU[t] = nbU * sgn(fc) * lmU * AW;
        sV[t] = nbV * sgn(fc) * lmV * AS;
      }
      aAW = nbU * (cdAU * cw * uoc + turnU) + dbU * (cdAU * cw + cbU);  // AreaW, AreaS adjoints
      aAS = nbV * (cdAV * cw * voc + turnV) + dbV * (cdAV * cw + cbV);
      aMU = nbU * (bUu * u + rdt * uNm1) + dbU * bP1u;  // SeaIceMassU, SeaIceMassV adjoints
      aMV = nbV * (bVv * v + rdt * vNm1) + dbV * bP1v;
      if (ada) {  // beta adjoints, times the averaging weight and 1 / deltatDyn
        sLBU[t] = ((T)0.5 * rdt) * (nbU * mU * u + dbU * mU);
        sLBV[t] = ((T)0.5 * rdt) * (nbV * mV * v + dbV * mV);
      }
      if (bas) {  // the basal coefficient's share of the denominators
        sCBU[t] = dbU * AW;
        sCBV[t] = dbV * AS;
      }
      if (sdr) {  // the lateral drag's mass factor and its speed (zero where AreaW/S = 0)
        aMU = aMU + ddU * (((T)sc.sdCoeff * sdmU) / sdnU);
        aMV = aMV + ddV * (((T)sc.sdCoeff * sdmV) / sdnV);
        sLSU[t] = AW > (T)0 ? (T)-0.5 * (ddU * sdU / sdnU) : (T)0;
        sLSV[t] = AS > (T)0 ? (T)-0.5 * (ddV * sdV / sdnV) : (T)0;
      }
    }
    __syncthreads();

    // ---- S2: stresses, drag and Coriolis at c, sigma12 at z
    T Padd = 0, mCadd = 0, Db = 0, zbp = 0, epb0 = 0, emb0 = 0, dbp = 0;
    T lo1 = 0, lo2 = 0, lo12 = 0, wU = 0, wV = 0, qU = 0, qV = 0, ltb = 0, aHadd = 0, laC = 0;
    if (ar >= 3 && br >= 3 && ar <= RI - 3 && br <= RJ - 3) {
      if (bas) {  // basal coefficient adjoint at c from its four faces
        const T lcb = (T)0.5 * (sCBU[t] + sCBU[t + RJ]) + (T)0.5 * (sCBV[t] + sCBV[t + 1]);
        ltb = lcb * cTB;
        aHadd = lcb * cH;
      }
      if (sdr) {  // speed adjoint at c from its four faces, then d speed / d (u, v)
        const T lsp = (sLSU[t] + sLSU[t + RJ]) + (sLSV[t] + sLSV[t + 1]);
        const T lq = spd > (T)0 ? lsp * (T)0.125 / spd : (T)0;  // speed = sqrt(q) / 2
        qU = (T)2 * lq * su;
        qV = (T)2 * lq * sv;
      }
      const T Xc = sA[t], Xip = sA[t + RJ], Xjm = sA[t - 1];
      const T Yc = sB[t], Yjp = sB[t + 1], Yim = sB[t - RJ];
      T cdb = (T)0.5 * (sCXU[t] + sCXU[t + RJ]) + (T)0.5 * (sCXV[t] + sCXV[t + 1]);
      if (f_turn) {  // the drag of this c-point in the turning terms of its four faces
        // (sAB/sBB of this cell are read before this thread overwrites them)
        const T kus = sU[t] + sU[t + RJ], kvs = sV[t] + sV[t + 1];
        cdb = cdb + (T)0.25 * (T)sc.sinWat * (sBB[t] * kvs - sAB[t] * kus);
        wU = sCD[t] * kus;  // and the ice velocities in them, weighted by the drag
        wV = sCD[t] * kvs;
      }
      const T fvb = (T)0.5 * (sQ[t] + sQ[t + RJ]);
      const T fub = -(T)0.5 * (sNV[t] + sNV[t + 1]);
      // neighbour velocities (sU/sV hold the turning factors when sinWat != 0)
      const T uip = f_turn ? ldg(xu, cip) : sU[t + RJ];
      const T vjp = f_turn ? ldg(xv, cjp) : sV[t + 1];
      mCadd = fvb * fc * (T)0.5 * (v + vjp) + fub * fc * (T)0.5 * (u + uip);
      sFV[t] = fvb;
      sFU[t] = fub;
      const T cbe = cdb * Mc;
      const bool drag = (D * D) * ww > (T)sc.cDragMin2;
      const T wb = (drag && sw > (T)0) ? cbe * D * (T)0.5 / sw : (T)0;
      Db = drag ? cbe * sw * (T)sc.rhoSea : (T)0;
      sAB[t] = wb * (T)0.5 * sa;
      sBB[t] = wb * (T)0.5 * sb;
      const T s11b = met(6) * (Xc - Xip);  // dyV
      const T s22b = met(7) * (Yc - Yjp);  // dxU
      const T s1t = l1 + (T)0.5 * (s11b + s22b);
      const T s2t = l2 + (T)0.5 * (s11b - s22b);
      // relaxation weights: fixed, or this substep's alpha at c (adaptive)
      const T den1c = ada ? (T)1.0 / aC : (T)sc.den1, am1c = ada ? aC - (T)1 : (T)sc.alphaM1;
      const T H1 = s1t * den1c * Mc, H2 = s2t * den1c * Mc;
      lo1 = H1 * am1c;
      lo2 = H2 * am1c;
      const T pb = -H1;
      const T pfac = (T)sc.tens1m * (T)2 * (T)sc.pr / (T)sc.tens1p;
      Padd = pb * (T)sc.tens1m * (T)sc.pr1m;
      zbp = H1 * (T)2 * ep + H2 * (T)0.5 * em + pb * pfac * dC;
      epb0 = H1 * (T)2 * zeta;
      emb0 = H2 * (T)0.5 * zeta;
      dbp = pb * pfac * zeta;
      const T s12t = l12 + met(8) * (Xjm - Xc) + met(9) * (Yim - Yc);  // dxV, dyU
      T den2z = (T)sc.den2, am1z = (T)sc.alphaM1, aZ = 0;
      if (ada) {  // alpha at z (veris/evp_solver.py:97-100)
        aZ = (T)0.5 * (sAl[t] + sAl[t - 1]);
        aZ = (T)0.5 * (aZ + (T)0.5 * (sAl[t - RJ] + sAl[t - RJ - 1]));
        den2z = (T)1.0 / aZ;
        am1z = aZ - (T)1;
      }
      const T G = s12t * den2z;
      lo12 = G * am1z;
      const T nz = (Mc + sM[t - RJ]) + (sM[t - 1] + sM[t - RJ - 1]);
      // 1 / nz for the 0/1 mask sums 1..4 (0 -> 0), without a division
      const T nrm = nz == (T)4 ? (T)0.25
                    : nz == (T)3 ? (T)(1.0 / 3.0)
                    : nz == (T)2 ? (T)0.5
                    : nz == (T)1 ? (T)1 : (nz > (T)0 ? (T)1 / nz : (T)0);
      const T zz = nrm * ((sZ[t] + sZ[t - RJ]) + (sZ[t - 1] + sZ[t - RJ - 1]));
      sCD[t] = G * (T)0.5 * e12r * nrm;  // zeta adjoint share of z = this cell, weighted
      sEP[t] = G * (T)0.5 * zz;          // e12 adjoint from sigma12
      if (ada) {
        // d sigma_new / d alpha = (sigma_old - forcing) / alpha^2 for sigma1, sigma2 at c
        // and sigma12 at z; beta of the four faces of this c-point
        const T divg = (((T)2 * zeta) * ep) - prs;
        const T tens = ((T)2 * zeta) * em;
        laC = (H1 * (so1 - divg) + H2 * (so2 - tens * (T)0.25)) * den1c +
              ((sLBU[t] + sLBU[t + RJ]) + (sLBV[t] + sLBV[t + 1]));
        const T shear = ((T)2 * zz) * e12r;
        sLAZ[t] = G * (so12 - shear * (T)0.25) * den2z;
      }
    }
    __syncthreads();

    // ---- S3: zeta and Delta at c
    if (ar >= 3 && br >= 3 && ar <= RI - 4 && br <= RJ - 4) {
      T zb = zbp + ((sCD[t] + sCD[t + RJ]) + (sCD[t + 1] + sCD[t + RJ + 1]));
      if (ada) {  // alpha at c from its four z-points, then to zeta and the mass
        const T la = laC + (T)0.25 * ((sLAZ[t] + sLAZ[t + 1]) + (sLAZ[t + RJ] + sLAZ[t + RJ + 1]));
        zb = zb + la * cZa;
        mCadd = mCadd + la * cMa;
      }
      const T rreg = (T)1 / (dC + (T)sc.deltaMin);
      Padd += zb * (T)0.5 * (T)sc.tens1p * rreg;
      const T Dbar = dbp - zb * zeta * rreg;
      const T dsqb = dC > (T)0 ? Dbar * (T)0.5 / dC : (T)0;
      const T epb = epb0 + dsqb * (T)2 * ep;
      const T emb = emb0 + dsqb * (T)sc.rpd * (T)2 * em;
      sA[t] = (epb + emb) * met(3) * mInC;  // e11 adjoint (times its stencil factor)
      sB[t] = (epb - emb) * met(4) * mInC;  // e22 adjoint
      sQ[t] = dsqb * (T)sc.rpd * met(5);    // adjoint of rAz e12^2 at the four z-corners
      if (f_kterms) {  // e22 += uave k1C, e11 += vave k2C (veris/dynamics_routines.py:151-158)
        sNV[t] = ((epb - emb) * mInC) * ldg(k.k1C, c) * (T)0.5;
        sCXU[t] = ((epb + emb) * mInC) * ldg(k.k2C, c) * (T)0.5;
      }
      if (f_turn) {
        sU[t] = wU;
        sV[t] = wV;
      }
      if (sdr) {
        sZ[t] = qU;
        sCXV[t] = qV;
      }
      if (bas) sLTB[t] = ltb;
    }
    __syncthreads();

    // ---- S4: input-velocity adjoints; results of the tile's own cells
    if (!own) return;
    auto eadj = [&](int z, T rz) {
      const T qs = (sQ[z] + sQ[z - RJ]) + (sQ[z - 1] + sQ[z - RJ - 1]);
      return sEP[z] + qs * (T)2 * rz * sE12[z];
    };
    auto mzf = [&](int z) { return (sM[z] * sM[z - RJ]) * (sM[z - 1] * sM[z - RJ - 1]); };
    const T rdyU = met(0), rdxV = met(1), rAz = met(2);
    if (packed) {
      rAz_jp = sRow[2 * RJ + br + 1];
      rdyU_jp = sRow[br + 1];
      rAz_ip = rAz;
      rdxV_ip = rdxV;
    }
    const T ec = eadj(t, rAz), ejp = eadj(t + 1, rAz_jp), eip = eadj(t + RJ, rAz_ip);
    const T mzc = mzf(t), mzjp = mzf(t + 1), mzip = mzf(t + RJ);
    // no-slip boundary term of e12 (absent under free slip)
    const T huc = sdr ? (T)0 : MUc - sMU[t - 1], hujp = sdr ? (T)0 : sMU[t + 1] - MUc;
    const T hvc = sdr ? (T)0 : MVc - sMV[t - RJ], hvip = sdr ? (T)0 : sMV[t + RJ] - MVc;
    const T A1 = (T)0.5 * mzc * rdyU + huc * rdyU;
    const T A2jp = -(T)0.5 * mzjp * rdyU_jp + hujp * rdyU_jp;
    const T A3 = (T)0.5 * mzc * rdxV + hvc * rdxV;
    const T A4ip = -(T)0.5 * mzip * rdxV_ip + hvip * rdxV_ip;
    const T fmc = fc * mC;
    T lu_out = nbU * bUu * mU + (sAB[t] + sAB[t - RJ]) * mInUc + (sA[t - RJ] - sA[t]) +
               ec * A1 + ejp * A2jp + (T)0.5 * fmc * sFU[t] + (T)0.5 * fmc_im * sFU[t - RJ];
    T lv_out = nbV * bVv * mV + (sBB[t] + sBB[t - 1]) * mInVc + (sB[t - 1] - sB[t]) +
               ec * A3 + eip * A4ip + (T)0.5 * fmc * sFV[t] + (T)0.5 * fmc_jm * sFV[t - 1];
    if (f_kterms) {  // the metric terms of e11, e22 (c) and e12 (z)
      const T k1Z = ldg(k.k1Z, c), k2Z = ldg(k.k2Z, c);
      const T k2Z_jp = ldg(k.k2Z, cjp), k1Z_ip = ldg(k.k1Z, cip);
      lu_out = lu_out + (sNV[t] + sNV[t - RJ]) - (T)0.25 * (ec * mzc * k2Z + ejp * mzjp * k2Z_jp);
      lv_out = lv_out + (sCXU[t] + sCXU[t - 1]) - (T)0.25 * (ec * mzc * k1Z + eip * mzip * k1Z_ip);
    }
    if (f_turn) {  // the ice velocity in the turning terms of the four neighbouring faces
      lu_out = lu_out + (T)0.25 * (T)sc.sinWat * (sU[t] + sU[t - 1]);
      lv_out = lv_out - (T)0.25 * (T)sc.sinWat * (sV[t] + sV[t - RJ]);
    }
    if (sdr) {  // the ice speed at the two c-points of this u (v) point
      lu_out = lu_out + (sZ[t] + sZ[t - RJ]);
      lv_out = lv_out + (sCXV[t] + sCXV[t - 1]);
    }
    if (SO && sc.sobc) {
      // the second-order boundary shear of the z-points whose stencil holds this
      // u (v): uavz at z = c, c+1 (v: c, c+RJ), u two cells down at z = c+2RJ
      // and one up at z = c-RJ (v: c+2, c-1), as the reference shifts them
      auto hu2f = [&](int z) {
        return ((sMU[z] - sMU[z - RJ]) / (T)3.0) * (sMU[z - 2 * RJ] * sMU[z - RJ] + sMU[z + RJ] * sMU[z]);
      };
      auto hv2f = [&](int z) {
        return ((sMV[z] - sMV[z - 1]) / (T)3.0) * (sMV[z - 2] * sMV[z - 1] + sMV[z + 1] * sMV[z]);
      };
      T rAz_i2 = rAz, rAz_im = rAz, rdyU_i2 = rdyU, rdyU_im = rdyU;  // along i: rows hold them
      T rAz_j2, rAz_jm, rdxV_j2, rdxV_jm;
      if (packed) {
        rAz_j2 = sRow[2 * RJ + br + 2];
        rAz_jm = sRow[2 * RJ + br - 1];
        rdxV_j2 = sRow[RJ + br + 2];
        rdxV_jm = sRow[RJ + br - 1];
      } else {
        const unsigned ci2 = mo + wrap(gi + 2, g.NX) * g.NY + gj, cj2 = mo + gi * g.NY + wrap(gj + 2, g.NY);
        rAz_i2 = ldg(k.rAz, ci2);
        rAz_im = ldg(k.rAz, cim);
        rdyU_i2 = ldg(k.rdyU, ci2);
        rdyU_im = ldg(k.rdyU, cim);
        rAz_j2 = ldg(k.rAz, cj2);
        rAz_jm = ldg(k.rAz, cjm);
        rdxV_j2 = ldg(k.rdxV, cj2);
        rdxV_jm = ldg(k.rdxV, cjm);
      }
      const T e_i2 = eadj(t + 2 * RJ, rAz_i2), e_im = eadj(t - RJ, rAz_im);
      const T e_j2 = eadj(t + 2, rAz_j2), e_jm = eadj(t - 1, rAz_jm);
      lu_out = lu_out + (T)1.5 * (ec * rdyU * hu2f(t) + ejp * rdyU_jp * hu2f(t + 1)) -
               (T)0.5 * (e_i2 * rdyU_i2 * hu2f(t + 2 * RJ) * sMU[t + RJ] +
                         e_im * rdyU_im * hu2f(t - RJ) * sMU[t - RJ]);
      lv_out = lv_out + (T)1.5 * (ec * rdxV * hv2f(t) + eip * rdxV_ip * hv2f(t + RJ)) -
               (T)0.5 * (e_j2 * rdxV_j2 * hv2f(t + 2) * sMV[t + 1] +
                         e_jm * rdxV_jm * hv2f(t - 1) * sMV[t - 1]);
    }
    if (bas) {  // the basal coefficient's velocity argument at the two c-points
      lu_out = lu_out + (((T)0.5 * (u * mInUc)) * mInUc) * (sLTB[t] + sLTB[t - RJ]);
      lv_out = lv_out + (((T)0.5 * (v * mInVc)) * mInVc) * (sLTB[t] + sLTB[t - 1]);
      stg(acc.h, c, ldg(acc.h, c) + aHadd);
    }
    stg(L.out[0], c, lu_out);
    stg(L.out[1], c, lv_out);
    stg(L.out[2], c, lo1);
    stg(L.out[3], c, lo2);
    stg(L.out[4], c, lo12);
    stg(acc.P, c, aP + Padd);
    stg(acc.mC, c, amC + mCadd);
    if (fc < (T)0) stg(acc.DS, c, aDS + Db);
    else stg(acc.DN, c, aDN + Db);
    if (comp) {  // nbU = nbV = 0 elsewhere
      stg(acc.mU, c, amU + aMU);
      stg(acc.WFX, c, aWX + nbU);
      stg(acc.mV, c, amV + aMV);
      stg(acc.WFY, c, aWY + nbV);
      if (acc.u0) {  // optional outputs (uniform branch)
        stg(acc.u0, c, ldg(acc.u0, c) + nbU * mU * rdt);
        stg(acc.v0, c, ldg(acc.v0, c) + nbV * mV * rdt);
      }
      if (acc.AW) {
        stg(acc.AW, c, ldg(acc.AW, c) + aAW);
        stg(acc.AS, c, ldg(acc.AS, c) + aAS);
      }
    }
  };
  if (blk->packed != 0) body(std::true_type{});
  else body(std::false_type{});
}

struct AdjTileCfg {
  int ri, rj;
};
// 512- and 768-thread shapes: the FP64 kernel needs ~156 VGPRs, more than a
// 1024-thread workgroup can give each lane.
constexpr AdjTileCfg kAdjTiles[] = {{16, 32}, {24, 32}, {12, 64}, {32, 32}, {20, 32}};
constexpr int kAdjNumTiles = sizeof(kAdjTiles) / sizeof(kAdjTiles[0]);

template <typename T, int RI, int RJ, int MODE>
void adj_launch_shape(const AdjArgs<T>* a, const EvpArgBlock<T>* b, LamSet<T> L, const Geom& g,
                      int n_members, hipStream_t st) {
  constexpr int R = MODE == 2 ? 5 : 4, TI = RI - 2 * R, TJ = RJ - 2 * R;
  const int tiles_j = (g.NY + TJ - 1) / TJ;
  const int n_tiles = ((g.NX + TI - 1) / TI) * tiles_j;
  const int per_xcd = (n_tiles + 7) / 8;
  hipLaunchKernelGGL((adj_substep_kernel<T, double, RI, RJ, MODE>), dim3(per_xcd * 8, n_members),
                     dim3(RI * RJ), 0, st, a, b, L, n_tiles, tiles_j, per_xcd);
}

template <typename T>
void adj_launch(int cfg, const AdjArgs<T>* a, const EvpArgBlock<T>* b, LamSet<T> L, const Geom& g,
                int n_members, int mode, hipStream_t st) {
  if constexpr (sizeof(T) == 8) {
    if (mode >= 1) {
      // the extra-branch FP64 kernels: their LDS does not fit a 768-thread
      // workgroup, so they exist as 16 x 32 only
      if (mode == 2) adj_launch_shape<T, 16, 32, 2>(a, b, L, g, n_members, st);
      else adj_launch_shape<T, 16, 32, 1>(a, b, L, g, n_members, st);
      return;
    }
  }
  if (mode == 2) {  // FP32 with the second-order term: 24 x 32 only
    if constexpr (sizeof(T) == 4) adj_launch_shape<T, 24, 32, 2>(a, b, L, g, n_members, st);
    return;
  }
  if (mode == 1) {
    if constexpr (sizeof(T) == 4) {
      switch (cfg) {
        case 1:
        case 4: adj_launch_shape<T, 24, 32, 1>(a, b, L, g, n_members, st); break;
        case 2: adj_launch_shape<T, 12, 64, 1>(a, b, L, g, n_members, st); break;
        case 3: adj_launch_shape<T, 32, 32, 1>(a, b, L, g, n_members, st); break;
        default: adj_launch_shape<T, 16, 32, 1>(a, b, L, g, n_members, st); break;
      }
    }
    return;
  }
  auto by_shape = [&](auto md) {
    constexpr int M = decltype(md)::value;
    switch (cfg) {
      case 1: adj_launch_shape<T, 24, 32, M>(a, b, L, g, n_members, st); break;
      case 2: adj_launch_shape<T, 12, 64, M>(a, b, L, g, n_members, st); break;
      case 4: adj_launch_shape<T, 20, 32, M>(a, b, L, g, n_members, st); break;
      case 3:
        if constexpr (sizeof(T) == 4) adj_launch_shape<T, 32, 32, M>(a, b, L, g, n_members, st);
        else adj_launch_shape<T, 24, 32, M>(a, b, L, g, n_members, st);
        break;
      default: adj_launch_shape<T, 16, 32, M>(a, b, L, g, n_members, st); break;
    }
  };
  if (mode == 3) by_shape(std::integral_constant<int, 3>{});
  else by_shape(std::integral_constant<int, 0>{});
}

// ---- pre-EVP adjoint: strength, mass, wind forcing
template <typename T>
struct PreAdjArgs {
  // forward inputs
  const T *h, *A, *M, *MU, *MV, *pstarF, *u, *v, *uW, *vW, *fC, *rdxC, *rdyC, *ssh, *sp, *AW, *AS;
  // adjoints of the EVP inputs
  const T *aP, *amC, *amU, *amV, *aWFX, *aWFY;
  const T* ahb;  // thickness adjoint of the basal drag (NULL: none)
  // outputs
  T *apstar, *ah, *ahs, *aCaN, *aCaS;
  T pstar, cStar, rhoI, rhoS, rhoAir, wmin, wmin2, grav, rrho, sinWin, cosWin;
  const T* load;  // SeaIceLoad in the surface tilt (useRealFreshWaterFlux), else NULL
  T loadFac;      // gravity * seaIceLoadFac
  int relWind;
};

// d tau_c / d Ca at c-point (i, j) (veris/dynsolver.py:7-47)
template <typename T>
__device__ inline void tau_c_dca(const Geom& g, const PreAdjArgs<T>& a, size_t mo, int i, int j,
                                 T* dtx, T* dty) {
  const int ip = wp(i, g.NX), jp = wp(j, g.NY);
  const size_t c = mo + AT(i, j);
  T urel = a.uW[c], vrel = a.vW[c];
  if (a.relWind) {
    urel = urel - (T)0.5 * (a.u[c] + a.u[mo + AT(ip, j)]);
    vrel = vrel - (T)0.5 * (a.v[c] + a.v[mo + AT(i, jp)]);
  }
  const T wsq = urel * urel + vrel * vrel;
  const T ws = wsq < a.wmin2 ? a.wmin : sqrt(wsq);
  const T sg = sgn(a.fC[c]);
  *dtx = (a.rhoAir * ws) * (a.cosWin * urel - (sg * a.sinWin) * vrel);
  *dty = (a.rhoAir * ws) * (a.cosWin * vrel + (sg * a.sinWin) * urel);
}

template <typename T>
__global__ __launch_bounds__(BLK) void pre_evp_adjoint_kernel(Geom g, PreAdjArgs<T> a) {
  CELL_SETUP_G
  // equal members packed back to back: member blockIdx.y starts at element mo
  const size_t mo = (size_t)blockIdx.y * n_;
  // strength: P = P* h exp(-C (1 - A)) M
  const T E = exp((-a.cStar) * ((T)1 - a.A[mo + c]));
  const T ps = a.pstarF ? a.pstarF[mo + c] : a.pstar;
  a.apstar[mo + c] = a.aP[mo + c] * a.h[mo + c] * E * a.M[mo + c];
  T hb = a.aP[mo + c] * ps * E * a.M[mo + c];
  // wind forcing: the tilt term -massU rdxC (phi - phi[i-1]) adds to the mass adjoints
  // (the load of useRealFreshWaterFlux is a state input here: Growth wrote it)
  auto phi_at = [&](int ii, int jj) {
    const T p = a.load ? a.sp[mo + AT(ii, jj)] + a.load[mo + AT(ii, jj)] * a.loadFac
                       : a.sp[mo + AT(ii, jj)];
    return a.grav * a.ssh[mo + AT(ii, jj)] + p * a.rrho;
  };
  const T phi = phi_at(i, j);
  const T amU_c = a.amU[mo + c] - a.aWFX[mo + c] * a.rdxC[mo + c] * (phi - phi_at(im, j));
  const T amU_ip = a.amU[mo + AT(ip, j)] - a.aWFX[mo + AT(ip, j)] * a.rdxC[mo + AT(ip, j)] * (phi_at(ip, j) - phi);
  const T amV_c = a.amV[mo + c] - a.aWFY[mo + c] * a.rdyC[mo + c] * (phi - phi_at(i, jm));
  const T amV_jp = a.amV[mo + AT(i, jp)] - a.aWFY[mo + AT(i, jp)] * a.rdyC[mo + AT(i, jp)] * (phi_at(i, jp) - phi);
  // mass: massC = rhoI h + rhoS hs, massU = (massC + massC[i-1])/2, massV likewise in j
  const T amC = a.amC[mo + c] + (T)0.5 * (amU_c + amU_ip) + (T)0.5 * (amV_c + amV_jp);
  a.ah[mo + c] = (hb + a.rhoI * amC) + (a.ahb ? a.ahb[mo + c] : (T)0);
  a.ahs[mo + c] = a.rhoS * amC;
  // air drag: WindForcingX = 0.5 (tau_c + tau_c[i-1]) MU AreaW + tilt, tau_c linear in Ca
  T dtx, dty;
  tau_c_dca(g, a, mo, i, j, &dtx, &dty);
  const T gx = (T)0.5 * (a.aWFX[mo + c] * a.AW[mo + c] * a.MU[mo + c] +
                         a.aWFX[mo + AT(ip, j)] * a.AW[mo + AT(ip, j)] * a.MU[mo + AT(ip, j)]);
  const T gy = (T)0.5 * (a.aWFY[mo + c] * a.AS[mo + c] * a.MV[mo + c] +
                         a.aWFY[mo + AT(i, jp)] * a.AS[mo + AT(i, jp)] * a.MV[mo + AT(i, jp)]);
  const T contrib = gx * dtx + gy * dty;
  a.aCaN[mo + c] = a.fC[mo + c] < (T)0 ? (T)0 : contrib;
  a.aCaS[mo + c] = a.fC[mo + c] < (T)0 ? contrib : (T)0;
}

// ---- host side
struct AdjLayout {
  size_t fb, blk, total;
  int nck, ce, n;
  size_t off_args, off_fblk, off_rblk, off_r2blk, off_ckpt, off_pp, off_seg, off_lam, off_pblk,
      off_pack, seg_stride;
};

// Workspace of a subcycle of n substeps with a checkpoint every ce, for nm
// equal members packed back to back (1 = one domain).
template <typename T>
AdjLayout adj_layout(const Geom& g, int n, int ce, int nm = 1) {
  AdjLayout L;
  L.fb = align256(field_elems(g) * (size_t)nm * sizeof(T));
  L.blk = align256(sizeof(EvpArgBlock<T>));
  L.ce = ce;
  L.n = n;
  L.nck = (n + ce - 1) / ce;
  L.off_args = 0;
  L.off_fblk = align256(sizeof(AdjArgs<T>));
  L.off_rblk = L.off_fblk + (size_t)n * L.blk;
  L.off_r2blk = L.off_rblk + (size_t)n * L.blk;  // two-substep recompute blocks
  L.off_ckpt = L.off_r2blk + (size_t)n * L.blk;
  // checkpoint slots 1 .. nck-1, slot nck = copy of the final state (slot 0 unused)
  L.off_pp = L.off_ckpt + (size_t)5 * (L.nck + 1) * L.fb;
  L.off_seg = L.off_pp + (size_t)10 * L.fb;
  L.seg_stride = (size_t)5 * (std::min(ce, n) + 1) * L.fb;  // one set of segment slots
  // three sets of segment slots (slot 0 of each unused): two alternate so that
  // the recompute of one segment overlaps the reverse sweep of the next, the
  // third keeps the last segment's states from the forward; a subcycle of one
  // segment (ckpt_every >= n_substeps: every state kept, nothing recomputed)
  // needs only that one
  L.off_lam = L.off_seg + (size_t)(L.nck == 1 ? 1 : 3) * L.seg_stride;
  // argument block of the pack kernel, then the packed step constants of the
  // forward and recompute launches
  L.off_pblk = L.off_lam + (size_t)10 * L.fb;
  L.off_pack = L.off_pblk + L.blk;
  L.total = L.off_pack + evp_pack_bytes<T>(field_elems(g) * (size_t)nm, (size_t)g.NY * nm);
  return L;
}


template <typename T>
T* wsf(void* ws, size_t off, size_t fb, int q) {
  return (T*)((char*)ws + off + (size_t)q * fb);
}

// Tile shape of the reverse kernel (index into kAdjTiles; environment
// VERIS_ADJ_TILE overrides the choice): 20 x 32, the fastest measured where
// latency bounds the kernel (1 degree, 60k cells, either precision: 217
// workgroups where 24 x 32 gives 161, 3% faster); in FP32 from 200k cells
// on, 32 x 32, whose smaller ring of recomputed cells pays once the launch
// fills the GPU (8 members of 240 x 240: 6% faster).  The extra-branch
// instantiations run 16 x 32 (FP64) or 24 x 32 (FP32) for either small shape.
template <typename T>
int adj_tile_cfg(size_t cells) {
  static int env = -2;
  if (env == -2) {
    const char* e = getenv("VERIS_ADJ_TILE");
    const int v = e ? atoi(e) : -1;
    env = (v >= 0 && v < kAdjNumTiles) ? v : -1;
  }
  if (env >= 0) return env;
  return (sizeof(T) == 4 && cells >= 200000) ? 3 : 4;
}

// Whether the recompute of a segment runs on a second stream beside the
// reverse sweep of the next (environment VERIS_ADJ_OVERLAP=0 serialises it).
bool adj_overlap() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("VERIS_ADJ_OVERLAP");
    v = (e && atoi(e) == 0) ? 0 : 1;
  }
  return v == 1;
}

// Whether the checkpointed forward runs four substeps per launch before the
// last segment where it runs two (environment VERIS_ADJ_QUAD=0 keeps two).
bool adj_quad() {
  const char* e = getenv("VERIS_ADJ_QUAD");
  return !(e && atoi(e) == 0);
}

// A second stream and two events per device, for the fork/join of the
// reverse sweep (created once, default priority, kept for the process).
struct SideStream {
  hipStream_t stream;
  hipEvent_t fork, join;
};

int side_stream(SideStream* out) {
  static std::mutex mu;
  static std::map<int, SideStream> per_dev;
  int dev = 0;
  VH_TRY_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto it = per_dev.find(dev);
  if (it == per_dev.end()) {
    SideStream ss{};
    VH_TRY_HIP(hipStreamCreateWithFlags(&ss.stream, hipStreamNonBlocking));
    VH_TRY_HIP(hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming));
    VH_TRY_HIP(hipEventCreateWithFlags(&ss.join, hipEventDisableTiming));
    it = per_dev.emplace(dev, ss).first;
  }
  *out = it->second;
  return VH_OK;
}

// Single-substep shape of the forward and recompute launches: the requested
// one when it has k = 1, else the default.
int pick_adj_tile(const Geom& g, int req) {
  if (req > 0 && kEvpTiles[req].k == 1) return req;
  return (long long)g.NX * g.NY < 400000 ? 5 : 3;
}

int check_adj(const vh_domain* d, const vh_settings* s, const vh_fields* f, int n, int ce) {
  VH_REQUIRE(d && s && f, VH_EINVAL, "null domain, settings or fields");
  VH_REQUIRE(n >= 1 && ce >= 1, VH_EINVAL, "n_substeps=%d ckpt_every=%d", n, ce);
  VH_REQUIRE(n <= (1 << 20), VH_EINVAL, "n_substeps=%d too large", n);
  VH_REQUIRE(d->tile >= 0 && d->tile < kEvpNumTiles && (d->tile == 0 || kEvpTiles[d->tile].k <= 2),
             VH_EINVAL, "adjoint: tile %d (needs a shape of one or two substeps, or 0)", d->tile);
  VH_REQUIRE(s->useEVP != 0.0, VH_EUNSUPPORTED, "adjoint: the EVP solver must be on");

  VH_REQUIRE(s->basalDragK2 == 0.0 || (f->R_low && f->Area && f->hIceMean), VH_EINVAL,
             "adjoint: basal drag needs R_low, Area, hIceMean");
  const bool kt = f->k1AtC || f->k2AtC || f->k1AtZ || f->k2AtZ;
  VH_REQUIRE(!kt || (f->k1AtC && f->k2AtC && f->k1AtZ && f->k2AtZ), VH_EINVAL,
             "adjoint: k1AtC/k2AtC/k1AtZ/k2AtZ must be all set or all NULL");
  const void* req[kEvpMaxRequired];
  const int nreq = evp_required_fields(f, req);
  for (int q = 0; q < nreq; ++q)
    VH_REQUIRE(req[q] != nullptr, VH_EINVAL, "adjoint: missing input field");
  return VH_OK;
}

inline int kterms_of(const vh_fields* f) {
  return (f->k1AtC || f->k2AtC || f->k1AtZ || f->k2AtZ) ? 1 : 0;
}

template <typename T>
EvpArgBlock<T> proto_block(const Geom& g, const vh_settings* s, const vh_fields* f) {
  EvpArgBlock<T> ab{};
  fill_evp_scalars(s, kterms_of(f), s->basalDragK2 != 0.0 ? 1 : 0, &ab.sc);
  fill_evp_const(f, &ab.c);
  ab.g = g;
  ab.nlaunch = 1;
  ab.members = nullptr;
  ab.mstride = (long long)field_elems(g);  // member blockIdx.y of equal packed members
  return ab;
}

// Packs the step constants into the workspace (evp_pack_kernel) and returns
// the prototype block of the forward launches, pointing at them; the blocks
// take the pack kernel's outcome from the returned device block.  With
// `pack` false the packed constants and their outcome are the ones the
// checkpointed forward left in the same workspace (the reverse sweep runs on
// the forward's fields).
template <typename T>
EvpArgBlock<T> packed_proto(const Geom& g, const vh_settings* s, const vh_fields* f,
                            const AdjLayout& L, void* ws, const EvpArgBlock<T>** pack_status,
                            int nm, hipStream_t st, bool pack = true) {
  EvpArgBlock<T> proto = proto_block<T>(g, s, f);
  evp_pack_attach(&proto, (char*)ws + L.off_pack, field_elems(g) * (size_t)nm, g.NY * nm);
  EvpArgBlock<T>* pblk = (EvpArgBlock<T>*)((char*)ws + L.off_pblk);
  if (pack) {
    hipLaunchKernelGGL(evp_write_args_kernel<T>, dim3(1), dim3(64), 0, st, pblk, proto);
    evp_pack_launch(pblk, (long long)field_elems(g), nm, st);
  }
  *pack_status = pblk;
  return proto;
}

template <typename T>
BlockPlan<T> block_plan(const AdjLayout& L, const vh_fields* f, void* const outs[5], void* ws) {
  BlockPlan<T> p{};
  const void* in0[5] = {f->uIce, f->vIce, f->sigma1, f->sigma2, f->sigma12};
  for (int q = 0; q < 5; ++q) {
    p.in0[q] = (const T*)in0[q];
    p.outs[q] = outs ? (T*)outs[q] : nullptr;
  }
  p.ws = (char*)ws;
  p.fb = L.fb;
  p.off_ckpt = L.off_ckpt;
  p.nck = L.nck;
  p.off_pp = L.off_pp;
  p.off_seg = L.off_seg;
  p.seg_stride = L.seg_stride;
  p.n = L.n;
  p.ce = L.ce;
  p.kstep = 1;
  return p;
}

// Shape of the two-substep forward and recompute launches of the adjoint
// (a k = 2 shape), or 0 when they run one substep per launch.  Two substeps
// per launch need the default settings (the default reverse kernel reads only
// the velocities of the states, which two-substep launches then give bit for
// bit) and an even checkpoint interval and subcycle; they run when the caller
// names a k = 2 shape in vh_domain.tile (stating that vh_evp_overlaps_periodic
// holds for the inputs) or, with tile 0, on a domain without a filled overlap
// (20 x 32 x 2 below 400k cells, where one launch is latency bound, 32 x 32 x 2
// above).  Environment VERIS_ADJ_PAIRS=0 (read at every call) turns them off.
// Returns -1 for a k = 2 request that the settings or sizes rule out.
int adj_pair_cfg(const Geom& g, int tile, bool plain, int n, int ce, size_t cells) {
  const char* e = getenv("VERIS_ADJ_PAIRS");
  const bool requested = tile > 0 && kEvpTiles[tile].k == 2;
  if (e && atoi(e) == 0) return 0;
  if (!plain || n % 2 != 0 || ce % 2 != 0) return requested ? -1 : 0;
  if (requested) return tile;
  if (tile != 0 || g.fill_x || g.fill_y) return 0;
  return cells < 400000 ? 18 : 9;
}

void tiles_of(int cfg, const Geom& g, int* n_tiles, int* tiles_j) {
  const int TI = evp_tile_extent_i(cfg), TJ = evp_tile_extent_j(cfg);
  *tiles_j = (g.NY + TJ - 1) / TJ;
  *n_tiles = ((g.NX + TI - 1) / TI) * *tiles_j;
}

template <typename T>
int run_recorded(const char* kind, const Geom& g, const AdjLayout& L, int cfg, int nm, void* ws,
                 const std::function<void(hipStream_t)>& record, hipStream_t st) {
  if (graphs_enabled()) {
    char key[192];
    snprintf(key, sizeof(key), "%s/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%p", kind, (int)sizeof(T), cfg,
             L.n, L.ce, nm, g.NX, g.NY, g.nx, g.fill_x, g.fill_y, ws);
    return graph_run(key, record, st);
  }
  record(st);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int forward_ckpt(const Geom& g, const vh_domain* d, const vh_settings* s, const vh_fields* f,
                 int n, int ce, int nm, void* const outs[5], void* ws, size_t wsb, hipStream_t st) {
  const AdjLayout L = adj_layout<T>(g, n, ce, nm);
  VH_REQUIRE(ws && wsb >= L.total, VH_EWORKSPACE, "adjoint workspace: need %zu bytes, got %zu",
             L.total, wsb);
  const void* req[kEvpMaxRequired];
  const int nreq = evp_required_fields(f, req);
  for (int q = 0; q < 5; ++q) {
    VH_REQUIRE(outs[q] != nullptr, VH_EINVAL, "null output");
    for (int r = 0; r < nreq; ++r)
      VH_REQUIRE(outs[q] != req[r], VH_EINVAL, "output aliases an input field");
    for (int r = 0; r < q; ++r) VH_REQUIRE(outs[q] != outs[r], VH_EINVAL, "two outputs alias");
  }
  const int cfg = pick_adj_tile(g, d->tile);
  int n_tiles, tiles_j;
  tiles_of(cfg, g, &n_tiles, &tiles_j);
  EvpArgBlock<T>* fblk = (EvpArgBlock<T>*)((char*)ws + L.off_fblk);
  AdjArgs<T> unused{};
  const EvpArgBlock<T>* pack_status = nullptr;
  const EvpArgBlock<T> proto = packed_proto<T>(g, s, f, L, ws, &pack_status, nm, st);
  const bool so = s->noSlip != 0.0 && s->secondOrderBC != 0.0;
  const bool plain = evp_plain_kernels_enabled() && evp_plain_settings(proto.sc);
  // two substeps per launch when the adjoint allows it (adj_pair_cfg)
  const int pcfg = adj_pair_cfg(g, d->tile, plain, n, ce, field_elems(g) * (size_t)nm);
  VH_REQUIRE(pcfg >= 0, VH_EUNSUPPORTED,
             "adjoint: tile %d runs two substeps per launch, which needs the default settings "
             "and even n_substeps (%d) and ckpt_every (%d)", d->tile, n, ce);
  BlockPlan<T> plan = block_plan<T>(L, f, outs, ws);
  if (pcfg) plan.kstep = 2;
  // four substeps per launch before the last segment, where the pairs run
  // and the domain is small enough that the four-substep shape is the faster
  const int qcfg = (pcfg && L.nck >= 2 && ce >= 4 && adj_quad() &&
                    field_elems(g) * (size_t)nm < 400000)
                       ? evp_quad_cfg()
                       : 0;
  const int q4 = ce / 4, lps = q4 + (ce % 4) / 2;
  hipLaunchKernelGGL(adj_write_blocks_kernel<T>, dim3((n + 63) / 64), dim3(64), 0, st, fblk, L.blk,
                     proto, plan, qcfg ? 4 : pcfg ? 2 : 0, (AdjArgs<T>*)nullptr, unused,
                     pack_status);
  VH_TRY_HIP(hipGetLastError());
  T* final_slot = wsf<T>(ws, L.off_ckpt, L.fb, 5 * L.nck);
  const int n_all = (int)(field_elems(g) * (size_t)nm);
  int p_tiles = 0, p_tiles_j = 0, q_tiles = 0, q_tiles_j = 0;
  if (pcfg) tiles_of(pcfg, g, &p_tiles, &p_tiles_j);
  if (qcfg) tiles_of(qcfg, g, &q_tiles, &q_tiles_j);
  const int head = qcfg ? (L.nck - 1) * lps : 0;
  const int n_launch = qcfg ? head + (n - (L.nck - 1) * ce) / 2 : pcfg ? n / 2 : n;
  auto record = [&](hipStream_t cs) {
    for (int kk = 0; kk < n_launch; ++kk) {
      const EvpArgBlock<T>* b = (const EvpArgBlock<T>*)((char*)fblk + (size_t)kk * L.blk);
      if (kk < head && kk % lps < q4) evp_launch_pair<T>(qcfg, b, q_tiles, q_tiles_j, cs, nm);
      else if (pcfg) evp_launch_pair<T>(pcfg, b, p_tiles, p_tiles_j, cs, nm);
      else evp_launch_single<T, double>(cfg, b, n_tiles, tiles_j, cs, nm, so, plain);
    }
    hipLaunchKernelGGL(adj_keep_final_kernel<T>, dim3((unsigned)((n_all + BLK - 1) / BLK)),
                       dim3(BLK), 0, cs,
                       (const EvpArgBlock<T>*)((char*)fblk + (size_t)(n_launch - 1) * L.blk),
                       n_all, final_slot, L.fb);
  };
  return run_recorded<T>(
      "adjf", g, L, (((cfg * 2 + (so ? 1 : 0)) * 2 + (plain ? 1 : 0)) * 32 + pcfg) * 32 + qcfg, nm, ws,
      record, st);
}

template <typename T>
int adjoint(const Geom& g, const vh_domain* d, const vh_settings* s, const vh_fields* f, int n,
            int ce, int nm, const void* const lam_final[5], const vh_evp_adjoint_out* o, void* ws,
            size_t wsb, hipStream_t st) {
  const AdjLayout L = adj_layout<T>(g, n, ce, nm);
  VH_REQUIRE(ws && wsb >= L.total, VH_EWORKSPACE, "adjoint workspace: need %zu bytes, got %zu",
             L.total, wsb);
  AdjArgs<T> av{};
  fill_evp_const(f, &av.c);
  fill_evp_scalars(s, kterms_of(f), s->basalDragK2 != 0.0 ? 1 : 0, &av.sc);
  av.g = g;
  av.acc = AdjAcc<T>{(T*)o->adj_SeaIceStrength, (T*)o->adj_SeaIceMassC, (T*)o->adj_SeaIceMassU,
                     (T*)o->adj_SeaIceMassV, (T*)o->adj_AreaW, (T*)o->adj_AreaS,
                     (T*)o->adj_WindForcingX, (T*)o->adj_WindForcingY, (T*)o->adj_uNm1,
                     (T*)o->adj_vNm1, (T*)o->adj_waterDrag_north, (T*)o->adj_waterDrag_south,
                     (T*)o->adj_hIceMean};
  const void* accp[kAdjAccFields] = {av.acc.P,   av.acc.mC, av.acc.mU, av.acc.mV, av.acc.AW,
                                     av.acc.AS,  av.acc.WFX, av.acc.WFY, av.acc.u0, av.acc.v0,
                                     av.acc.DN,  av.acc.DS, av.acc.h};
  const void* outs[5] = {o->adj_uIce, o->adj_vIce, o->adj_sigma1, o->adj_sigma2, o->adj_sigma12};
  const bool basal = s->basalDragK2 != 0.0;
  for (int q = 0; q < kAdjAccFields; ++q)
    VH_REQUIRE(accp[q] != nullptr || q == 4 || q == 5 || q == 8 || q == 9 || (q == 12 && !basal),
               VH_EINVAL, "adjoint: null output");
  VH_REQUIRE((av.acc.AW == nullptr) == (av.acc.AS == nullptr) &&
                 (av.acc.u0 == nullptr) == (av.acc.v0 == nullptr),
             VH_EINVAL, "adjoint: adj_AreaW/adj_AreaS and adj_uNm1/adj_vNm1 go in pairs");
  for (const void* p : outs) VH_REQUIRE(p != nullptr, VH_EINVAL, "adjoint: null output");
  for (int q = 0; q < 5; ++q) {
    av.lam_final[q] = (const T*)lam_final[q];
    av.outs[q] = (T*)outs[q];
  }
  AdjArgs<T>* a_dev = (AdjArgs<T>*)((char*)ws + L.off_args);
  EvpArgBlock<T>* rblk = (EvpArgBlock<T>*)((char*)ws + L.off_rblk);
  const EvpArgBlock<T>* pack_status = nullptr;
  const EvpArgBlock<T> proto = packed_proto<T>(g, s, f, L, ws, &pack_status, nm, st, false);
  hipLaunchKernelGGL(adj_write_blocks_kernel<T>, dim3((n + 63) / 64), dim3(64), 0, st, rblk, L.blk,
                     proto, block_plan<T>(L, f, nullptr, ws), 1, a_dev, av, pack_status);
  VH_TRY_HIP(hipGetLastError());
  const bool plain = evp_plain_kernels_enabled() && evp_plain_settings(av.sc);
  // the recompute runs two substeps per launch when the adjoint allows it
  const int pcfg = adj_pair_cfg(g, d->tile, plain, n, ce, field_elems(g) * (size_t)nm);
  VH_REQUIRE(pcfg >= 0, VH_EUNSUPPORTED,
             "adjoint: tile %d runs two substeps per launch, which needs the default settings "
             "and even n_substeps (%d) and ckpt_every (%d)", d->tile, n, ce);
  EvpArgBlock<T>* r2blk = (EvpArgBlock<T>*)((char*)ws + L.off_r2blk);
  if (pcfg) {
    hipLaunchKernelGGL(adj_write_blocks_kernel<T>, dim3((n + 63) / 64), dim3(64), 0, st, r2blk,
                       L.blk, proto, block_plan<T>(L, f, nullptr, ws), 3, (AdjArgs<T>*)nullptr,
                       av, pack_status);
    VH_TRY_HIP(hipGetLastError());
  }
  int p_tiles = 0, p_tiles_j = 0;
  if (pcfg) tiles_of(pcfg, g, &p_tiles, &p_tiles_j);

  const int cfg = pick_adj_tile(g, d->tile);
  int n_tiles, tiles_j;
  tiles_of(cfg, g, &n_tiles, &tiles_j);
  const int acfg = adj_tile_cfg<T>(field_elems(g) * (size_t)nm);
  const bool overlap = adj_overlap();
  // the instantiation: 1 for free slip with lateral drag, basal drag or
  // adaptive EVP, 2 when the second-order no-slip term is on as well
  const bool sobc = s->noSlip != 0.0 && s->secondOrderBC != 0.0;
  const int sd = sobc ? 2
                 : (s->noSlip == 0.0 || s->basalDragK2 != 0.0 || s->useAdaptiveEVP != 0.0) ? 1 : 0;
  T* lam[2][5];
  for (int q = 0; q < 5; ++q) {
    lam[0][q] = wsf<T>(ws, L.off_lam, L.fb, q);
    lam[1][q] = wsf<T>(ws, L.off_lam, L.fb, 5 + q);
  }
  const int n_all = (int)(field_elems(g) * (size_t)nm);
  const dim3 grid((unsigned)((n_all + BLK - 1) / BLK));
  auto sets = [&](int cur) {
    LamSet<T> ls;
    for (int q = 0; q < 5; ++q) {
      ls.in[q] = lam[cur][q];
      ls.out[q] = lam[1 - cur][q];
    }
    return ls;
  };
  SideStream side;
  int rc = side_stream(&side);
  if (rc) return rc;
  // the reverse instantiation: MODE 3 (the default settings with their runtime
  // branches compiled out) in FP32, where it is 2-3% faster; in FP64 the
  // default instantiation, which measured faster than MODE 3 (2.6 against 3.2
  // ms for the 1-degree 120-substep adjoint)
  const int rmode = sd == 0 && plain && sizeof(T) == 4 ? 3 : sd;
  auto recompute = [&](int sgm, hipStream_t cs) {
    const int k0 = sgm * ce, len = std::min(ce, n - k0);
    if (pcfg) {  // states jj+1 and jj+2 from state jj, up to state len-1
      for (int jj = 0; jj + 1 < len; jj += 2)
        evp_launch_pair<T>(pcfg, (const EvpArgBlock<T>*)((char*)r2blk + (size_t)(k0 + jj) * L.blk),
                           p_tiles, p_tiles_j, cs, nm);
      return;
    }
    for (int jj = 0; jj + 1 < len; ++jj)
      evp_launch_single<T, double>(
          cfg, (const EvpArgBlock<T>*)((char*)rblk + (size_t)(k0 + jj) * L.blk), n_tiles, tiles_j,
          cs, nm, sobc, plain);
  };
  // Segment sgm - 1 is recomputed on a second stream while segment sgm is
  // reversed (they use different slot sets); the fork follows the reverse of
  // segment sgm + 1, the last reader of the set that the recompute overwrites.
  auto record = [&](hipStream_t cs) {
    int cur = 0;
    LamSet<T> init;
    for (int q = 0; q < 5; ++q) {
      init.in[q] = nullptr;
      init.out[q] = lam[cur][q];
    }
    hipLaunchKernelGGL(adj_init_kernel<T>, grid, dim3(BLK), 0, cs, a_dev, init, n_all);
    // the last segment's states are where the checkpointed forward left them
    for (int sgm = L.nck - 1; sgm >= 0; --sgm) {
      if (sgm >= 1 && overlap) {
        hipEventRecord(side.fork, cs);
        hipStreamWaitEvent(side.stream, side.fork, 0);
        recompute(sgm - 1, side.stream);
        hipEventRecord(side.join, side.stream);
      }
      const int k0 = sgm * ce, len = std::min(ce, n - k0);
      for (int jj = len - 1; jj >= 0; --jj) {
        const EvpArgBlock<T>* b = (const EvpArgBlock<T>*)((char*)rblk + (size_t)(k0 + jj) * L.blk);
        adj_launch<T>(acfg, a_dev, b, sets(cur), g, nm, rmode, cs);
        cur = 1 - cur;
      }
      if (sgm >= 1 && overlap) hipStreamWaitEvent(cs, side.join, 0);
      if (sgm >= 1 && !overlap) recompute(sgm - 1, cs);
    }
    hipLaunchKernelGGL(adj_finish_kernel<T>, grid, dim3(BLK), 0, cs, a_dev, sets(cur), n_all);
  };
  // the key holds every choice baked into the recorded launches: shapes,
  // overlap and the extra-branch instantiation
  return run_recorded<T>(
      "adjr", g, L,
      ((((cfg * 16 + acfg) * 2 + (overlap ? 1 : 0)) * 4 + sd) * 2 + (plain ? 1 : 0)) * 32 + pcfg,
      nm, ws, record, st);
}

template <typename T>
int pre_evp_adjoint(const Geom& g, const vh_settings* s, const vh_fields* f, const void* pstarF,
                    const vh_evp_adjoint_out* in, const vh_pre_evp_adjoint_out* o, int nm,
                    hipStream_t st) {
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
  a.ahb = evp_ptr<T>(in->adj_hIceMean);
  a.apstar = (T*)o->adj_pStar;
  a.ah = (T*)o->adj_hIceMean;
  a.ahs = (T*)o->adj_hSnowMean;
  a.aCaN = (T*)o->adj_airDrag_north;
  a.aCaS = (T*)o->adj_airDrag_south;
  const void* req[] = {a.h, a.A, a.M, a.MU, a.MV, a.u, a.v, a.uW, a.vW, a.fC, a.rdxC, a.rdyC,
                       a.ssh, a.sp, a.AW, a.AS, a.aP, a.amC, a.amU, a.amV, a.aWFX, a.aWFY,
                       a.apstar, a.ah, a.ahs, a.aCaN, a.aCaS};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "pre_evp_adjoint: null pointer");
  a.load = nullptr;
  a.loadFac = (T)(s->gravity * s->seaIceLoadFac);
  if (s->useRealFreshWaterFlux != 0.0) {
    a.load = evp_ptr<T>(f->SeaIceLoad);
    VH_REQUIRE(a.load != nullptr, VH_EINVAL, "pre_evp_adjoint: null SeaIceLoad");
  }
  a.pstar = (T)s->pStar;
  a.cStar = (T)s->cStar;
  a.rhoI = (T)s->rhoIce;
  a.rhoS = (T)s->rhoSnow;
  a.rhoAir = (T)s->rhoAir;
  a.wmin = (T)s->wSpeedMin;
  a.wmin2 = (T)(s->wSpeedMin * s->wSpeedMin);
  a.grav = (T)s->gravity;
  a.rrho = (T)s->recip_rhoSea;
  const double ang = s->airTurnAngle * (M_PI / 180.0);
  a.sinWin = (T)sin(ang);
  a.cosWin = (T)cos(ang);
  a.relWind = s->useRelativeWind != 0.0;
  const int n = g.NX * g.NY;
  hipLaunchKernelGGL(pre_evp_adjoint_kernel<T>, dim3((n + BLK - 1) / BLK, nm), dim3(BLK), 0, st, g,
                     a);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

#undef CELL_SETUP
#undef CELL_SETUP_G
#undef AT

}  // namespace
}  // namespace vh

namespace vh {
namespace {

// Geometry of one member and the member count of an equal-member batch
// (every field holds n_members padded boxes back to back).
int batch_geom(const vh_domain* d, int nm, Geom* g) {
  int rc = make_geom(d, g);
  if (rc) return rc;
  VH_REQUIRE(nm >= 1 && nm <= 65535, VH_EINVAL, "n_members=%d", nm);
  // the kernels address every packed field with 32-bit byte offsets
  VH_REQUIRE(field_elems(*g) * (size_t)nm * (d->dtype == VH_F64 ? 8 : 4) < ((size_t)1 << 32),
             VH_ESHAPE, "%d members of %d x %d cells exceed 4 GiB per packed field", nm, g->nx,
             g->ny);
  return VH_OK;
}

int ws_bytes(const vh_domain* d, int nm, int n, int ce, size_t* bytes) {
  Geom g;
  int rc = batch_geom(d, nm, &g);
  if (rc) return rc;
  VH_REQUIRE(bytes && n >= 1 && ce >= 1, VH_EINVAL, "bad arguments");
  *bytes = d->dtype == VH_F64 ? adj_layout<double>(g, n, ce, nm).total
                              : adj_layout<float>(g, n, ce, nm).total;
  return VH_OK;
}

int fwd(const vh_domain* d, int nm, const vh_settings* s, const vh_fields* f, int n, int ce,
        void* const outs[5], void* ws, size_t wsb, void* stream) {
  Geom g;
  int rc = batch_geom(d, nm, &g);
  if (rc) return rc;
  rc = check_adj(d, s, f, n, ce);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64) return forward_ckpt<double>(g, d, s, f, n, ce, nm, outs, ws, wsb, st);
  return forward_ckpt<float>(g, d, s, f, n, ce, nm, outs, ws, wsb, st);
}

int rev(const vh_domain* d, int nm, const vh_settings* s, const vh_fields* f, int n, int ce,
        const void* const lam[5], const vh_evp_adjoint_out* out, void* ws, size_t wsb,
        void* stream) {
  Geom g;
  int rc = batch_geom(d, nm, &g);
  if (rc) return rc;
  rc = check_adj(d, s, f, n, ce);
  if (rc) return rc;
  VH_REQUIRE(out != nullptr, VH_EINVAL, "null output struct");
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64) return adjoint<double>(g, d, s, f, n, ce, nm, lam, out, ws, wsb, st);
  return adjoint<float>(g, d, s, f, n, ce, nm, lam, out, ws, wsb, st);
}

int pre_rev(const vh_domain* d, int nm, const vh_settings* s, const vh_fields* f,
            const void* psf, const vh_evp_adjoint_out* evp_adj, const vh_pre_evp_adjoint_out* out,
            void* stream) {
  Geom g;
  int rc = batch_geom(d, nm, &g);
  if (rc) return rc;
  VH_REQUIRE(s && f && evp_adj && out, VH_EINVAL, "bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64) return pre_evp_adjoint<double>(g, s, f, psf, evp_adj, out, nm, st);
  return pre_evp_adjoint<float>(g, s, f, psf, evp_adj, out, nm, st);
}

}  // namespace
}  // namespace vh

extern "C" int vh_evp_adjoint_workspace_bytes(const vh_domain* d, int32_t n_substeps,
                                              int32_t ckpt_every, size_t* bytes) {
  return vh::ws_bytes(d, 1, n_substeps, ckpt_every, bytes);
}

extern "C" int vh_evp_forward_checkpointed(const vh_domain* d, const vh_settings* s,
                                           const vh_fields* f, int32_t n_substeps,
                                           int32_t ckpt_every, void* uIce, void* vIce,
                                           void* sigma1, void* sigma2, void* sigma12,
                                           void* workspace, size_t workspace_bytes, void* stream) {
  void* outs[5] = {uIce, vIce, sigma1, sigma2, sigma12};
  return vh::fwd(d, 1, s, f, n_substeps, ckpt_every, outs, workspace, workspace_bytes, stream);
}

extern "C" int vh_evp_adjoint(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                              int32_t n_substeps, int32_t ckpt_every, const void* adj_uIce,
                              const void* adj_vIce, const void* adj_sigma1, const void* adj_sigma2,
                              const void* adj_sigma12, const vh_evp_adjoint_out* out,
                              void* workspace, size_t workspace_bytes, void* stream) {
  const void* lam[5] = {adj_uIce, adj_vIce, adj_sigma1, adj_sigma2, adj_sigma12};
  return vh::rev(d, 1, s, f, n_substeps, ckpt_every, lam, out, workspace, workspace_bytes, stream);
}

extern "C" int vh_pre_evp_adjoint(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                                  const void* pStarField, const vh_evp_adjoint_out* evp_adj,
                                  const vh_pre_evp_adjoint_out* out, void* stream) {
  return vh::pre_rev(d, 1, s, f, pStarField, evp_adj, out, stream);
}

extern "C" int vh_evp_adjoint_batched_workspace_bytes(const vh_domain* d, int32_t n_members,
                                                      int32_t n_substeps, int32_t ckpt_every,
                                                      size_t* bytes) {
  return vh::ws_bytes(d, n_members, n_substeps, ckpt_every, bytes);
}

extern "C" int vh_evp_forward_checkpointed_batched(const vh_domain* d, int32_t n_members,
                                                   const vh_settings* s, const vh_fields* f,
                                                   int32_t n_substeps, int32_t ckpt_every,
                                                   void* uIce, void* vIce, void* sigma1,
                                                   void* sigma2, void* sigma12, void* workspace,
                                                   size_t workspace_bytes, void* stream) {
  void* outs[5] = {uIce, vIce, sigma1, sigma2, sigma12};
  return vh::fwd(d, n_members, s, f, n_substeps, ckpt_every, outs, workspace, workspace_bytes,
                 stream);
}

extern "C" int vh_evp_adjoint_batched(const vh_domain* d, int32_t n_members, const vh_settings* s,
                                      const vh_fields* f, int32_t n_substeps, int32_t ckpt_every,
                                      const void* adj_uIce, const void* adj_vIce,
                                      const void* adj_sigma1, const void* adj_sigma2,
                                      const void* adj_sigma12, const vh_evp_adjoint_out* out,
                                      void* workspace, size_t workspace_bytes, void* stream) {
  const void* lam[5] = {adj_uIce, adj_vIce, adj_sigma1, adj_sigma2, adj_sigma12};
  return vh::rev(d, n_members, s, f, n_substeps, ckpt_every, lam, out, workspace, workspace_bytes,
                 stream);
}

extern "C" int vh_pre_evp_adjoint_batched(const vh_domain* d, int32_t n_members,
                                          const vh_settings* s, const vh_fields* f,
                                          const void* pStarField,
                                          const vh_evp_adjoint_out* evp_adj,
                                          const vh_pre_evp_adjoint_out* out, void* stream) {
  return vh::pre_rev(d, n_members, s, f, pStarField, evp_adj, out, stream);
}
