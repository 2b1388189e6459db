// Hand-written adjoint of the EVP subcycle (veris/evp_solver.py:22-319) with
// segment checkpointing, and of the pre-EVP forcing it depends on
// (veris/dynamics_routines.py:5-18, veris/area_mass.py:17-28,
// veris/dynsolver.py:7-84).
//
// Forward (vh_evp_forward_checkpointed): n single-substep launches of the
// fused substep kernel; the state after every `ckpt_every` substeps stays in
// the workspace.  Reverse (vh_evp_adjoint): for each segment, last first, the
// substeps are recomputed from the segment's checkpoint by the same kernel in
// its STORE form, which also keeps each substep's intermediates (e12, zeta,
// Delta, e11+e22, e11-e22, water drag); then each substep is reversed by
// three per-cell kernels in gather form -- each cell sums what its neighbours
// contribute, so no atomics are needed:
//   S  momentum update and fill (evaluated at the three u-points and three
//      v-points a c-point touches), c-point stress divergence, stress
//      relaxation, drag and Coriolis, z-point sigma12 relaxation
//   D  c-point zeta and Delta
//   V  z-point shear strain (at the three z-points a u/v-point touches) and
//      the adjoint of the substep's input velocities
// Every launch reads its pointers from argument blocks in the workspace that
// one small kernel writes per call, so the n forward launches and the whole
// reverse sweep are each replayed from one hipGraph.
//
// Non-smooth points: d sqrt(x)/dx is taken as 0 at x = 0 (Delta = 0 on
// motionless or masked cells, zero relative velocity); the denominator guard
// and the drag regime switch contribute no derivative.  Supported settings:
// non-adaptive EVP, noSlip, no turning angle (sinWat = 0), no basal drag, no
// metric terms -- the defaults.
#include <algorithm>
#include <cstdio>

#include "evp_kernel.h"

namespace vh {
namespace {

constexpr int BLK = 256;

template <typename T>
struct AdjAcc {  // parameter adjoints, accumulated over the substeps
  T *P, *mC, *mU, *mV, *AW, *AS, *WFX, *WFY, *u0, *v0, *DN, *DS;
};
constexpr int kAdjAccFields = 12;

template <typename T>
struct AdjWork {  // adjoint intermediates of one substep (workspace)
  T *UDIR, *VDIR, *ZBP, *EPB, *EMB, *DBP, *AB, *BB, *FUB, *FVB, *ZZB, *E12P, *E11B, *E22B, *QB;
};
constexpr int kAdjWorkFields = 15;

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
  size_t fb, off_ckpt, off_pp, off_seg, off_interm;
  int n, ce;
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

// forward location of field q of the state after k substeps
template <typename T>
__device__ inline T* fwd_loc(const BlockPlan<T>& p, int k, int q) {
  if (k == 0) return (T*)p.in0[q];
  if (k == p.n) return p.outs[q];
  if (k % p.ce == 0) return slot(p, p.off_ckpt, 5 * (k / p.ce) + q);
  return slot(p, p.off_pp, 5 * ((k - 1) % 2) + q);
}

// mode 0: forward blocks (substep k: state k -> state k+1);
// mode 1: recompute blocks (substep k inside its segment: segment slots, with
// the intermediates of segment position k % ce) and the reverse arguments.
template <typename T>
__global__ void adj_write_blocks_kernel(EvpArgBlock<T>* blocks, size_t stride, EvpArgBlock<T> proto,
                                        BlockPlan<T> p, int mode, AdjArgs<T>* aa, AdjArgs<T> av) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (mode == 1 && k == 0) *aa = av;
  if (k >= p.n) return;
  EvpArgBlock<T> b = proto;
  if (mode == 0) {
    for (int q = 0; q < 5; ++q) {
      b.in[q] = fwd_loc(p, k, q);
      b.out[q] = fwd_loc(p, k + 1, q);
    }
    for (int r = 0; r < 6; ++r) b.interm[r] = nullptr;
  } else {
    const int jj = k % p.ce, sgm = k / p.ce;
    for (int q = 0; q < 5; ++q) {
      b.in[q] = jj == 0 ? (k == 0 ? p.in0[q] : slot(p, p.off_ckpt, 5 * sgm + q))
                        : slot(p, p.off_seg, 5 * jj + q);
      b.out[q] = slot(p, p.off_seg, 5 * (jj + 1) + q);
    }
    for (int r = 0; r < 6; ++r) b.interm[r] = slot(p, p.off_interm, 6 * jj + r);
  }
  *(EvpArgBlock<T>*)((char*)blocks + (size_t)k * stride) = b;
}

// ---- reverse sweep: set-up and result copy
template <typename T>
__global__ __launch_bounds__(BLK) void adj_init_kernel(const AdjArgs<T>* __restrict__ a,
                                                       LamSet<T> L) {
  CELL_SETUP
  const AdjAcc<T>& acc = a->acc;
  T* const accp[kAdjAccFields] = {acc.P, acc.mC, acc.mU, acc.mV, acc.AW, acc.AS,
                                  acc.WFX, acc.WFY, acc.u0, acc.v0, acc.DN, acc.DS};
#pragma unroll
  for (int q = 0; q < kAdjAccFields; ++q) accp[q][c] = (T)0;
#pragma unroll
  for (int q = 0; q < 5; ++q) L.out[q][c] = a->lam_final[q] ? a->lam_final[q][c] : (T)0;
}

template <typename T>
__global__ __launch_bounds__(BLK) void adj_finish_kernel(const AdjArgs<T>* __restrict__ a,
                                                         LamSet<T> L) {
  CELL_SETUP
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

template <typename T>
struct MomAdj {
  T nb;   // adjoint of the numerator (lambda * mask / den)
  T db;   // adjoint of the denominator
  T cdx;  // adjoint of the averaged drag coefficient
  T cdA;  // the averaged drag coefficient
};

// Adjoint of the u-point (Y = 0) or v-point (Y = 1) momentum update at (i, j)
// (veris/evp_solver.py:152-187 with the fill of :190-191); zero on cells the
// fill overwrites.
template <typename T, int Y>
__device__ inline MomAdj<T> mom_point(const Geom& g, const EvpConst<T>& k, const EvpScalars& sc,
                                      const T* CD, const T* xn, const T* lam, int i, int j) {
  const int c = AT(i, j);
  const int cm = Y ? AT(i, wm(j, g.NY)) : AT(wm(i, g.NX), j);
  const T l = gather_filled(g, lam, i, j);
  const T cd0 = ldg(CD, c), cd1 = ldg(CD, cm);
  const T m = ldg(Y ? k.massV : k.massU, c);
  const T ar = ldg(Y ? k.AreaS : k.AreaW, c);
  const T oc = ldg(Y ? k.vOc : k.uOc, c);
  const T mk = ldg(Y ? k.MV : k.MU, c);
  const T x = ldg(xn, c);
  const T cw = (T)sc.cosWat;
  const T cdA = (T)0.5 * (cd0 + cd1);
  const T den = m * (T)sc.betaP1 + ar * (cdA * cw);
  const bool guard = den == (T)0;
  const T r = (T)1 / (guard ? (T)1 : den);
  const T lc = is_computed(g, i, j) ? l : (T)0;
  MomAdj<T> o;
  o.nb = lc * mk * r;
  o.db = guard ? (T)0 : -lc * x * r;
  o.cdx = (o.nb * oc + o.db) * (ar * cw);
  o.cdA = cdA;
  return o;
}

// ---- S: momentum, stresses (veris/evp_solver.py:79-103), drag (:108-110),
//         Coriolis (:142-143), stress divergence (veris/dynamics_routines.py:284-302)
template <typename T>
__global__ __launch_bounds__(BLK) void adj_stress_kernel(const AdjArgs<T>* __restrict__ a,
                                                         const EvpArgBlock<T>* __restrict__ blk,
                                                         LamSet<T> L, AdjWork<T> w) {
  CELL_SETUP
  const EvpConst<T> k = a->c;
  const EvpScalars& sc = a->sc;
  const AdjAcc<T> acc = a->acc;
  const T* u = blk->in[0];
  const T* v = blk->in[1];
  const T* un = blk->out[0];
  const T* vn = blk->out[1];
  const T* E12 = blk->interm[0];
  const T* Z = blk->interm[1];
  const T* DL = blk->interm[2];
  const T* EP = blk->interm[3];
  const T* EM = blk->interm[4];
  const T* CD = blk->interm[5];
  const int cip = AT(ip, j), cjp = AT(i, jp), cim = AT(im, j), cjm = AT(i, jm), cmm = AT(im, jm);
  // momentum adjoints at the u-points c, ip, jm and the v-points c, jp, im
  const MomAdj<T> U = mom_point<T, 0>(g, k, sc, CD, un, L.in[0], i, j);
  const MomAdj<T> Uip = mom_point<T, 0>(g, k, sc, CD, un, L.in[0], ip, j);
  const MomAdj<T> Ujm = mom_point<T, 0>(g, k, sc, CD, un, L.in[0], i, jm);
  const MomAdj<T> V = mom_point<T, 1>(g, k, sc, CD, vn, L.in[1], i, j);
  const MomAdj<T> Vjp = mom_point<T, 1>(g, k, sc, CD, vn, L.in[1], i, jp);
  const MomAdj<T> Vim = mom_point<T, 1>(g, k, sc, CD, vn, L.in[1], im, j);
  // c-point inputs
  const T uc = ldg(u, c), uip = ldg(u, cip), vc = ldg(v, c), vjp = ldg(v, cjp);
  const T uoc = ldg(k.uOc, c), uoip = ldg(k.uOc, cip), voc = ldg(k.vOc, c), vojp = ldg(k.vOc, cjp);
  const T miu = ldg(k.mInU, c), miuip = ldg(k.mInU, cip), miv = ldg(k.mInV, c),
          mivjp = ldg(k.mInV, cjp);
  const T Mc = ldg(k.M, c), Mim = ldg(k.M, cim), Mjm = ldg(k.M, cjm), Mmm = ldg(k.M, cmm);
  const T fc = ldg(k.fC, c);
  const T rrAu = ldg(k.rrAu, c), rrAuip = ldg(k.rrAu, cip), rrAujm = ldg(k.rrAu, cjm);
  const T rrAv = ldg(k.rrAv, c), rrAvjp = ldg(k.rrAv, cjp), rrAvim = ldg(k.rrAv, cim);
  const T dyV = ldg(k.dyV, c), dxU = ldg(k.dxU, c), dxV = ldg(k.dxV, c), dyU = ldg(k.dyU, c);
  const T mUc = ldg(k.massU, c), mVc = ldg(k.massV, c);
  const T uNm1 = ldg(k.uNm1, c), vNm1 = ldg(k.vNm1, c);
  const T l1 = ldg(L.in[2], c), l2 = ldg(L.in[3], c), l12 = ldg(L.in[4], c);
  const T zeta = ldg(Z, c), dC = ldg(DL, c), ep = ldg(EP, c), em = ldg(EM, c), e12 = ldg(E12, c);
  const T zim = ldg(Z, cim), zjm = ldg(Z, cjm), zmm = ldg(Z, cmm);
  const T aP = ldg(acc.P, c), amC = ldg(acc.mC, c), aDN = ldg(acc.DN, c), aDS = ldg(acc.DS, c);

  const T cw = (T)sc.cosWat, bU = (T)sc.betaU, bP1 = (T)sc.betaP1, rdt = (T)sc.rdt;
  if (is_computed(g, i, j)) {
    const T amU = ldg(acc.mU, c), au0 = ldg(acc.u0, c), aWX = ldg(acc.WFX, c), aAW = ldg(acc.AW, c);
    const T amV = ldg(acc.mV, c), av0 = ldg(acc.v0, c), aWY = ldg(acc.WFY, c), aAS = ldg(acc.AS, c);
    stg(acc.mU, c, amU + (U.nb * (bU * uc + rdt * uNm1) + U.db * bP1));
    stg(acc.u0, c, au0 + U.nb * mUc * rdt);
    stg(acc.WFX, c, aWX + U.nb);
    stg(acc.AW, c, aAW + (U.nb * (U.cdA * cw * uoc) + U.db * (U.cdA * cw)));
    stg(acc.mV, c, amV + (V.nb * (bU * vc + rdt * vNm1) + V.db * bP1));
    stg(acc.v0, c, av0 + V.nb * mVc * rdt);
    stg(acc.WFY, c, aWY + V.nb);
    stg(acc.AS, c, aAS + (V.nb * (V.cdA * cw * voc) + V.db * (V.cdA * cw)));
  }
  stg(w.UDIR, c, U.nb * bU * mUc);
  stg(w.VDIR, c, V.nb * bU * mVc);
  const T Xc = U.nb * rrAu, Xip = Uip.nb * rrAuip, Xjm = Ujm.nb * rrAujm;
  const T Yc = V.nb * rrAv, Yjp = Vjp.nb * rrAvjp, Yim = Vim.nb * rrAvim;

  // drag and Coriolis at c
  const T cdb = (T)0.5 * (U.cdx + Uip.cdx) + (T)0.5 * (V.cdx + Vjp.cdx);
  const T fvb = (T)0.5 * (U.nb + Uip.nb);
  const T fub = -(T)0.5 * (V.nb + Vjp.nb);
  stg(acc.mC, c, amC + (fvb * fc * (T)0.5 * (vc + vjp) + fub * fc * (T)0.5 * (uc + uip)));
  stg(w.FVB, c, fvb);
  stg(w.FUB, c, fub);
  const T cbe = cdb * Mc;
  const bool south = fc < (T)0;
  const T D = (T)(south ? sc.cwS : sc.cwN) * (T)sc.rhoSea;
  const T ad = (uc - uoc) * miu + (uip - uoip) * miuip;
  const T bd = (vc - voc) * miv + (vjp - vojp) * mivjp;
  const T ww = (T)0.25 * (ad * ad + bd * bd);
  const T sw = sqrt(ww);
  const bool drag = (D * D) * ww > (T)sc.cDragMin2;
  const T wb = (drag && sw > (T)0) ? cbe * D * (T)0.5 / sw : (T)0;
  const T Db = drag ? cbe * sw * (T)sc.rhoSea : (T)0;
  stg(w.AB, c, wb * (T)0.5 * ad);
  stg(w.BB, c, wb * (T)0.5 * bd);
  stg(acc.DS, c, aDS + (south ? Db : (T)0));
  stg(acc.DN, c, aDN + (south ? (T)0 : Db));

  // sigma1, sigma2 at c
  const T s11b = dyV * (Xc - Xip);
  const T s22b = dxU * (Yc - Yjp);
  const T s1t = l1 + (T)0.5 * (s11b + s22b);
  const T s2t = l2 + (T)0.5 * (s11b - s22b);
  const T H1 = s1t * (T)sc.den1 * Mc, H2 = s2t * (T)sc.den2 * Mc;
  stg(L.out[2], c, H1 * (T)sc.alphaM1);
  stg(L.out[3], c, H2 * (T)sc.alphaM1);
  const T pb = -H1;
  const T pfac = (T)sc.tens1m * (T)2 * (T)sc.pr / (T)sc.tens1p;
  stg(acc.P, c, aP + pb * (T)sc.tens1m * (T)sc.pr1m);
  stg(w.ZBP, c, H1 * (T)2 * ep + H2 * (T)0.5 * em + pb * pfac * dC);
  stg(w.EPB, c, H1 * (T)2 * zeta);
  stg(w.EMB, c, H2 * (T)0.5 * zeta);
  stg(w.DBP, c, pb * pfac * zeta);

  // sigma12 at z = c
  const T s12t = l12 + dxV * (Xjm - Xc) + dyU * (Yim - Yc);
  const T G = s12t * (T)sc.den2;
  stg(L.out[4], c, G * (T)sc.alphaM1);
  const T nz = (Mc + Mim) + (Mjm + Mmm);
  const T nrm = nz > (T)0 ? (T)1 / nz : (T)0;
  const T zz = nrm * ((zeta + zim) + (zjm + zmm));
  stg(w.ZZB, c, G * (T)0.5 * e12 * nrm);  // carries the z-average weight of z = c
  stg(w.E12P, c, G * (T)0.5 * zz);
}

// ---- D: c-point zeta and Delta (veris/dynamics_routines.py:220-268)
template <typename T>
__global__ __launch_bounds__(BLK) void adj_visc_kernel(const AdjArgs<T>* __restrict__ a,
                                                       const EvpArgBlock<T>* __restrict__ blk,
                                                       AdjWork<T> w) {
  CELL_SETUP
  const EvpConst<T> k = a->c;
  const EvpScalars& sc = a->sc;
  T* const accP = a->acc.P;
  const T* const* it = blk->interm;
  const T z0 = ldg(w.ZZB, c), z1 = ldg(w.ZZB, AT(ip, j)), z2 = ldg(w.ZZB, AT(i, jp)),
          z3 = ldg(w.ZZB, AT(ip, jp));
  const T zbp = ldg(w.ZBP, c), dbp = ldg(w.DBP, c), epb0 = ldg(w.EPB, c), emb0 = ldg(w.EMB, c);
  const T zeta = ldg(it[1], c), dC = ldg(it[2], c), ep = ldg(it[3], c), em = ldg(it[4], c);
  const T mi = ldg(k.mInC, c), rdxU = ldg(k.rdxU, c), rdyV = ldg(k.rdyV, c), rrA = ldg(k.rrA, c);
  const T aP = ldg(accP, c);
  const T zb = zbp + ((z0 + z1) + (z2 + z3));
  const T rreg = (T)1 / (dC + (T)sc.deltaMin);
  stg(accP, c, aP + zb * (T)0.5 * (T)sc.tens1p * rreg);
  const T Db = dbp - zb * zeta * rreg;
  const T dsqb = dC > (T)0 ? Db * (T)0.5 / dC : (T)0;
  const T epb = epb0 + dsqb * (T)2 * ep;
  const T emb = emb0 + dsqb * (T)sc.rpd * (T)2 * em;
  stg(w.E11B, c, (epb + emb) * rdxU * mi);
  stg(w.E22B, c, (epb - emb) * rdyV * mi);
  stg(w.QB, c, dsqb * (T)sc.rpd * rrA);
}

// Adjoint of e12 at z-point (ii, jj) times its coefficients in
// e12(z) = A1 u + A2 u[j-1] + A3 v + A4 v[i-1] (veris/dynamics_routines.py:160-175):
// returns the adjoint e and the coefficients.
template <typename T>
struct E12Adj {
  T e, A1, A2, A3, A4;
};

template <typename T>
__device__ inline E12Adj<T> e12_adj(const Geom& g, const EvpConst<T>& k, const T* E12,
                                    const AdjWork<T>& w, int ii, int jj) {
  const int im = wm(ii, g.NX), jm = wm(jj, g.NY);
  const int c = AT(ii, jj), cim = AT(im, jj), cjm = AT(ii, jm), cmm = AT(im, jm);
  const T q0 = ldg(w.QB, c), q1 = ldg(w.QB, cim), q2 = ldg(w.QB, cjm), q3 = ldg(w.QB, cmm);
  const T e12p = ldg(w.E12P, c), rAz = ldg(k.rAz, c), e12 = ldg(E12, c);
  const T M0 = ldg(k.M, c), M1 = ldg(k.M, cim), M2 = ldg(k.M, cjm), M3 = ldg(k.M, cmm);
  const T MU0 = ldg(k.MU, c), MU2 = ldg(k.MU, cjm), MV0 = ldg(k.MV, c), MV1 = ldg(k.MV, cim);
  const T ry = ldg(k.rdyU, c), rx = ldg(k.rdxV, c);
  E12Adj<T> o;
  o.e = e12p + ((q0 + q1) + (q2 + q3)) * (T)2 * rAz * e12;
  const T mz = (M0 * M1) * (M2 * M3);
  const T hu = MU0 - MU2, hv = MV0 - MV1;
  o.A1 = (T)0.5 * mz * ry + hu * ry;
  o.A2 = -(T)0.5 * mz * ry + hu * ry;
  o.A3 = (T)0.5 * mz * rx + hv * rx;
  o.A4 = -(T)0.5 * mz * rx + hv * rx;
  return o;
}

// ---- V: shear strain and the adjoint of the substep's input velocities
template <typename T>
__global__ __launch_bounds__(BLK) void adj_vel_kernel(const AdjArgs<T>* __restrict__ a,
                                                      const EvpArgBlock<T>* __restrict__ blk,
                                                      LamSet<T> L, AdjWork<T> w) {
  CELL_SETUP
  const EvpConst<T> k = a->c;
  const T* E12 = blk->interm[0];
  const E12Adj<T> zc = e12_adj(g, k, E12, w, i, j);
  const E12Adj<T> zjp = e12_adj(g, k, E12, w, i, jp);
  const E12Adj<T> zip = e12_adj(g, k, E12, w, ip, j);
  const int cim = AT(im, j), cjm = AT(i, jm);
  const T udir = ldg(w.UDIR, c), vdir = ldg(w.VDIR, c);
  const T ab0 = ldg(w.AB, c), ab1 = ldg(w.AB, cim), bb0 = ldg(w.BB, c), bb1 = ldg(w.BB, cjm);
  const T e11a = ldg(w.E11B, cim), e11b = ldg(w.E11B, c), e22a = ldg(w.E22B, cjm),
          e22b = ldg(w.E22B, c);
  const T fub0 = ldg(w.FUB, c), fub1 = ldg(w.FUB, cim), fvb0 = ldg(w.FVB, c), fvb1 = ldg(w.FVB, cjm);
  const T fc = ldg(k.fC, c), fim = ldg(k.fC, cim), fjm = ldg(k.fC, cjm);
  const T mC = ldg(k.massC, c), mCim = ldg(k.massC, cim), mCjm = ldg(k.massC, cjm);
  const T miu = ldg(k.mInU, c), miv = ldg(k.mInV, c);
  stg(L.out[0], c,
      udir + (ab0 + ab1) * miu + (e11a - e11b) + zc.e * zc.A1 + zjp.e * zjp.A2 +
          (T)0.5 * fc * mC * fub0 + (T)0.5 * fim * mCim * fub1);
  stg(L.out[1], c,
      vdir + (bb0 + bb1) * miv + (e22a - e22b) + zc.e * zc.A3 + zip.e * zip.A4 +
          (T)0.5 * fc * mC * fvb0 + (T)0.5 * fjm * mCjm * fvb1);
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
  T pstar, cStar, rhoI, rhoS, rhoAir, wmin, wmin2, grav, rrho, sinWin, cosWin;
  int relWind;
};

// d tau_c / d Ca at c-point (i, j) (veris/dynsolver.py:7-47)
template <typename T>
__device__ inline void tau_c_dca(const Geom& g, const PreAdjArgs<T>& a, int i, int j, T* dtx,
                                 T* dty) {
  const int ip = wp(i, g.NX), jp = wp(j, g.NY);
  const int c = AT(i, j);
  T urel = a.uW[c], vrel = a.vW[c];
  if (a.relWind) {
    urel = urel - (T)0.5 * (a.u[c] + a.u[AT(ip, j)]);
    vrel = vrel - (T)0.5 * (a.v[c] + a.v[AT(i, jp)]);
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
  // strength: P = P* h exp(-C (1 - A)) M
  const T E = exp((-a.cStar) * ((T)1 - a.A[c]));
  const T ps = a.pstarF ? a.pstarF[c] : a.pstar;
  a.apstar[c] = a.aP[c] * a.h[c] * E * a.M[c];
  T hb = a.aP[c] * ps * E * a.M[c];
  // wind forcing: the tilt term -massU rdxC (phi - phi[i-1]) adds to the mass adjoints
  auto phi_at = [&](int ii, int jj) {
    return a.grav * a.ssh[AT(ii, jj)] + a.sp[AT(ii, jj)] * a.rrho;
  };
  const T phi = phi_at(i, j);
  const T amU_c = a.amU[c] - a.aWFX[c] * a.rdxC[c] * (phi - phi_at(im, j));
  const T amU_ip = a.amU[AT(ip, j)] - a.aWFX[AT(ip, j)] * a.rdxC[AT(ip, j)] * (phi_at(ip, j) - phi);
  const T amV_c = a.amV[c] - a.aWFY[c] * a.rdyC[c] * (phi - phi_at(i, jm));
  const T amV_jp = a.amV[AT(i, jp)] - a.aWFY[AT(i, jp)] * a.rdyC[AT(i, jp)] * (phi_at(i, jp) - phi);
  // mass: massC = rhoI h + rhoS hs, massU = (massC + massC[i-1])/2, massV likewise in j
  const T amC = a.amC[c] + (T)0.5 * (amU_c + amU_ip) + (T)0.5 * (amV_c + amV_jp);
  a.ah[c] = hb + a.rhoI * amC;
  a.ahs[c] = a.rhoS * amC;
  // air drag: WindForcingX = 0.5 (tau_c + tau_c[i-1]) MU AreaW + tilt, tau_c linear in Ca
  T dtx, dty;
  tau_c_dca(g, a, i, j, &dtx, &dty);
  const T gx = (T)0.5 * (a.aWFX[c] * a.AW[c] * a.MU[c] +
                         a.aWFX[AT(ip, j)] * a.AW[AT(ip, j)] * a.MU[AT(ip, j)]);
  const T gy = (T)0.5 * (a.aWFY[c] * a.AS[c] * a.MV[c] +
                         a.aWFY[AT(i, jp)] * a.AS[AT(i, jp)] * a.MV[AT(i, jp)]);
  const T contrib = gx * dtx + gy * dty;
  a.aCaN[c] = a.fC[c] < (T)0 ? (T)0 : contrib;
  a.aCaS[c] = a.fC[c] < (T)0 ? contrib : (T)0;
}

// ---- host side
struct AdjLayout {
  size_t fb, blk, total;
  int nck, ce, n;
  size_t off_args, off_fblk, off_rblk, off_ckpt, off_pp, off_seg, off_interm, off_work, off_lam;
};

template <typename T>
AdjLayout adj_layout(const Geom& g, int n, int ce) {
  AdjLayout L;
  L.fb = align256(field_elems(g) * sizeof(T));
  L.blk = align256(sizeof(EvpArgBlock<T>));
  L.ce = ce;
  L.n = n;
  L.nck = (n + ce - 1) / ce;
  L.off_args = 0;
  L.off_fblk = align256(sizeof(AdjArgs<T>));
  L.off_rblk = L.off_fblk + (size_t)n * L.blk;
  L.off_ckpt = L.off_rblk + (size_t)n * L.blk;
  L.off_pp = L.off_ckpt + (size_t)5 * L.nck * L.fb;  // checkpoint slot 0 is unused
  L.off_seg = L.off_pp + (size_t)10 * L.fb;
  L.off_interm = L.off_seg + (size_t)5 * (ce + 1) * L.fb;  // segment slot 0 is unused
  L.off_work = L.off_interm + (size_t)6 * ce * L.fb;
  L.off_lam = L.off_work + (size_t)kAdjWorkFields * L.fb;
  L.total = L.off_lam + (size_t)10 * L.fb;
  return L;
}

template <typename T>
T* wsf(void* ws, size_t off, size_t fb, int q) {
  return (T*)((char*)ws + off + (size_t)q * fb);
}

int pick_adj_tile(const Geom& g, int req) {
  if (req > 0) return req;
  return (long long)g.NX * g.NY < 400000 ? 5 : 3;
}

int check_adj(const vh_domain* d, const vh_settings* s, const vh_fields* f, int n, int ce) {
  VH_REQUIRE(d && s && f, VH_EINVAL, "null domain, settings or fields");
  VH_REQUIRE(n >= 1 && ce >= 1, VH_EINVAL, "n_substeps=%d ckpt_every=%d", n, ce);
  VH_REQUIRE(n <= (1 << 20), VH_EINVAL, "n_substeps=%d too large", n);
  VH_REQUIRE(d->tile >= 0 && d->tile < kEvpNumTiles && (d->tile == 0 || kEvpTiles[d->tile].k == 1),
             VH_EINVAL, "adjoint: tile %d (needs a single-substep tile shape or 0)", d->tile);
  VH_REQUIRE(s->useAdaptiveEVP == 0.0 && s->noSlip != 0.0 && s->secondOrderBC == 0.0 &&
                 s->useEVP != 0.0,
             VH_EUNSUPPORTED, "adjoint: only non-adaptive noSlip EVP is supported");
  VH_REQUIRE(s->sinWat == 0.0, VH_EUNSUPPORTED, "adjoint: turning angle (sinWat != 0) not supported");
  VH_REQUIRE(s->basalDragK2 == 0.0, VH_EUNSUPPORTED, "adjoint: basal drag not supported");
  VH_REQUIRE(!f->k1AtC && !f->k2AtC && !f->k1AtZ && !f->k2AtZ, VH_EUNSUPPORTED,
             "adjoint: metric terms k1/k2 not supported (pass NULL for all-zero fields)");
  const void* req[kEvpMaxRequired];
  const int nreq = evp_required_fields(f, req);
  for (int q = 0; q < nreq; ++q)
    VH_REQUIRE(req[q] != nullptr, VH_EINVAL, "adjoint: missing input field");
  return VH_OK;
}

template <typename T>
EvpArgBlock<T> proto_block(const Geom& g, const vh_settings* s, const vh_fields* f) {
  EvpArgBlock<T> ab{};
  fill_evp_scalars(s, 0, 0, &ab.sc);
  fill_evp_const(f, &ab.c);
  ab.g = g;
  ab.nlaunch = 1;
  ab.members = nullptr;
  return ab;
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
  p.off_pp = L.off_pp;
  p.off_seg = L.off_seg;
  p.off_interm = L.off_interm;
  p.n = L.n;
  p.ce = L.ce;
  return p;
}

void tiles_of(int cfg, const Geom& g, int* n_tiles, int* tiles_j) {
  const int TI = evp_tile_extent_i(cfg), TJ = evp_tile_extent_j(cfg);
  *tiles_j = (g.NY + TJ - 1) / TJ;
  *n_tiles = ((g.NX + TI - 1) / TI) * *tiles_j;
}

template <typename T>
int run_recorded(const char* kind, const Geom& g, const AdjLayout& L, int cfg, void* ws,
                 const std::function<void(hipStream_t)>& record, hipStream_t st) {
  if (graphs_enabled()) {
    char key[192];
    snprintf(key, sizeof(key), "%s/%d/%d/%d/%d/%d/%d/%d/%d/%d/%p", kind, (int)sizeof(T), cfg, L.n,
             L.ce, g.NX, g.NY, g.nx, g.fill_x, g.fill_y, ws);
    return graph_run(key, record, st);
  }
  record(st);
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

template <typename T>
int forward_ckpt(const Geom& g, const vh_domain* d, const vh_settings* s, const vh_fields* f,
                 int n, int ce, void* const outs[5], void* ws, size_t wsb, hipStream_t st) {
  const AdjLayout L = adj_layout<T>(g, n, ce);
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
  hipLaunchKernelGGL(adj_write_blocks_kernel<T>, dim3((n + 63) / 64), dim3(64), 0, st, fblk, L.blk,
                     proto_block<T>(g, s, f), block_plan<T>(L, f, outs, ws), 0,
                     (AdjArgs<T>*)nullptr, unused);
  VH_TRY_HIP(hipGetLastError());
  auto record = [&](hipStream_t cs) {
    for (int kk = 0; kk < n; ++kk)
      evp_launch_single<T, double, false>(
          cfg, (const EvpArgBlock<T>*)((char*)fblk + (size_t)kk * L.blk), n_tiles, tiles_j, cs);
  };
  return run_recorded<T>("adjf", g, L, cfg, ws, record, st);
}

template <typename T>
int adjoint(const Geom& g, const vh_domain* d, const vh_settings* s, const vh_fields* f, int n,
            int ce, const void* const lam_final[5], const vh_evp_adjoint_out* o, void* ws,
            size_t wsb, hipStream_t st) {
  const AdjLayout L = adj_layout<T>(g, n, ce);
  VH_REQUIRE(ws && wsb >= L.total, VH_EWORKSPACE, "adjoint workspace: need %zu bytes, got %zu",
             L.total, wsb);
  AdjArgs<T> av{};
  fill_evp_const(f, &av.c);
  fill_evp_scalars(s, 0, 0, &av.sc);
  av.g = g;
  av.acc = AdjAcc<T>{(T*)o->adj_SeaIceStrength, (T*)o->adj_SeaIceMassC, (T*)o->adj_SeaIceMassU,
                     (T*)o->adj_SeaIceMassV, (T*)o->adj_AreaW, (T*)o->adj_AreaS,
                     (T*)o->adj_WindForcingX, (T*)o->adj_WindForcingY, (T*)o->adj_uNm1,
                     (T*)o->adj_vNm1, (T*)o->adj_waterDrag_north, (T*)o->adj_waterDrag_south};
  const void* accp[kAdjAccFields] = {av.acc.P, av.acc.mC, av.acc.mU, av.acc.mV,
                                     av.acc.AW, av.acc.AS, av.acc.WFX, av.acc.WFY,
                                     av.acc.u0, av.acc.v0, av.acc.DN, av.acc.DS};
  const void* outs[5] = {o->adj_uIce, o->adj_vIce, o->adj_sigma1, o->adj_sigma2, o->adj_sigma12};
  for (const void* p : accp) VH_REQUIRE(p != nullptr, VH_EINVAL, "adjoint: null output");
  for (const void* p : outs) VH_REQUIRE(p != nullptr, VH_EINVAL, "adjoint: null output");
  for (int q = 0; q < 5; ++q) {
    av.lam_final[q] = (const T*)lam_final[q];
    av.outs[q] = (T*)outs[q];
  }
  AdjArgs<T>* a_dev = (AdjArgs<T>*)((char*)ws + L.off_args);
  EvpArgBlock<T>* rblk = (EvpArgBlock<T>*)((char*)ws + L.off_rblk);
  hipLaunchKernelGGL(adj_write_blocks_kernel<T>, dim3((n + 63) / 64), dim3(64), 0, st, rblk, L.blk,
                     proto_block<T>(g, s, f), block_plan<T>(L, f, nullptr, ws), 1, a_dev, av);
  VH_TRY_HIP(hipGetLastError());

  const int cfg = pick_adj_tile(g, d->tile);
  int n_tiles, tiles_j;
  tiles_of(cfg, g, &n_tiles, &tiles_j);
  AdjWork<T> w;
  T** wp_[kAdjWorkFields] = {&w.UDIR, &w.VDIR, &w.ZBP, &w.EPB, &w.EMB, &w.DBP, &w.AB, &w.BB,
                             &w.FUB, &w.FVB, &w.ZZB, &w.E12P, &w.E11B, &w.E22B, &w.QB};
  for (int q = 0; q < kAdjWorkFields; ++q) *wp_[q] = wsf<T>(ws, L.off_work, L.fb, q);
  T* lam[2][5];
  for (int q = 0; q < 5; ++q) {
    lam[0][q] = wsf<T>(ws, L.off_lam, L.fb, q);
    lam[1][q] = wsf<T>(ws, L.off_lam, L.fb, 5 + q);
  }
  const dim3 grid((unsigned)((field_elems(g) + BLK - 1) / BLK));
  auto sets = [&](int cur) {
    LamSet<T> ls;
    for (int q = 0; q < 5; ++q) {
      ls.in[q] = lam[cur][q];
      ls.out[q] = lam[1 - cur][q];
    }
    return ls;
  };
  auto record = [&](hipStream_t cs) {
    int cur = 0;
    LamSet<T> init;
    for (int q = 0; q < 5; ++q) {
      init.in[q] = nullptr;
      init.out[q] = lam[cur][q];
    }
    hipLaunchKernelGGL(adj_init_kernel<T>, grid, dim3(BLK), 0, cs, a_dev, init);
    for (int sgm = L.nck - 1; sgm >= 0; --sgm) {
      const int k0 = sgm * ce, len = std::min(ce, n - k0);
      for (int jj = 0; jj < len; ++jj)
        evp_launch_single<T, double, true>(
            cfg, (const EvpArgBlock<T>*)((char*)rblk + (size_t)(k0 + jj) * L.blk), n_tiles,
            tiles_j, cs);
      for (int jj = len - 1; jj >= 0; --jj) {
        const EvpArgBlock<T>* b = (const EvpArgBlock<T>*)((char*)rblk + (size_t)(k0 + jj) * L.blk);
        const LamSet<T> ls = sets(cur);
        hipLaunchKernelGGL(adj_stress_kernel<T>, grid, dim3(BLK), 0, cs, a_dev, b, ls, w);
        hipLaunchKernelGGL(adj_visc_kernel<T>, grid, dim3(BLK), 0, cs, a_dev, b, w);
        hipLaunchKernelGGL(adj_vel_kernel<T>, grid, dim3(BLK), 0, cs, a_dev, b, ls, w);
        cur = 1 - cur;
      }
    }
    hipLaunchKernelGGL(adj_finish_kernel<T>, grid, dim3(BLK), 0, cs, a_dev, sets(cur));
  };
  return run_recorded<T>("adjr", g, L, cfg, ws, record, st);
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
#undef CELL_SETUP_G
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
  rc = vh::check_adj(d, s, f, n_substeps, ckpt_every);
  if (rc) return rc;
  void* outs[5] = {uIce, vIce, sigma1, sigma2, sigma12};
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64)
    return vh::forward_ckpt<double>(g, d, s, f, n_substeps, ckpt_every, outs, workspace,
                                    workspace_bytes, st);
  return vh::forward_ckpt<float>(g, d, s, f, n_substeps, ckpt_every, outs, workspace,
                                 workspace_bytes, st);
}

extern "C" int vh_evp_adjoint(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                              int32_t n_substeps, int32_t ckpt_every, const void* adj_uIce,
                              const void* adj_vIce, const void* adj_sigma1, const void* adj_sigma2,
                              const void* adj_sigma12, const vh_evp_adjoint_out* out,
                              void* workspace, size_t workspace_bytes, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  rc = vh::check_adj(d, s, f, n_substeps, ckpt_every);
  if (rc) return rc;
  VH_REQUIRE(out != nullptr, VH_EINVAL, "null output struct");
  const void* lam[5] = {adj_uIce, adj_vIce, adj_sigma1, adj_sigma2, adj_sigma12};
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64)
    return vh::adjoint<double>(g, d, s, f, n_substeps, ckpt_every, lam, out, workspace,
                               workspace_bytes, st);
  return vh::adjoint<float>(g, d, s, f, n_substeps, ckpt_every, lam, out, workspace,
                            workspace_bytes, st);
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
