// Single-domain EVP subcycle (veris/evp_solver.py:257-319): host side of the
// C ABI vh_evp_solver.  The substep kernel is evp_substep_kernel
// (evp_kernel.h); this file is built with -ffp-contract=off so the FP64
// instantiation evaluates the reference expressions in their own order.
//
// Launch plan per call:
//   1. one tiny kernel copies the argument block (all field pointers, the
//      ping-pong state pointers and the scalars) into the caller's workspace;
//   2. the n substep launches read it from there.  For n >= 8 they are
//      replayed from a hipGraph captured once per (device, geometry, tile, n,
//      workspace address), so moving inputs and outputs never forces a
//      re-capture and a 400-substep subcycle costs one graph launch.
#include <map>
#include <mutex>
#include <tuple>

#include "evp_kernel.h"

namespace vh {
namespace {

struct TileCfg {
  int ri, rj, k;  // workgroup shape and substeps per launch
};
// index 0 = library default (chosen from the domain size)
constexpr TileCfg kTiles[] = {{0, 0, 0},   {8, 64, 1},  {12, 64, 1}, {16, 64, 1}, {8, 32, 1},
                              {16, 32, 1}, {32, 32, 1}, {6, 64, 1},  {16, 64, 2}, {32, 32, 2},
                              {12, 64, 2}, {32, 32, 3}, {16, 64, 3}, {32, 32, 4}, {16, 32, 2},
                              {24, 32, 2}};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

int pick_tile(const Geom& g, int req) {
  if (req > 0 && req < kNumTiles) return req;
  const long long cells = (long long)g.NX * g.NY;
  return cells < 400000 ? 5 : 3;
}

template <typename T, typename TD, int RI, int RJ, int K>
void launch_one(const EvpArgBlock<T>* ab, int launch, const Geom& g, hipStream_t st) {
  constexpr int TI = RI - 4 * K, TJ = RJ - 4 * K;
  const int tiles_i = (g.NX + TI - 1) / TI, tiles_j = (g.NY + TJ - 1) / TJ;
  const int n_tiles = tiles_i * tiles_j;
  const int per_xcd = (n_tiles + 7) / 8;
  hipLaunchKernelGGL((evp_substep_kernel<T, TD, RI, RJ, K, false>), dim3(per_xcd * 8),
                     dim3(RI * RJ), 0, st, ab, launch, n_tiles, tiles_j, per_xcd);
}

template <typename T, typename TD>
void launch_cfg(int cfg, const EvpArgBlock<T>* ab, int launch, const Geom& g, hipStream_t st) {
  switch (cfg) {
    case 1: launch_one<T, TD, 8, 64, 1>(ab, launch, g, st); break;
    case 2: launch_one<T, TD, 12, 64, 1>(ab, launch, g, st); break;
    case 3: launch_one<T, TD, 16, 64, 1>(ab, launch, g, st); break;
    case 4: launch_one<T, TD, 8, 32, 1>(ab, launch, g, st); break;
    case 5: launch_one<T, TD, 16, 32, 1>(ab, launch, g, st); break;
    case 6: launch_one<T, TD, 32, 32, 1>(ab, launch, g, st); break;
    case 7: launch_one<T, TD, 6, 64, 1>(ab, launch, g, st); break;
    case 8: launch_one<T, TD, 16, 64, 2>(ab, launch, g, st); break;
    case 9: launch_one<T, TD, 32, 32, 2>(ab, launch, g, st); break;
    case 10: launch_one<T, TD, 12, 64, 2>(ab, launch, g, st); break;
    case 11: launch_one<T, TD, 32, 32, 3>(ab, launch, g, st); break;
    case 12: launch_one<T, TD, 16, 64, 3>(ab, launch, g, st); break;
    case 13: launch_one<T, TD, 32, 32, 4>(ab, launch, g, st); break;
    case 14: launch_one<T, TD, 16, 32, 2>(ab, launch, g, st); break;
    default: launch_one<T, TD, 24, 32, 2>(ab, launch, g, st); break;
  }
}

// A subcycle of nsub substeps with configuration cfg is nsub/K launches of
// cfg followed by nsub%K single-substep launches of a K=1 shape.
constexpr int kRemainderCfg = 5;

int n_launches(int cfg, int nsub) { return nsub / kTiles[cfg].k + nsub % kTiles[cfg].k; }

template <typename T, typename TD>
void launch_substeps(int cfg, const EvpArgBlock<T>* ab, int nsub, const Geom& g, hipStream_t st) {
  const int K = kTiles[cfg].k;
  int l = 0;
  for (int q = 0; q < nsub / K; ++q) launch_cfg<T, TD>(cfg, ab, l++, g, st);
  for (int q = 0; q < nsub % K; ++q) launch_cfg<T, TD>(kRemainderCfg, ab, l++, g, st);
}

// ---- graph cache
using GraphKey = std::tuple<int /*dev*/, int /*dtype*/, int /*cfg*/, int /*nsub*/, int, int, int,
                            int, const void* /*arg block*/>;
std::mutex g_mu;
std::map<GraphKey, hipGraphExec_t> g_graphs;
std::map<int, hipStream_t> g_capture_streams;

int graphs_enabled() {
  static int en = -1;
  if (en < 0) {
    const char* e = getenv("VERIS_EVP_GRAPH");
    en = (e && e[0] == '0') ? 0 : 1;
  }
  return en;
}

template <typename T, typename TD>
int run_graph(int cfg, const EvpArgBlock<T>* ab_dev, int nsub, const Geom& g, int dtype,
              hipStream_t st) {
  int dev = 0;
  VH_TRY_HIP(hipGetDevice(&dev));
  GraphKey key{dev, dtype, cfg, nsub, g.NX, g.NY, g.fill_x * 2 + g.fill_y, g.nx, (const void*)ab_dev};
  hipGraphExec_t exec = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_graphs.find(key);
    if (it != g_graphs.end()) exec = it->second;
  }
  if (!exec) {
    hipStream_t cs;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      auto it = g_capture_streams.find(dev);
      if (it == g_capture_streams.end()) {
        VH_TRY_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        g_capture_streams[dev] = cs;
      } else {
        cs = it->second;
      }
    }
    hipGraph_t graph = nullptr;
    VH_TRY_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    launch_substeps<T, TD>(cfg, ab_dev, nsub, g, cs);
    VH_TRY_HIP(hipStreamEndCapture(cs, &graph));
    VH_TRY_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    VH_TRY_HIP(hipGraphDestroy(graph));
    std::lock_guard<std::mutex> lk(g_mu);
    auto ins = g_graphs.emplace(key, exec);
    if (!ins.second) {  // another thread won the race
      hipGraphExecDestroy(exec);
      exec = ins.first->second;
    }
  }
  VH_TRY_HIP(hipGraphLaunch(exec, st));
  return VH_OK;
}

template <typename T>
const T* P_(const void* p) {
  return static_cast<const T*>(p);
}

template <typename T>
int run_evp(const Geom& g, const vh_settings* s, const vh_fields* fl, int nsub, void* uo, void* vo,
            void* s1o, void* s2o, void* s12o, void* ws, size_t ws_bytes, int tile_req,
            int dtype, hipStream_t st) {
  const size_t fb = align256(field_elems(g) * sizeof(T));
  const size_t need = 5 * fb + align256(sizeof(EvpArgBlock<T>));
  VH_REQUIRE(ws != nullptr && ws_bytes >= need, VH_EWORKSPACE,
             "evp workspace: need %zu bytes, got %zu", need, ws_bytes);
  const void* req[] = {fl->uIce, fl->vIce, fl->sigma1, fl->sigma2, fl->sigma12,
                       fl->SeaIceStrength, fl->SeaIceMassC, fl->SeaIceMassU, fl->SeaIceMassV,
                       fl->AreaW, fl->AreaS, fl->WindForcingX, fl->WindForcingY, fl->uOcean,
                       fl->vOcean, fl->fCori, fl->iceMask, fl->iceMaskU, fl->iceMaskV,
                       fl->maskInC, fl->maskInU, fl->maskInV, fl->recip_dxU, fl->recip_dyV,
                       fl->recip_dyU, fl->recip_dxV, fl->dxU, fl->dyU, fl->dxV, fl->dyV,
                       fl->recip_rAu, fl->recip_rAv, fl->rAz, fl->recip_rA};
  for (const void* p : req) VH_REQUIRE(p != nullptr, VH_EINVAL, "evp_solver: missing input field");
  const void* outs[] = {uo, vo, s1o, s2o, s12o};
  for (int q = 0; q < 5; ++q) {
    VH_REQUIRE(outs[q] != nullptr, VH_EINVAL, "evp_solver: null output");
    for (const void* p : req)
      VH_REQUIRE(outs[q] != p, VH_EINVAL, "evp_solver: output aliases an input field");
    for (int r = 0; r < q; ++r)
      VH_REQUIRE(outs[q] != outs[r], VH_EINVAL, "evp_solver: two outputs alias");
  }
  VH_REQUIRE(s->useAdaptiveEVP == 0.0, VH_EUNSUPPORTED, "useAdaptiveEVP is not supported");
  VH_REQUIRE(s->noSlip != 0.0, VH_EUNSUPPORTED, "noSlip=False (side drag) is not supported");
  VH_REQUIRE(s->secondOrderBC == 0.0, VH_EUNSUPPORTED, "secondOrderBC is not supported");
  VH_REQUIRE(s->useEVP != 0.0, VH_EUNSUPPORTED, "useEVP=False (free drift) is not supported");
  const int kt = (fl->k1AtC || fl->k2AtC || fl->k1AtZ || fl->k2AtZ) ? 1 : 0;
  VH_REQUIRE(!kt || (fl->k1AtC && fl->k2AtC && fl->k1AtZ && fl->k2AtZ), VH_EINVAL,
             "k1AtC/k2AtC/k1AtZ/k2AtZ must be all set or all NULL");
  const int basal = s->basalDragK2 != 0.0 ? 1 : 0;
  if (basal)
    VH_REQUIRE(fl->R_low && fl->Area && fl->hIceMean, VH_EINVAL,
               "basal drag needs R_low, Area, hIceMean");

  EvpArgBlock<T> ab{};
  EvpScalars& sc = ab.sc;
  sc.rpd = 1.0 / (s->PlasDefCoeff * s->PlasDefCoeff);
  sc.deltaMin = s->deltaMin;
  sc.tens1p = 1 + s->tensileStrFac;
  sc.tens1m = 1 - s->tensileStrFac;
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
  sc.kterms = kt;
  sc.turn = s->sinWat != 0.0 ? 1 : 0;
  sc.basal = basal;

  EvpConst<T>& f = ab.c;
  f.uNm1 = P_<T>(fl->uIce);
  f.vNm1 = P_<T>(fl->vIce);
  f.P = P_<T>(fl->SeaIceStrength);
  f.massC = P_<T>(fl->SeaIceMassC);
  f.massU = P_<T>(fl->SeaIceMassU);
  f.massV = P_<T>(fl->SeaIceMassV);
  f.AreaW = P_<T>(fl->AreaW);
  f.AreaS = P_<T>(fl->AreaS);
  f.WFX = P_<T>(fl->WindForcingX);
  f.WFY = P_<T>(fl->WindForcingY);
  f.uOc = P_<T>(fl->uOcean);
  f.vOc = P_<T>(fl->vOcean);
  f.fC = P_<T>(fl->fCori);
  f.M = P_<T>(fl->iceMask);
  f.MU = P_<T>(fl->iceMaskU);
  f.MV = P_<T>(fl->iceMaskV);
  f.mInC = P_<T>(fl->maskInC);
  f.mInU = P_<T>(fl->maskInU);
  f.mInV = P_<T>(fl->maskInV);
  f.rdxU = P_<T>(fl->recip_dxU);
  f.rdyV = P_<T>(fl->recip_dyV);
  f.rdyU = P_<T>(fl->recip_dyU);
  f.rdxV = P_<T>(fl->recip_dxV);
  f.dxU = P_<T>(fl->dxU);
  f.dyU = P_<T>(fl->dyU);
  f.dxV = P_<T>(fl->dxV);
  f.dyV = P_<T>(fl->dyV);
  f.rrAu = P_<T>(fl->recip_rAu);
  f.rrAv = P_<T>(fl->recip_rAv);
  f.rAz = P_<T>(fl->rAz);
  f.rrA = P_<T>(fl->recip_rA);
  f.k1C = P_<T>(fl->k1AtC);
  f.k2C = P_<T>(fl->k2AtC);
  f.k1Z = P_<T>(fl->k1AtZ);
  f.k2Z = P_<T>(fl->k2AtZ);
  f.Rlow = P_<T>(fl->R_low);
  f.Area = P_<T>(fl->Area);
  f.hIce = P_<T>(fl->hIceMean);

  const T* inp[5] = {P_<T>(fl->uIce), P_<T>(fl->vIce), P_<T>(fl->sigma1), P_<T>(fl->sigma2),
                     P_<T>(fl->sigma12)};
  T* outp[5] = {(T*)uo, (T*)vo, (T*)s1o, (T*)s2o, (T*)s12o};
  const size_t nb = field_elems(g) * sizeof(T);
  if (nsub == 0) {
    for (int q = 0; q < 5; ++q)
      VH_TRY_HIP(hipMemcpyAsync(outp[q], inp[q], nb, hipMemcpyDeviceToDevice, st));
    return VH_OK;
  }
  for (int q = 0; q < 5; ++q) {
    ab.in[q] = inp[q];
    ab.out[q] = outp[q];
    ab.alt[q] = (T*)((char*)ws + q * fb);
  }
  const int cfg = pick_tile(g, tile_req);
  ab.g = g;
  ab.nlaunch = n_launches(cfg, nsub);
  ab.members = nullptr;
  ab.tiles = nullptr;
  EvpArgBlock<T>* ab_dev = (EvpArgBlock<T>*)((char*)ws + 5 * fb);
  hipLaunchKernelGGL(evp_write_args_kernel<T>, dim3(1), dim3(64), 0, st, ab_dev, ab);
  VH_TRY_HIP(hipGetLastError());
  if (nsub >= 8 && graphs_enabled()) {
    int rc = run_graph<T, double>(cfg, ab_dev, nsub, g, dtype, st);
    if (rc) return rc;
  } else {
    launch_substeps<T, double>(cfg, ab_dev, nsub, g, st);
  }
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

}  // namespace
}  // namespace vh

extern "C" int vh_evp_workspace_bytes(const vh_domain* d, size_t* bytes) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(bytes != nullptr, VH_EINVAL, "null bytes");
  const size_t eb = vh::elem_bytes(d->dtype);
  const size_t ab = d->dtype == VH_F64 ? sizeof(vh::EvpArgBlock<double>) : sizeof(vh::EvpArgBlock<float>);
  *bytes = 5 * vh::align256(vh::field_elems(g) * eb) + vh::align256(ab);
  return VH_OK;
}

extern "C" int vh_evp_tile_count(void) { return vh::kNumTiles - 1; }

extern "C" int vh_evp_tile_shape(int32_t tile, int32_t* ri, int32_t* rj, int32_t* k) {
  VH_REQUIRE(tile >= 1 && tile < vh::kNumTiles && ri && rj && k, VH_EINVAL, "tile %d", tile);
  *ri = vh::kTiles[tile].ri;
  *rj = vh::kTiles[tile].rj;
  *k = vh::kTiles[tile].k;
  return VH_OK;
}

extern "C" int vh_evp_solver(const vh_domain* d, const vh_settings* s, const vh_fields* f,
                             int32_t n_substeps, void* uIce, void* vIce, void* sigma1, void* sigma2,
                             void* sigma12, void* workspace, size_t workspace_bytes, void* stream) {
  vh::Geom g;
  int rc = vh::make_geom(d, &g);
  if (rc) return rc;
  VH_REQUIRE(s != nullptr && f != nullptr, VH_EINVAL, "null settings or fields");
  VH_REQUIRE(n_substeps >= 0, VH_EINVAL, "n_substeps=%d", n_substeps);
  VH_REQUIRE(d->tile >= 0 && d->tile < vh::kNumTiles, VH_EINVAL, "tile %d", d->tile);
  hipStream_t st = (hipStream_t)stream;
  if (d->dtype == VH_F64)
    return vh::run_evp<double>(g, s, f, n_substeps, uIce, vIce, sigma1, sigma2, sigma12, workspace,
                               workspace_bytes, d->tile, d->dtype, st);
  return vh::run_evp<float>(g, s, f, n_substeps, uIce, vIce, sigma1, sigma2, sigma12, workspace,
                            workspace_bytes, d->tile, d->dtype, st);
}
