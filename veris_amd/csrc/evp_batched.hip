// Batched EVP subcycle: host side of vh_evp_solver_batched and its FP64
// instantiation (built with -ffp-contract=off, like the single-domain path).
// The FP32 instantiation lives in evp_batched_fast.hip (contraction allowed).
//
// Members of uneven sizes are packed into shared arrays.  A device member
// table (offset, geometry, index of the member's first tile) lets one launch
// cover every member: a workgroup finds its member by binary search and its
// tile inside the member arithmetically, and a member's tiles form one
// contiguous run of the XCD-ordered tile sequence.  The member table and the
// argument block are written into the caller's workspace by small kernels
// that receive them by value (so no host buffer has to outlive the call);
// the substep launches are replayed from a hipGraph.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "evp_kernel.h"

namespace vh {

void evp_batched_launch_f32(int cfg, const EvpArgBlock<float>* ab, int which, int launch,
                            int n_tiles, bool so, bool plain, hipStream_t st);
void evp_batched_launch_f32_f64(int cfg, const EvpArgBlock<float>* ab, int which, int launch,
                                int n_tiles, bool so, bool plain, hipStream_t st);
void evp_batched_launch_f32_s64(int cfg, const EvpArgBlock<float, double>* ab, int which,
                                int launch, int n_tiles, bool so, bool plain, hipStream_t st);

namespace {

int member_geom(const vh_member& m, Geom* g) {
  vh_domain d{m.nx, m.ny, 2, 2, m.cyclic_x, m.fill_mode, VH_F64, 0};
  return make_geom(&d, g);
}

int n_tiles_of(int cfg, const Geom& g) {
  const int TI = evp_tile_extent_i(cfg), TJ = evp_tile_extent_j(cfg);
  return ((g.NX + TI - 1) / TI) * ((g.NY + TJ - 1) / TJ);
}

// Workspace: ping-pong state, argument block, member table (the minimum,
// `base`), then the packed constants (`packed`) and the tile tables of the
// main and remainder shapes (`total`).  Without room for the packed constants
// the launches read the full fields; without room for the tables they search
// the member table per tile.
struct Layout {
  size_t fb, ab, mem, base, packed, total;
  int rows;             // metric rows of all members (sum of padded ny)
  int n_main, n_rem;    // tiles of the main and remainder shapes
};

int batch_cfg(const vh_batch* b) { return (b->tile > 0) ? b->tile : 3; }

// T: storage of the step constants, S: of the state (ping-pong slots)
template <typename T, typename S = T>
Layout layout_of(const vh_batch* b) {
  Layout L;
  L.fb = align256((size_t)b->total_elems * sizeof(S));
  L.ab = align256(sizeof(EvpArgBlock<T, S>));
  L.mem = align256((size_t)b->n_members * sizeof(MemberDesc));
  L.base = 5 * L.fb + L.ab + L.mem;
  L.rows = 0;
  L.n_main = 0;
  L.n_rem = 0;
  for (int m = 0; m < b->n_members; ++m) {
    Geom g;
    member_geom(b->members[m], &g);
    L.rows += g.NY;
    L.n_main += n_tiles_of(batch_cfg(b), g);
    L.n_rem += n_tiles_of(kEvpRemainderCfg, g);
  }
  L.packed = L.base + evp_pack_bytes<T>((size_t)b->total_elems, (size_t)L.rows);
  L.total = L.packed + align256((size_t)(L.n_main + L.n_rem) * sizeof(TileDesc));
  return L;
}

int check_batch(const vh_batch* b) {
  VH_REQUIRE(b != nullptr && b->members != nullptr && b->n_members > 0, VH_EINVAL, "empty batch");
  VH_REQUIRE(b->dtype == VH_F64 || b->dtype == VH_F32, VH_EDTYPE, "dtype %d", b->dtype);
  VH_REQUIRE(b->tile >= 0 && b->tile < kEvpNumTiles, VH_EINVAL, "tile %d", b->tile);
  VH_REQUIRE(b->arith == 0 || ((b->arith == 1 || b->arith == 2) && b->dtype == VH_F32), VH_EINVAL,
             "arith %d (1 = FP64 arithmetic, 2 = FP64 state and arithmetic: FP32 storage only)",
             b->arith);
  std::vector<std::pair<long long, long long>> spans;
  long long tiles = 0;
  for (int m = 0; m < b->n_members; ++m) {
    Geom g;
    int rc = member_geom(b->members[m], &g);
    if (rc) return rc;
    const long long n = (long long)g.NX * g.NY;
    VH_REQUIRE(b->members[m].offset >= 0 && b->members[m].offset + n <= b->total_elems,
               VH_ESHAPE, "member %d box [%lld, %lld) outside the packed fields (%lld)", m,
               (long long)b->members[m].offset, (long long)b->members[m].offset + n,
               (long long)b->total_elems);
    spans.push_back({b->members[m].offset, b->members[m].offset + n});
    tiles += n_tiles_of(7, g);  // the shape with the most tiles
  }
  VH_REQUIRE(tiles < (1LL << 30), VH_ESHAPE, "batch too large (%lld tiles)", tiles);
  // the substep kernel addresses every packed field with 32-bit byte offsets
  const long long esz = (b->dtype == VH_F64 || b->arith == 2) ? 8 : 4;
  VH_REQUIRE(b->total_elems * esz < (1LL << 32), VH_ESHAPE,
             "batch too large: %lld elements of %lld bytes exceed 4 GiB per packed field",
             (long long)b->total_elems, esz);
  std::sort(spans.begin(), spans.end());
  for (size_t q = 1; q < spans.size(); ++q)
    VH_REQUIRE(spans[q].first >= spans[q - 1].second, VH_EINVAL, "member boxes overlap");
  return VH_OK;
}

inline void launch_batched(int cfg, int, bool so, bool plain, const EvpArgBlock<double>* ab,
                           int which, int launch, int n_tiles, hipStream_t st) {
  evp_launch_cfg<double, double, true>(cfg, ab, which, launch, n_tiles, 0, st, so, plain);
}

inline void launch_batched(int cfg, int arith, bool so, bool plain, const EvpArgBlock<float>* ab,
                           int which, int launch, int n_tiles, hipStream_t st) {
  if (arith == 1) evp_batched_launch_f32_f64(cfg, ab, which, launch, n_tiles, so, plain, st);
  else evp_batched_launch_f32(cfg, ab, which, launch, n_tiles, so, plain, st);
}

inline void launch_batched(int cfg, int, bool so, bool plain,
                           const EvpArgBlock<float, double>* ab, int which, int launch,
                           int n_tiles, hipStream_t st) {
  evp_batched_launch_f32_s64(cfg, ab, which, launch, n_tiles, so, plain, st);
}

template <typename T, typename S = T>
int run_batched(const vh_batch* b, const vh_settings* s, const vh_fields* fl, int nsub, void* uo,
                void* vo, void* s1o, void* s2o, void* s12o, void* ws, size_t wsb, hipStream_t st) {
  const Layout L = layout_of<T, S>(b);
  VH_REQUIRE(ws != nullptr && wsb >= L.base, VH_EWORKSPACE,
             "batched evp workspace: need %zu bytes, got %zu", L.total, wsb);
  const bool pack = wsb >= L.packed;
  const bool tables = wsb >= L.total;
  const void* req[kEvpMaxRequired];
  const int nreq = evp_required_fields(fl, req);
  for (int q = 0; q < nreq; ++q)
    VH_REQUIRE(req[q] != nullptr, VH_EINVAL, "evp_solver_batched: missing input field");
  const void* outs[] = {uo, vo, s1o, s2o, s12o};
  for (int q = 0; q < 5; ++q) {
    VH_REQUIRE(outs[q] != nullptr, VH_EINVAL, "evp_solver_batched: null output");
    for (int r = 0; r < nreq; ++r)
      VH_REQUIRE(outs[q] != req[r], VH_EINVAL, "evp_solver_batched: output aliases an input");
    for (int r = 0; r < q; ++r)
      VH_REQUIRE(outs[q] != outs[r], VH_EINVAL, "evp_solver_batched: two outputs alias");
  }
  if (s->noSlip == 0.0 && s->use_coastline != 0.0)
    VH_REQUIRE(fl->Fu && fl->Fv, VH_EINVAL, "lateral drag with use_coastline needs Fu, Fv");
  const int kt = (fl->k1AtC || fl->k2AtC || fl->k1AtZ || fl->k2AtZ) ? 1 : 0;
  VH_REQUIRE(!kt || (fl->k1AtC && fl->k2AtC && fl->k1AtZ && fl->k2AtZ), VH_EINVAL,
             "k1AtC/k2AtC/k1AtZ/k2AtZ must be all set or all NULL");
  const int basal = s->basalDragK2 != 0.0 ? 1 : 0;
  if (basal)
    VH_REQUIRE(fl->R_low && fl->Area && fl->hIceMean, VH_EINVAL,
               "basal drag needs R_low, Area, hIceMean");

  const S* inp[5] = {evp_ptr<S>(fl->uIce), evp_ptr<S>(fl->vIce), evp_ptr<S>(fl->sigma1),
                     evp_ptr<S>(fl->sigma2), evp_ptr<S>(fl->sigma12)};
  S* outp[5] = {(S*)uo, (S*)vo, (S*)s1o, (S*)s2o, (S*)s12o};
  if (nsub == 0) {
    for (int q = 0; q < 5; ++q)
      VH_TRY_HIP(hipMemcpyAsync(outp[q], inp[q], (size_t)b->total_elems * sizeof(S),
                                hipMemcpyDeviceToDevice, st));
    return VH_OK;
  }

  const int cfg = batch_cfg(b);
  VH_REQUIRE(s->noSlip == 0.0 || s->secondOrderBC == 0.0 || kEvpTiles[cfg].k == 1,
             VH_EUNSUPPORTED, "secondOrderBC needs a single-substep tile shape (tile %d has k=%d)",
             cfg, kEvpTiles[cfg].k);
  char* base = (char*)ws;
  EvpArgBlock<T, S>* ab_dev = (EvpArgBlock<T, S>*)(base + 5 * L.fb);
  MemberDesc* mem_dev = (MemberDesc*)(base + 5 * L.fb + L.ab);

  // member table, written by value in chunks
  int n_main = 0, n_rem = 0, roff = 0;
  long long max_cells = 0;
  for (int first = 0; first < b->n_members; first += kMembersPerWrite) {
    MemberChunk ch{};
    ch.first = first;
    ch.count = std::min(kMembersPerWrite, b->n_members - first);
    for (int q = 0; q < ch.count; ++q) {
      MemberDesc& md = ch.m[q];
      md.off = b->members[first + q].offset;
      member_geom(b->members[first + q], &md.g);
      md.tb[0] = n_main;
      md.tb[1] = n_rem;
      md.roff = roff;
      roff += md.g.NY;
      max_cells = std::max(max_cells, (long long)md.g.NX * md.g.NY);
      n_main += n_tiles_of(cfg, md.g);
      n_rem += n_tiles_of(kEvpRemainderCfg, md.g);
    }
    hipLaunchKernelGGL(evp_write_members_kernel, dim3(1), dim3(kMembersPerWrite), 0, st, mem_dev,
                       ch);
  }
  VH_TRY_HIP(hipGetLastError());

  EvpArgBlock<T, S> ab{};
  fill_evp_scalars(s, kt, basal, &ab.sc);
  fill_evp_const(fl, &ab.c);
  for (int q = 0; q < 5; ++q) {
    ab.in[q] = inp[q];
    ab.out[q] = outp[q];
    ab.alt[q] = (S*)(base + q * L.fb);
  }
  ab.g = Geom{};
  ab.nlaunch = evp_n_launches(cfg, nsub);
  ab.members = mem_dev;
  ab.n_members = b->n_members;
  if (pack) evp_pack_attach(&ab, base + L.base, (size_t)b->total_elems, L.rows);
  VH_REQUIRE(n_main == L.n_main && n_rem == L.n_rem, VH_EINVAL, "batched evp: tile count mismatch");
  if (tables) {
    ab.tiles[0] = (const TileDesc*)(base + L.packed);
    ab.tiles[1] = ab.tiles[0] + n_main;
  }
  hipLaunchKernelGGL((evp_write_args_kernel<T, S>), dim3(1), dim3(64), 0, st, ab_dev, ab);
  if (pack) evp_pack_launch(ab_dev, max_cells, b->n_members, st);
  if (tables) {
    const int tabs[2][2] = {{cfg, n_main}, {kEvpRemainderCfg, n_rem}};
    for (int w = 0; w < 2; ++w)
      if (tabs[w][1] > 0)
        hipLaunchKernelGGL((evp_tile_table_kernel<T, S>), dim3((tabs[w][1] + 255) / 256), dim3(256), 0,
                           st, ab_dev, w, evp_tile_extent_i(tabs[w][0]),
                           evp_tile_extent_j(tabs[w][0]), tabs[w][1]);
  }
  VH_TRY_HIP(hipGetLastError());
  const int K = kEvpTiles[cfg].k;
  const bool so = ab.sc.sobc != 0;
  const bool plain = evp_plain_kernels_enabled() && evp_plain_settings(ab.sc);
  auto record = [&](hipStream_t cs) {
    int l = 0;
    for (int q = 0; q < nsub / K; ++q)
      launch_batched(cfg, b->arith, so, plain, ab_dev, 0, l++, n_main, cs);
    for (int q = 0; q < nsub % K; ++q)
      launch_batched(kEvpRemainderCfg, b->arith, so, plain, ab_dev, 1, l++, n_rem, cs);
  };
  if (nsub >= 8 && graphs_enabled()) {
    char key[192];
    snprintf(key, sizeof(key), "evpb/%d/%d/%d/%d/%d/%d/%d/%d/%p", (int)sizeof(T), b->arith,
             (int)so, (int)plain, cfg, nsub, n_main, n_rem, (const void*)ab_dev);
    int rc = graph_run(key, record, st);
    if (rc) return rc;
  } else {
    record(st);
  }
  VH_TRY_HIP(hipGetLastError());
  return VH_OK;
}

}  // namespace
}  // namespace vh

extern "C" int vh_evp_batched_workspace_bytes(const vh_batch* b, size_t* bytes) {
  int rc = vh::check_batch(b);
  if (rc) return rc;
  VH_REQUIRE(bytes != nullptr, VH_EINVAL, "null bytes");
  *bytes = b->dtype == VH_F64 ? vh::layout_of<double>(b).total
           : b->arith == 2     ? vh::layout_of<float, double>(b).total
                               : vh::layout_of<float>(b).total;
  return VH_OK;
}

extern "C" int vh_evp_solver_batched(const vh_batch* b, const vh_settings* s, const vh_fields* f,
                                     int32_t n_substeps, void* uIce, void* vIce, void* sigma1,
                                     void* sigma2, void* sigma12, void* workspace,
                                     size_t workspace_bytes, void* stream) {
  int rc = vh::check_batch(b);
  if (rc) return rc;
  VH_REQUIRE(s != nullptr && f != nullptr, VH_EINVAL, "null settings or fields");
  VH_REQUIRE(n_substeps >= 0, VH_EINVAL, "n_substeps=%d", n_substeps);
  hipStream_t st = (hipStream_t)stream;
  if (b->dtype == VH_F64)
    return vh::run_batched<double>(b, s, f, n_substeps, uIce, vIce, sigma1, sigma2, sigma12,
                                   workspace, workspace_bytes, st);
  if (b->arith == 2)
    return vh::run_batched<float, double>(b, s, f, n_substeps, uIce, vIce, sigma1, sigma2,
                                          sigma12, workspace, workspace_bytes, st);
  return vh::run_batched<float>(b, s, f, n_substeps, uIce, vIce, sigma1, sigma2, sigma12,
                                workspace, workspace_bytes, st);
}
