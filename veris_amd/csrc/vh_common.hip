// Library-wide entry points (version, error text) and the domain check.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include "vh_common.h"

namespace vh {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int make_geom(const vh_domain* d, Geom* g) {
  VH_REQUIRE(d != nullptr && g != nullptr, VH_EINVAL, "null domain");
  VH_REQUIRE(d->olx == 2 && d->oly == 2, VH_ESHAPE, "olx=%d oly=%d: only 2 is supported",
             d->olx, d->oly);
  VH_REQUIRE(d->nx >= 4 && d->ny >= 4, VH_ESHAPE, "nx=%d ny=%d: need at least 4 x 4 interior cells",
             d->nx, d->ny);
  VH_REQUIRE(d->dtype == VH_F64 || d->dtype == VH_F32, VH_EDTYPE, "dtype %d", d->dtype);
  VH_REQUIRE(d->fill_mode == 0 || d->fill_mode == 1, VH_EINVAL, "fill_mode %d", d->fill_mode);
  g->nx = d->nx;
  g->ny = d->ny;
  g->ol = 2;
  g->NX = d->nx + 4;
  g->NY = d->ny + 4;
  g->fill_x = d->fill_mode == 1 ? 1 : (d->cyclic_x ? 1 : 0);
  g->fill_y = d->fill_mode == 1 ? 1 : 0;
  return VH_OK;
}

namespace {
std::mutex g_graph_mu;
std::mutex g_capture_mu;  // held for a whole capture on the per-device capture stream
std::map<std::pair<int, std::string>, hipGraphExec_t> g_graphs;
std::map<int, hipStream_t> g_capture_streams;
}  // namespace

// Keys name the workspace address, so a caller that passes ever new
// workspaces would grow the cache without end: once a device holds
// kMaxGraphsPerDevice graphs, new keys launch their sequence directly on the
// caller's stream instead (the cached graphs stay valid: another thread may
// be about to launch one).
constexpr size_t kMaxGraphsPerDevice = 256;

size_t graphs_on(int dev) {
  size_t n = 0;
  for (const auto& kv : g_graphs) n += kv.first.first == dev;
  return n;
}

int graphs_enabled() {
  static int en = -1;
  if (en < 0) {
    const char* e = getenv("VERIS_EVP_GRAPH");
    en = (e && e[0] == '0') ? 0 : 1;
  }
  return en;
}

int evp_plain_kernels_enabled() {
  const char* e = getenv("VERIS_EVP_PLAIN");
  return (e && e[0] == '0') ? 0 : 1;
}

int graph_run(const std::string& key, const std::function<void(hipStream_t)>& record,
              hipStream_t st) {
  int dev = 0;
  VH_TRY_HIP(hipGetDevice(&dev));
  hipGraphExec_t exec = nullptr;
  hipStream_t cs = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    auto it = g_graphs.find({dev, key});
    if (it != g_graphs.end()) exec = it->second;
  }
  if (!exec) {
    // one capture at a time; another thread may have captured this key meanwhile
    std::lock_guard<std::mutex> cap(g_capture_mu);
    bool full = false;
    {
      std::lock_guard<std::mutex> lk(g_graph_mu);
      auto it = g_graphs.find({dev, key});
      if (it != g_graphs.end()) exec = it->second;
      auto ct = g_capture_streams.find(dev);
      if (ct != g_capture_streams.end()) cs = ct->second;
      full = graphs_on(dev) >= kMaxGraphsPerDevice;
    }
    if (!exec && full) {
      record(st);
      VH_TRY_HIP(hipGetLastError());
      return VH_OK;
    }
    if (!exec) {
      if (!cs) {
        VH_TRY_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        std::lock_guard<std::mutex> lk(g_graph_mu);
        g_capture_streams.emplace(dev, cs);
      }
      hipGraph_t graph = nullptr;
      VH_TRY_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
      record(cs);
      VH_TRY_HIP(hipStreamEndCapture(cs, &graph));
      hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      hipError_t de = hipGraphDestroy(graph);
      VH_TRY_HIP(ie);
      VH_TRY_HIP(de);
      std::lock_guard<std::mutex> lk(g_graph_mu);
      g_graphs.emplace(std::make_pair(dev, key), exec);
    }
  }
  VH_TRY_HIP(hipGraphLaunch(exec, st));
  return VH_OK;
}

}  // namespace vh

extern "C" int vh_version(void) { return VH_ABI_VERSION; }

extern "C" int vh_last_error(char* buf, size_t len) {
  size_t n = strlen(vh::g_err);
  if (buf && len) {
    size_t c = n < len - 1 ? n : len - 1;
    memcpy(buf, vh::g_err, c);
    buf[c] = '\0';
  }
  return (int)n;
}
