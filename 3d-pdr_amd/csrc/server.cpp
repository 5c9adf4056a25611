// Request loop for serving grid-point solves on one GPU with continuous
// batching.  No counterpart in the reference (calculate_abundances_ solves
// one caller-supplied batch and returns); this is the device-resident front
// end for callers that submit many independent requests over time.
//
// A request is npts points of one network with its own end time.  Pending
// requests wait in one FIFO; whenever a launch slot is free, the oldest
// pending request's network is batched with every later pending request of
// the same network that fits (points cap and device-memory budget), packed
// into pinned staging, copied to a block of the device arena, solved by the
// persistent kernel with per-point end times, and copied back -- all on the
// slot's own stream, so batches of different requests and networks overlap
// on the device.  Requests that arrive while others run are admitted as
// slots and arena space free up.  Admission is by device-memory footprint:
// a batch starts only when its inputs and outputs fit the arena.
//
// A batch goes to the device as one HIP graph: the slot's stream captures
// the input copy, the work-queue reset (a kernel, not a memset node), the
// persistent solver launch and the result copies; the slot keeps one
// executable graph and updates it in place for the next batch
// (hipGraphExecUpdate; re-instantiated only when the update is refused), so
// a batch costs one graph launch.  The solver's stream workspace is sized
// for a whole-device launch when the server is created, so nothing inside
// the capture allocates.  PDRCHEM_SERVER_GRAPHS=0 in the environment at
// creation selects plain stream launches instead.
//
// Requests may also hold device arrays (pdrchem_server_submit_device): their
// batch gathers inputs into the arena block by device-to-device copies and
// scatters abundances (and, if asked, per-point statuses and integrator
// counters) back the same way, so state that lives in HBM never crosses PCIe;
// only the per-point end times and the statuses go through pinned staging.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pdrchem.h"
#include "gen/net_reduced.h"
#include "gen/wg_full.h"
#include "gen/wg_mynetwork.h"
#include "pdr_internal.h"

using namespace pdrchem;

namespace {

constexpr double kYear = 3.1556926e7;  // calculate_abundances.c:90

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Shape {
  int nspec, nreac;
};
const Shape kShape[NET_COUNT] = {{pdrnet_reduced::NSPEC, pdrnet_reduced::NREAC},
                                 {pdrwg_full::NSPEC, pdrwg_full::NREAC},
                                 {pdrwg_mynetwork::NSPEC, pdrwg_mynetwork::NREAC}};

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// first-fit allocator over one device allocation
struct Arena {
  char *base = nullptr;
  size_t cap = 0;
  std::map<size_t, size_t> free_;  // offset -> size, coalesced
  bool init(size_t bytes) {
    cap = align256(bytes);
    if (hipMalloc(&base, cap) != hipSuccess) return false;
    free_[0] = cap;
    return true;
  }
  long long alloc(size_t n) {
    n = align256(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      const size_t off = it->first, rest = it->second - n;
      free_.erase(it);
      if (rest) free_[off + n] = rest;
      return (long long)off;
    }
    return -1;
  }
  void release(size_t off, size_t n) {
    n = align256(n);
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Req {
  int ticket, net, npts;
  double *ab;  // caller's arrays (ab in/out), valid until the request completes
  const double *rate, *dens;
  double tend_s, t_submit;
  bool device = false;        // arrays are device memory (no host staging)
  int *d_status = nullptr;    // device requests: per-point status out (optional)
  int *d_stats = nullptr;     // device requests: per-point [8] counters out (optional)
};

struct Done {
  int ticket, status;
  double latency_ms;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  bool device = false;  // the batch's requests hold device arrays
  int net = 0;
  size_t npts = 0, off = 0, bytes = 0;
  std::vector<Req> reqs;
  char *pin = nullptr;  // pinned staging of the whole block (host requests)
  size_t pin_cap = 0;
  char *pin2 = nullptr;  // pinned end times and statuses (device requests)
  size_t pin2_cap = 0;
  hipGraphExec_t exec = nullptr;  // executable graph of the last batch, updated per batch
};

}  // namespace

struct pdrchem_server {
  std::mutex mu;
  Arena arena;
  std::vector<Slot> slots;
  std::deque<Req> pending;
  std::vector<Done> done;
  int next_ticket = 0;
  int precision = PDRCHEM_MIXED;
  size_t max_batch_pts = 65536;
  double rtol = 1e-8, atol = 1e-30;
  bool graphs = true;
  long long graph_launches = 0, graph_instantiations = 0;
  std::string err;
};

namespace {

// device (and pinned) block layout of a batch
struct Layout {
  size_t ab, rate, dens, tend, status, stats, counter, total;
};
Layout layout(int net, size_t n) {
  const Shape &s = kShape[net];
  Layout L;
  L.ab = 0;
  L.rate = L.ab + align256(n * s.nspec * 8);
  L.dens = L.rate + align256(n * s.nreac * 8);
  L.tend = L.dens + align256(n * 8);
  L.status = L.tend + align256(n * 8);
  L.stats = L.status + align256(n * 4);
  L.counter = L.stats + align256(n * 8 * 4);
  L.total = L.counter + 256;
  return L;
}

bool launch_batch(pdrchem_server *S, Slot &sl) {
  const int net = sl.net;
  const Shape &s = kShape[net];
  const size_t n = sl.npts;
  const Layout L = layout(net, n);
  // pinned staging: the whole block for host requests; end times and
  // statuses only for device requests
  char **pin = sl.device ? &sl.pin2 : &sl.pin;
  size_t *cap = sl.device ? &sl.pin2_cap : &sl.pin_cap;
  const size_t need = sl.device ? align256(n * 8) + n * 4 : L.total;
  if (*cap < need) {
    if (*pin) (void)hipHostFree(*pin);
    *pin = nullptr;
    *cap = 0;
    if (hipHostMalloc((void **)pin, need, hipHostMallocDefault) != hipSuccess) return false;
    *cap = need;
  }
  double *tend_pin = (double *)(sl.device ? sl.pin2 : sl.pin + L.tend);
  char *status_pin = sl.device ? sl.pin2 + align256(n * 8) : sl.pin + L.status;
  size_t k = 0;
  for (const Req &r : sl.reqs) {
    if (!sl.device) {
      std::memcpy(sl.pin + L.ab + k * s.nspec * 8, r.ab, (size_t)r.npts * s.nspec * 8);
      std::memcpy(sl.pin + L.rate + k * s.nreac * 8, r.rate, (size_t)r.npts * s.nreac * 8);
      std::memcpy(sl.pin + L.dens + k * 8, r.dens, (size_t)r.npts * 8);
    }
    for (int i = 0; i < r.npts; i++) tend_pin[k + i] = r.tend_s;
    k += r.npts;
  }
  char *dev = S->arena.base + sl.off;
  SolveRequest q{};
  q.abundance = (double *)(dev + L.ab);
  q.rate = (const double *)(dev + L.rate);
  q.density = (const double *)(dev + L.dens);
  q.npts = (int)n;
  q.nspec = s.nspec;
  q.nreac = s.nreac;
  q.rtol = S->rtol;
  q.atol = S->atol;
  q.t0 = 0.0;
  q.tend = sl.reqs.front().tend_s;
  q.tend_pts = (const double *)(dev + L.tend);
  q.tout0 = 1.0e-4 * kYear;  // calculate_abundances.c:214
  q.nelect0 = s.nspec - 1;   // electron last (read_species.F90:164)
  q.mxstep = solver_options().mxstep;  // calculate_abundances.c:111 (1e7 unless set)
  q.xe_policy = solver_options().xe_policy;
  q.status = (int *)(dev + L.status);
  q.stats = sl.device ? (int *)(dev + L.stats) : nullptr;
  q.counter = (unsigned int *)(dev + L.counter);
  q.max_blocks = 0;
  q.mixed = S->precision == PDRCHEM_MIXED ? 1 : 0;
  // the batch's work on the slot's stream
  auto enqueue = [&]() -> bool {
    if (sl.device) {
      // inputs gathered from the requests' device arrays (device to device)
      size_t k = 0;
      for (const Req &r : sl.reqs) {
        const size_t m = (size_t)r.npts;
        if (hipMemcpyAsync(dev + L.ab + k * s.nspec * 8, r.ab, m * s.nspec * 8, hipMemcpyDeviceToDevice,
                           sl.stream) != hipSuccess ||
            hipMemcpyAsync(dev + L.rate + k * s.nreac * 8, r.rate, m * s.nreac * 8, hipMemcpyDeviceToDevice,
                           sl.stream) != hipSuccess ||
            hipMemcpyAsync(dev + L.dens + k * 8, r.dens, m * 8, hipMemcpyDeviceToDevice, sl.stream) != hipSuccess)
          return false;
        k += m;
      }
      if (hipMemcpyAsync(dev + L.tend, tend_pin, n * 8, hipMemcpyHostToDevice, sl.stream) != hipSuccess)
        return false;
    } else if (hipMemcpyAsync(dev, sl.pin, L.status, hipMemcpyHostToDevice, sl.stream) != hipSuccess) {
      return false;
    }
    if (launch_zero_u32(q.counter, sl.stream) != 0) return false;
    const int rc = net == NET_REDUCED ? launch_wave_solver(net, q, sl.stream) : launch_wg_solver(net, q, sl.stream);
    if (rc != 0) return false;
    if (sl.device) {
      // results scattered back to the requests' device arrays
      size_t k = 0;
      for (const Req &r : sl.reqs) {
        const size_t m = (size_t)r.npts;
        if (hipMemcpyAsync(r.ab, dev + L.ab + k * s.nspec * 8, m * s.nspec * 8, hipMemcpyDeviceToDevice,
                           sl.stream) != hipSuccess)
          return false;
        if (r.d_status && hipMemcpyAsync(r.d_status, dev + L.status + k * 4, m * 4, hipMemcpyDeviceToDevice,
                                         sl.stream) != hipSuccess)
          return false;
        if (r.d_stats && hipMemcpyAsync(r.d_stats, dev + L.stats + k * 32, m * 32, hipMemcpyDeviceToDevice,
                                        sl.stream) != hipSuccess)
          return false;
        k += m;
      }
    } else if (hipMemcpyAsync(sl.pin + L.ab, dev + L.ab, (size_t)n * s.nspec * 8, hipMemcpyDeviceToHost,
                              sl.stream) != hipSuccess) {
      return false;
    }
    // statuses back to the staging block (request status for both kinds)
    return hipMemcpyAsync(status_pin, dev + L.status, n * 4, hipMemcpyDeviceToHost, sl.stream) == hipSuccess;
  };
  if (!S->graphs) {
    if (!enqueue()) return false;
    return hipEventRecord(sl.done, sl.stream) == hipSuccess;
  }
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
  const bool queued = enqueue();
  const hipError_t ec = hipStreamEndCapture(sl.stream, &g);
  if (!queued || ec != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return false;
  }
  bool ready = false;
  if (sl.exec) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res;
    ready = hipGraphExecUpdate(sl.exec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
    if (!ready) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(sl.exec);
      sl.exec = nullptr;
    }
  }
  if (!ready) {
    if (hipGraphInstantiate(&sl.exec, g, nullptr, nullptr, 0) != hipSuccess) {
      sl.exec = nullptr;
      (void)hipGraphDestroy(g);
      return false;
    }
    S->graph_instantiations++;
  }
  (void)hipGraphDestroy(g);
  if (hipGraphLaunch(sl.exec, sl.stream) != hipSuccess) return false;
  S->graph_launches++;
  return hipEventRecord(sl.done, sl.stream) == hipSuccess;
}

// completed batches: results to the callers' arrays, arena block freed
void retire(pdrchem_server *S) {
  for (Slot &sl : S->slots) {
    if (!sl.busy || hipEventQuery(sl.done) != hipSuccess) continue;
    const Shape &s = kShape[sl.net];
    const Layout L = layout(sl.net, sl.npts);
    const double t = now_s();
    const int *status_pin = (const int *)(sl.device ? sl.pin2 + align256(sl.npts * 8) : sl.pin + L.status);
    size_t k = 0;
    for (const Req &r : sl.reqs) {
      if (!sl.device) std::memcpy(r.ab, sl.pin + L.ab + k * s.nspec * 8, (size_t)r.npts * s.nspec * 8);
      const int *st = status_pin + k;
      int status = 0;
      for (int i = 0; i < r.npts && !status; i++) status = st[i];
      S->done.push_back(Done{r.ticket, status, (t - r.t_submit) * 1e3});
      k += r.npts;
    }
    S->arena.release(sl.off, sl.bytes);
    sl.reqs.clear();
    sl.busy = false;
  }
}

// start batches on free slots while pending requests fit
bool admit(pdrchem_server *S) {
  for (Slot &sl : S->slots) {
    if (sl.busy || S->pending.empty()) continue;
    const int net = S->pending.front().net;
    const bool device = S->pending.front().device;
    std::vector<Req> take;
    size_t n = 0;
    for (auto it = S->pending.begin(); it != S->pending.end();) {
      // the oldest request always goes, even if the cap was lowered below its
      // size after it was queued (it then forms a batch of its own)
      if (it->net == net && it->device == device && (take.empty() || n + it->npts <= S->max_batch_pts)) {
        n += it->npts;
        take.push_back(*it);
        it = S->pending.erase(it);
      } else {
        ++it;
      }
    }
    // shrink the batch from the back until it fits the arena
    long long off = -1;
    size_t bytes = 0;
    while (!take.empty()) {
      bytes = layout(net, n).total;
      off = S->arena.alloc(bytes);
      if (off >= 0) break;
      n -= take.back().npts;
      // back into the queue at its arrival position (tickets increase)
      auto pos = S->pending.begin();
      while (pos != S->pending.end() && pos->ticket < take.back().ticket) ++pos;
      S->pending.insert(pos, take.back());
      take.pop_back();
    }
    if (take.empty()) return true;  // the oldest request waits for arena space
    sl.net = net;
    sl.device = device;
    sl.npts = n;
    sl.off = (size_t)off;
    sl.bytes = bytes;
    sl.reqs = std::move(take);
    sl.busy = true;
    if (!launch_batch(S, sl)) {
      // nothing of this batch runs: its requests complete with the set-up
      // failure status of calculate_abundances_ (1)
      (void)hipStreamSynchronize(sl.stream);
      const double t = now_s();
      for (const Req &r : sl.reqs) S->done.push_back(Done{r.ticket, 1, (t - r.t_submit) * 1e3});
      S->arena.release(sl.off, sl.bytes);
      sl.reqs.clear();
      sl.busy = false;
      S->err = "launch of a batch failed";
      return false;
    }
  }
  return true;
}

}  // namespace

extern "C" {

pdrchem_server *pdrchem_server_create(size_t device_bytes, int nslots, int precision) {
  pdrchem_server *S = new pdrchem_server();
  if (nslots < 1) nslots = 1;
  if (!S->arena.init(device_bytes ? device_bytes : (size_t)1 << 30)) {
    delete S;
    return nullptr;
  }
  S->slots.resize(nslots);
  for (Slot &sl : S->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
      pdrchem_server_destroy(S);
      return nullptr;
    }
  }
  S->precision = precision == PDRCHEM_FP64 ? PDRCHEM_FP64 : PDRCHEM_MIXED;
  S->rtol = chemistry_module_mp_relative_abundance_tolerance_;
  S->atol = chemistry_module_mp_absolute_abundance_tolerance_;
  const char *e = std::getenv("PDRCHEM_SERVER_GRAPHS");
  S->graphs = !(e && std::strcmp(e, "0") == 0);
  // solver workspace of every slot sized for a whole-device launch of any
  // network, so that no launch inside a capture allocates
  const int mixed = S->precision == PDRCHEM_MIXED ? 1 : 0;
  size_t ws = 0;
  for (int net = 0; net < NET_COUNT; net++)
    ws = std::max(ws, std::max(wave_solver_workspace_bytes(net, mixed), wg_solver_workspace_bytes(net, mixed)));
  for (Slot &sl : S->slots) {
    if (!stream_workspace(sl.stream, ws)) {
      pdrchem_server_destroy(S);
      return nullptr;
    }
  }
  return S;
}

int pdrchem_server_submit(pdrchem_server *S, int net, int npts, double *abundance, const double *rate,
                          const double *density, double t_end_years) {
  if (!S || net < 0 || net >= NET_COUNT || npts <= 0 || !abundance || !rate || !density) return -1;
  std::lock_guard<std::mutex> lock(S->mu);
  if ((size_t)npts > S->max_batch_pts || layout(net, npts).total > S->arena.cap) {
    S->err = "request larger than the serving budget";
    return -1;
  }
  Req r{S->next_ticket++, net, npts, abundance, rate, density, t_end_years * kYear, now_s()};
  S->pending.push_back(r);
  if (!admit(S)) return -1;
  return r.ticket;
}

int pdrchem_server_submit_device(pdrchem_server *S, int net, int npts, double *d_abundance, const double *d_rate,
                                 const double *d_density, double t_end_years, int *d_status, int *d_stats) {
  if (!S || net < 0 || net >= NET_COUNT || npts <= 0 || !d_abundance || !d_rate || !d_density) return -1;
  std::lock_guard<std::mutex> lock(S->mu);
  if ((size_t)npts > S->max_batch_pts || layout(net, npts).total > S->arena.cap) {
    S->err = "request larger than the serving budget";
    return -1;
  }
  Req r{S->next_ticket++, net, npts, d_abundance, d_rate, d_density, t_end_years * kYear, now_s()};
  r.device = true;
  r.d_status = d_status;
  r.d_stats = d_stats;
  S->pending.push_back(r);
  if (!admit(S)) return -1;
  return r.ticket;
}

int pdrchem_server_max_batch(pdrchem_server *S) { return S ? (int)S->max_batch_pts : -1; }

int pdrchem_server_set_max_batch(pdrchem_server *S, int npts) {
  if (!S || npts < 1) return -1;
  std::lock_guard<std::mutex> lock(S->mu);
  S->max_batch_pts = (size_t)npts;
  return 0;
}

int pdrchem_server_progress(pdrchem_server *S, int wait) {
  if (!S) return -1;
  std::lock_guard<std::mutex> lock(S->mu);
  for (;;) {
    retire(S);
    if (!admit(S)) return -1;
    int busy = 0;
    for (const Slot &sl : S->slots) busy += sl.busy;
    if (!wait || (busy == 0 && S->pending.empty())) return busy + (int)S->pending.size();
    for (Slot &sl : S->slots)
      if (sl.busy) {
        (void)hipEventSynchronize(sl.done);
        break;
      }
  }
}

int pdrchem_server_poll(pdrchem_server *S, int max, int *tickets, int *status, double *latency_ms) {
  if (!S || max < 0) return -1;
  std::lock_guard<std::mutex> lock(S->mu);
  const int n = (int)std::min<size_t>((size_t)max, S->done.size());
  for (int i = 0; i < n; i++) {
    if (tickets) tickets[i] = S->done[i].ticket;
    if (status) status[i] = S->done[i].status;
    if (latency_ms) latency_ms[i] = S->done[i].latency_ms;
  }
  S->done.erase(S->done.begin(), S->done.begin() + n);
  return n;
}

void pdrchem_server_destroy(pdrchem_server *S) {
  if (!S) return;
  for (Slot &sl : S->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.pin) (void)hipHostFree(sl.pin);
    if (sl.pin2) (void)hipHostFree(sl.pin2);
  }
  if (S->arena.base) (void)hipFree(S->arena.base);
  delete S;
}

int pdrchem_server_graph_stats(pdrchem_server *S, long long *launches, long long *instantiations) {
  if (!S) return -1;
  std::lock_guard<std::mutex> lock(S->mu);
  if (launches) *launches = S->graph_launches;
  if (instantiations) *instantiations = S->graph_instantiations;
  return S->graphs ? 1 : 0;
}

const char *pdrchem_server_last_error(pdrchem_server *S) { return S ? S->err.c_str() : "null server"; }

}  // extern "C"
