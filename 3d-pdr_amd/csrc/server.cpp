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
// Standing launches (pdrchem_server_set_refill) replace the batches for
// host requests: each network keeps a ring of points in pinned host memory
// that the solver reads and writes in place, and a persistent launch that
// keeps taking points from it while the host appends new requests -- a
// request's points start as soon as any worker of the running launch is
// free, instead of waiting for a slot whose batch may hold a straggler.  The
// host publishes a request by copying it into the ring and advancing the
// published count; workers claim published points one at a time and write
// each point's status last, which is how the host sees it finish.  A
// launch's workers stop after an idle spell or a lifetime, so launches of
// the networks take turns on the device and every launch ends by itself;
// the host starts a successor on the lane's second stream before a launch's
// lifetime runs out, and a new launch whenever work is waiting and none
// runs.
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
#include <thread>
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

// a host request held in a standing-launch ring
struct LaneReq {
  int ticket, npts;
  unsigned int first;  // point number of its first point (ring position first & mask)
  double *ab;          // caller's array, written when the request completes
  double t_submit;
  int checked = 0;     // leading points seen finished
  int status = 0;
  bool done = false;
};

// one network's standing launches: the ring (pinned, coherent host memory,
// read and written by the solver in place), the published count and stop
// word (host memory), the device counter of points taken, two streams that
// alternate launches
struct Lane {
  bool ready = false;
  unsigned int cap = 0, mask = 0;
  char *ring = nullptr;
  size_t off_rate = 0, off_dens = 0, off_tend = 0, off_status = 0;
  unsigned int *ctl = nullptr;  // [0] published, [1] stop
  unsigned int *d_counter = nullptr;
  unsigned int head = 0, tail = 0;  // next point number to publish; oldest point still held
  std::deque<LaneReq> held;         // published requests, in ring order
  std::deque<Req> waiting;          // requests waiting for ring space
  hipStream_t st[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool launched[2] = {false, false};
  double t_launch[2] = {0.0, 0.0};
  long long launches = 0;
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
  // standing launches for host requests (0 points: batches)
  unsigned int refill_points = 0;
  double idle_s = 2e-3, life_s = 0.25;
  double tick_hz = 1e8;  // real-time clock of the device
  Lane lanes[NET_COUNT];
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

// ---------------------------------------------------------------------------
// standing launches
// ---------------------------------------------------------------------------
void lane_free(Lane &ln);

bool lane_setup(pdrchem_server *S, Lane &ln, int net) {
  if (ln.ready) return true;
  const Shape &s = kShape[net];
  const size_t cap = S->refill_points;
  ln.off_rate = align256(cap * s.nspec * 8);
  ln.off_dens = ln.off_rate + align256(cap * s.nreac * 8);
  ln.off_tend = ln.off_dens + align256(cap * 8);
  ln.off_status = ln.off_tend + align256(cap * 8);
  const size_t bytes = ln.off_status + align256(cap * 4);
  const int mixed = S->precision == PDRCHEM_MIXED ? 1 : 0;
  const size_t ws = net == NET_REDUCED ? wave_solver_workspace_bytes(net, mixed) : wg_solver_workspace_bytes(net, mixed);
  bool ok = hipHostMalloc((void **)&ln.ring, bytes, hipHostMallocCoherent) == hipSuccess &&
            hipHostMalloc((void **)&ln.ctl, 256, hipHostMallocCoherent) == hipSuccess &&
            hipMalloc(&ln.d_counter, sizeof(unsigned int)) == hipSuccess &&
            hipMemset(ln.d_counter, 0, sizeof(unsigned int)) == hipSuccess;
  for (int k = 0; ok && k < 2; k++)
    ok = hipStreamCreateWithFlags(&ln.st[k], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&ln.ev[k], hipEventDisableTiming) == hipSuccess && stream_workspace(ln.st[k], ws);
  if (!ok) {
    lane_free(ln);
    S->err = "standing launch set-up failed (pinned ring or device memory)";
    return false;
  }
  ln.ctl[0] = 0;
  ln.ctl[1] = 0;
  ln.cap = (unsigned int)cap;
  ln.mask = ln.cap - 1;
  ln.head = ln.tail = 0;
  ln.ready = true;
  return true;
}

void lane_free(Lane &ln) {
  if (ln.ctl) __atomic_store_n(&ln.ctl[1], 1u, __ATOMIC_RELEASE);
  for (int k = 0; k < 2; k++) {
    if (ln.st[k]) (void)hipStreamSynchronize(ln.st[k]);
    if (ln.ev[k]) (void)hipEventDestroy(ln.ev[k]);
    if (ln.st[k]) (void)hipStreamDestroy(ln.st[k]);
  }
  if (ln.d_counter) (void)hipFree(ln.d_counter);
  if (ln.ring) (void)hipHostFree(ln.ring);
  if (ln.ctl) (void)hipHostFree(ln.ctl);
  ln = Lane();
}

// copy n rows of `row` bytes between the ring (from point number k) and a
// flat array, in at most two pieces around the ring's end
void ring_rows(const Lane &ln, char *base, unsigned int k, size_t n, size_t row, char *flat, bool to_ring) {
  const size_t pos = k & ln.mask;
  const size_t n1 = std::min(n, (size_t)ln.cap - pos);
  char *r1 = base + pos * row;
  if (to_ring) {
    std::memcpy(r1, flat, n1 * row);
    if (n > n1) std::memcpy(base, flat + n1 * row, (n - n1) * row);
  } else {
    std::memcpy(flat, r1, n1 * row);
    if (n > n1) std::memcpy(flat + n1 * row, base, (n - n1) * row);
  }
}

// waiting requests into the ring while they fit, then one publication
void lane_admit(pdrchem_server *S, Lane &ln, int net) {
  const Shape &s = kShape[net];
  bool any = false;
  while (!ln.waiting.empty()) {
    const Req &r = ln.waiting.front();
    if ((size_t)(ln.head - ln.tail) + (size_t)r.npts > ln.cap) break;
    const unsigned int k = ln.head;
    ring_rows(ln, ln.ring, k, r.npts, (size_t)s.nspec * 8, (char *)r.ab, true);
    ring_rows(ln, ln.ring + ln.off_rate, k, r.npts, (size_t)s.nreac * 8, (char *)r.rate, true);
    ring_rows(ln, ln.ring + ln.off_dens, k, r.npts, 8, (char *)r.dens, true);
    double *tend = (double *)(ln.ring + ln.off_tend);
    int *status = (int *)(ln.ring + ln.off_status);
    for (int i = 0; i < r.npts; i++) {
      tend[(k + i) & ln.mask] = r.tend_s;
      status[(k + i) & ln.mask] = kPendingStatus;
    }
    LaneReq h;
    h.ticket = r.ticket;
    h.npts = r.npts;
    h.first = k;
    h.ab = r.ab;
    h.t_submit = r.t_submit;
    ln.held.push_back(h);
    ln.head += (unsigned int)r.npts;
    ln.waiting.pop_front();
    any = true;
  }
  if (any) __atomic_store_n(&ln.ctl[0], ln.head, __ATOMIC_RELEASE);
  (void)S;
}

// finished requests: abundances back to the callers, ring space freed in order
void lane_retire(pdrchem_server *S, Lane &ln, int net) {
  const Shape &s = kShape[net];
  const int *status = (const int *)(ln.ring + ln.off_status);
  for (LaneReq &h : ln.held) {
    if (h.done) continue;
    while (h.checked < h.npts) {
      const int st = __atomic_load_n(&status[(h.first + (unsigned int)h.checked) & ln.mask], __ATOMIC_ACQUIRE);
      if (st == kPendingStatus) break;
      if (h.status == 0) h.status = st;
      h.checked++;
    }
    if (h.checked < h.npts) continue;
    ring_rows(ln, ln.ring, h.first, h.npts, (size_t)s.nspec * 8, (char *)h.ab, false);
    h.done = true;
    S->done.push_back(Done{h.ticket, h.status, (now_s() - h.t_submit) * 1e3});
  }
  while (!ln.held.empty() && ln.held.front().done) {
    ln.tail += (unsigned int)ln.held.front().npts;
    ln.held.pop_front();
  }
}

bool lane_launch(pdrchem_server *S, Lane &ln, int net, int k) {
  const Shape &s = kShape[net];
  SolveRequest q{};
  q.abundance = (double *)ln.ring;
  q.rate = (const double *)(ln.ring + ln.off_rate);
  q.density = (const double *)(ln.ring + ln.off_dens);
  q.npts = (int)ln.cap;  // the launch fills the device
  q.nspec = s.nspec;
  q.nreac = s.nreac;
  q.rtol = S->rtol;
  q.atol = S->atol;
  q.t0 = 0.0;
  q.tend = 0.0;
  q.tend_pts = (const double *)(ln.ring + ln.off_tend);
  q.tout0 = 1.0e-4 * kYear;  // calculate_abundances.c:214
  q.nelect0 = s.nspec - 1;
  q.mxstep = solver_options().mxstep;
  q.xe_policy = solver_options().xe_policy;
  q.status = (int *)(ln.ring + ln.off_status);
  q.stats = nullptr;
  q.counter = ln.d_counter;
  q.mixed = S->precision == PDRCHEM_MIXED ? 1 : 0;
  q.published = &ln.ctl[0];
  q.stop = &ln.ctl[1];
  q.ring_mask = ln.mask;
  q.idle_ticks = (unsigned long long)(S->idle_s * S->tick_hz);
  q.life_ticks = (unsigned long long)(S->life_s * S->tick_hz);
  const int rc = net == NET_REDUCED ? launch_wave_solver(net, q, ln.st[k]) : launch_wg_solver(net, q, ln.st[k]);
  if (rc != 0 || hipEventRecord(ln.ev[k], ln.st[k]) != hipSuccess) {
    // nothing more runs on this lane: once its streams are idle no launch
    // can touch the ring, and its requests complete with the set-up failure
    // status of calculate_abundances_ (1), their arrays untouched
    __atomic_store_n(&ln.ctl[1], 1u, __ATOMIC_RELEASE);
    for (int j = 0; j < 2; j++) (void)hipStreamSynchronize(ln.st[j]);
    const double t = now_s();
    for (const LaneReq &h : ln.held)
      if (!h.done) S->done.push_back(Done{h.ticket, 1, (t - h.t_submit) * 1e3});
    for (const Req &r : ln.waiting) S->done.push_back(Done{r.ticket, 1, (t - r.t_submit) * 1e3});
    lane_free(ln);
    S->err = "standing launch failed";
    return false;
  }
  ln.launched[k] = true;
  ln.t_launch[k] = now_s();
  ln.launches++;
  return true;
}

// keep a launch running while points wait: one when none runs, and a
// successor on the other stream once the running one is half way through
// its lifetime (its workers then stop at the lifetime and hand their
// compute units over)
bool lane_pump(pdrchem_server *S, Lane &ln, int net) {
  bool work = false;
  for (const LaneReq &h : ln.held)
    if (!h.done) {
      work = true;
      break;
    }
  if (!work) return true;
  bool live[2];
  for (int k = 0; k < 2; k++) {
    if (ln.launched[k] && hipEventQuery(ln.ev[k]) == hipSuccess) ln.launched[k] = false;
    live[k] = ln.launched[k];
  }
  const double t = now_s();
  if (!live[0] && !live[1]) return lane_launch(S, ln, net, 0);
  if (live[0] != live[1]) {
    const int k = live[0] ? 0 : 1;
    if (t - ln.t_launch[k] > 0.5 * S->life_s) return lane_launch(S, ln, net, 1 - k);
  }
  return true;
}

bool lanes_step(pdrchem_server *S) {
  for (int net = 0; net < NET_COUNT; net++) {
    Lane &ln = S->lanes[net];
    if (!ln.ready) continue;
    lane_retire(S, ln, net);
    lane_admit(S, ln, net);
    if (!lane_pump(S, ln, net)) return false;
  }
  return true;
}

int lanes_outstanding(const pdrchem_server *S) {
  int n = 0;
  for (const Lane &ln : S->lanes) {
    n += (int)ln.waiting.size();
    for (const LaneReq &h : ln.held) n += h.done ? 0 : 1;
  }
  return n;
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
  if (!S->refill_points && ((size_t)npts > S->max_batch_pts || layout(net, npts).total > S->arena.cap)) {
    S->err = "request larger than the serving budget";
    return -1;
  }
  if (S->refill_points) {
    Lane &ln = S->lanes[net];
    if (!lane_setup(S, ln, net)) return -1;
    if ((unsigned int)npts > ln.cap) {
      S->err = "request larger than the standing launch's ring";
      return -1;
    }
    Req r{S->next_ticket++, net, npts, abundance, rate, density, t_end_years * kYear, now_s()};
    ln.waiting.push_back(r);
    lane_admit(S, ln, net);
    if (!lane_pump(S, ln, net)) return -1;
    return r.ticket;
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
  std::unique_lock<std::mutex> lock(S->mu);
  for (;;) {
    retire(S);
    if (!admit(S)) return -1;
    if (!lanes_step(S)) return -1;
    int busy = 0;
    for (const Slot &sl : S->slots) busy += sl.busy;
    const int lane_work = lanes_outstanding(S);
    if (!wait || (busy == 0 && S->pending.empty() && lane_work == 0))
      return busy + (int)S->pending.size() + lane_work;
    if (busy) {
      for (Slot &sl : S->slots)
        if (sl.busy) {
          (void)hipEventSynchronize(sl.done);
          break;
        }
    } else {
      // standing launches finish points one by one: poll the ring with the
      // lock released, so that other threads can submit meanwhile
      lock.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(100));
      lock.lock();
    }
  }
}

int pdrchem_server_poll(pdrchem_server *S, int max, int *tickets, int *status, double *latency_ms) {
  if (!S || max < 0) return -1;
  std::lock_guard<std::mutex> lock(S->mu);
  for (int net = 0; net < NET_COUNT; net++)
    if (S->lanes[net].ready) lane_retire(S, S->lanes[net], net);
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
  for (Lane &ln : S->lanes) lane_free(ln);
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

int pdrchem_server_set_refill(pdrchem_server *S, int ring_points, double idle_ms, double life_ms) {
  if (!S || ring_points < 0 || (ring_points & (ring_points - 1)) != 0 || ring_points > (1 << 24) ||
      (ring_points && (ring_points < 64 || !(idle_ms > 0.0) || !(life_ms > 0.0) || idle_ms > 6e4 || life_ms > 6e5)))
    return -1;
  std::lock_guard<std::mutex> lock(S->mu);
  if (S->refill_points == (unsigned int)ring_points) {
    if (ring_points) {
      S->idle_s = idle_ms * 1e-3;
      S->life_s = life_ms * 1e-3;
    }
    return 0;
  }
  for (const Lane &ln : S->lanes)
    if (!ln.waiting.empty() || !ln.held.empty()) {
      S->err = "standing launches hold requests";
      return -1;
    }
  for (Lane &ln : S->lanes) lane_free(ln);
  if (ring_points) {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) {
      S->err = "no real-time clock rate for the device";
      return -1;
    }
    S->tick_hz = 1e3 * khz;
    S->idle_s = idle_ms * 1e-3;
    S->life_s = life_ms * 1e-3;
  }
  S->refill_points = (unsigned int)ring_points;
  return 0;
}

int pdrchem_server_refill_stats(pdrchem_server *S, long long *launches, int *held_points) {
  if (!S) return -1;
  std::lock_guard<std::mutex> lock(S->mu);
  long long n = 0;
  int held = 0;
  for (const Lane &ln : S->lanes) {
    n += ln.launches;
    held += (int)(ln.head - ln.tail);
  }
  if (launches) *launches = n;
  if (held_points) *held_points = held;
  return (int)S->refill_points;
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
