// Engine host code: device timetable image, device query state, the
// round-step scheduler and the C ABI entry points that use the GPU.
//
// Scheduling model: a step runs one RAPTOR round (raptor.h:218-387) for
// every running slot, each at its own round index k.  A slot whose
// start-time epoch ends (its break or end_k) hands its destination round
// times to the host, which updates the query's Pareto set; at the next
// step the slot starts its next epoch (next_start_time + add_start,
// raptor.h:179-203) or, when the query is done, takes the next queued
// query (iteration-level continuous batching: slots never wait for other
// slots' epochs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <exception>
#include <cstring>
#include <deque>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "capi_util.h"
#include "handles.h"
#include "device.h"
#include "journey_legs.h"
#include "engine_access.h"
#include "kernels.h"
#include "search_host.h"
#include "timetable.h"


namespace nigiri_amd {

#define HIP_CHECK(expr)                                                       \
  do {                                                                        \
    hipError_t const err_ = (expr);                                           \
    if (err_ != hipSuccess) {                                                 \
      throw status_error{NIGIRI_AMD_EDEVICE,                                  \
                         std::string{#expr ": "} + hipGetErrorString(err_)}; \
    }                                                                         \
  } while (false)

struct dev_alloc {
  std::vector<void*> ptrs;
  std::uint64_t bytes = 0U;
  template <typename T>
  T* alloc(std::size_t n) {
    if (n == 0U) n = 1U;
    void* p = nullptr;
    auto const err = hipMalloc(&p, n * sizeof(T));
    if (err != hipSuccess) {
      throw status_error{NIGIRI_AMD_ENOMEM, std::string{"hipMalloc: "} + hipGetErrorString(err)};
    }
    ptrs.push_back(p);
    bytes += n * sizeof(T);
    return static_cast<T*>(p);
  }
  template <typename T>
  T* upload(std::vector<T> const& v, hipStream_t s) {
    auto* p = alloc<T>(v.size());
    if (!v.empty()) HIP_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return p;
  }
  ~dev_alloc() {
    for (auto* p : ptrs) (void)hipFree(p);
  }
};

template <typename T>
struct pinned {
  T* p = nullptr;
  std::size_t n = 0U;
  void reserve(std::size_t m) {
    if (m <= n) return;
    if (p != nullptr) (void)hipHostFree(p);
    p = nullptr;
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), std::max<std::size_t>(m, 1U) * sizeof(T),
                            hipHostMallocDefault));
    n = m;
  }
  ~pinned() {
    if (p != nullptr) (void)hipHostFree(p);
  }
};

}  // namespace nigiri_amd

using namespace nigiri_amd;

struct nigiri_amd_image {
  timetable const* tt = nullptr;
  int device = 0;
  dev_image d{};
  dev_alloc mem;
  // nigiri_amd_lower_bounds: the upward search's labels and the selection
  // of one row, grown on demand (calls serialised by the mutex)
  std::mutex lb_mu;
  std::uint32_t* d_lb_sel = nullptr;  // [1]
  std::uint32_t* d_lb_seed_idx = nullptr;
  std::uint16_t* d_lb_seed_min = nullptr;
  std::size_t lb_seed_cap = 0U;
  // fastest direct connections (nigiri_amd_fastest_direct and the batched
  // search): queries, their node sets and results, grown on demand (calls
  // serialised by the mutex)
  std::mutex fd_mu;
  void* d_fd = nullptr;
  std::size_t fd_cap = 0U;
  ~nigiri_amd_image() {
    if (d_lb_seed_idx != nullptr) (void)hipFree(d_lb_seed_idx);
    if (d_lb_seed_min != nullptr) (void)hipFree(d_lb_seed_min);
    if (d_fd != nullptr) (void)hipFree(d_fd);
  }
};

// host threads that write the legs of a step's journeys (with the calling
// thread); the lower-bound workers and query preparation use others
constexpr unsigned kLegWorkers = 9U;
// PHAST of the batched engine: a group with this many new lanes is swept in
// place; up to kPhastRows new slots of other groups are swept per pass as
// rows of the engine's PHAST buffer
constexpr std::uint32_t kPhastDenseLanes = 8U;
constexpr std::uint32_t kPhastRows = 512U;

// Fork-join pool of host threads for the per-journey work of a step (the
// legs of reconstructed journeys).  parallel_for runs f(0 .. n-1) on the
// workers and the calling thread and returns when all calls are done; the
// first exception is rethrown in the caller.
class host_pool {
 public:
  explicit host_pool(unsigned n) {
    for (unsigned i = 0; i != n; ++i) workers_.emplace_back([this]() { work(); });
  }
  ~host_pool() {
    {
      std::lock_guard<std::mutex> lk{mu_};
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }
  void parallel_for(std::size_t n, std::function<void(std::size_t)> f) {
    start(n, std::move(f));
    wait();
  }
  // starts f(0 .. n-1) on the workers; wait() joins (the caller helps)
  void start(std::size_t n, std::function<void(std::size_t)> f) {
    wait();
    std::lock_guard<std::mutex> lk{mu_};
    job_ = std::move(f);
    n_ = n;
    next_.store(0U);
    busy_ = workers_.size();
    error_ = nullptr;
    active_ = true;
    ++gen_;
    cv_.notify_all();
  }
  void wait() {
    if (!active_) return;
    drain();
    std::unique_lock<std::mutex> lk{mu_};
    done_cv_.wait(lk, [&]() { return busy_ == 0U; });
    active_ = false;
    job_ = nullptr;
    if (error_) {
      auto e = error_;
      error_ = nullptr;
      std::rethrow_exception(e);
    }
  }

 private:
  void drain() {
    for (auto i = next_.fetch_add(1U); i < n_; i = next_.fetch_add(1U)) {
      try {
        job_(i);
      } catch (...) {
        std::lock_guard<std::mutex> lk{mu_};
        if (!error_) error_ = std::current_exception();
      }
    }
  }
  void work() {
    std::uint64_t seen = 0U;
    while (true) {
      {
        std::unique_lock<std::mutex> lk{mu_};
        cv_.wait(lk, [&]() { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      drain();
      std::lock_guard<std::mutex> lk{mu_};
      if (--busy_ == 0U) done_cv_.notify_all();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::function<void(std::size_t)> job_;
  bool active_ = false;  // a job was started and not yet joined (caller thread only)
  std::size_t n_ = 0U;
  std::atomic<std::size_t> next_{0U};
  std::size_t busy_ = 0U;
  std::uint64_t gen_ = 0U;
  bool stop_ = false;
  std::exception_ptr error_;
};

struct nigiri_amd_engine {
  nigiri_amd_image* img = nullptr;
  std::uint32_t W = 64U, G = 1U, K = 16U, S = 64U;
  hipStream_t stream = nullptr;
  dev_alloc mem;
  dev_state st{};
  // host mirrors of per-slot parameters
  std::vector<slot_param> h_param;
  std::vector<std::uint32_t> h_dest, h_dist_loc;
  std::vector<std::uint16_t> h_dist_min;
  std::vector<std::vector<std::uint32_t>> slot_dest_bits;  // dest mask bits set per slot
  // device scratch for epoch setup
  std::uint32_t* d_reset_groups = nullptr;
  std::uint64_t* d_rt_mask = nullptr;
  std::uint64_t* d_bt_mask = nullptr;
  std::uint8_t* d_flags = nullptr;
  std::int16_t* d_worst = nullptr;
  std::uint32_t* d_sweep_groups = nullptr;
  std::uint32_t* d_slot_list = nullptr;  // [S] slots whose epoch ended
  // lower bounds of admitted slots (PHAST): selected groups / rows, their
  // lane masks, and the upward searches' labels (growable)
  std::uint16_t* d_lb_row = nullptr;   // [L] lower bounds given by the caller
  std::uint32_t* d_ph_sel = nullptr;   // [S + G] groups and rows swept
  std::uint32_t* d_ph_slots = nullptr; // [S] batched engine: the rows' slots
  std::uint64_t* d_ph_masks = nullptr; // [G] batched engine: new lanes of the groups swept in place
  std::uint16_t* d_ph_rows = nullptr;  // batched engine: [min(S, kPhastRows)][L] PHAST rows
  std::uint32_t* d_ph_seed_idx = nullptr;
  std::uint16_t* d_ph_seed_min = nullptr;
  std::size_t ph_seed_cap = 0U;
  // growable device buffers for starts / dest ops
  std::uint32_t* d_start_slot = nullptr;
  std::uint32_t* d_start_loc = nullptr;
  std::int16_t* d_start_delta = nullptr;
  std::size_t start_cap = 0U;
  std::uint64_t* d_dest_ops = nullptr;
  std::uint8_t* d_dest_set = nullptr;
  std::size_t dest_ops_cap = 0U;
  // journey reconstruction: jobs, egress offsets of intermodal jobs, records
  rc_job* d_rc_jobs = nullptr;
  rc_out* d_rc_out = nullptr;
  std::size_t rc_cap = 0U;
  std::uint32_t* d_rc_off_loc = nullptr;
  std::int16_t* d_rc_off_dur = nullptr;
  std::size_t rc_off_cap = 0U;
  std::vector<rc_out> h_rc_out;
  std::unique_ptr<host_pool> pool;  // writes the legs of a step's journeys
  std::vector<void*> grow_ptrs;
  // pinned host staging
  pinned<std::uint64_t> h_active;
  pinned<std::int16_t> h_rows;
  pinned<std::uint32_t> h_counts;
  // algo-level (single slot) pending starts
  std::vector<std::vector<std::pair<std::uint32_t, std::int32_t>>> pending_starts;
  std::vector<std::uint8_t> algo_new;    // reset_arrivals requested
  std::vector<std::uint8_t> algo_fresh;  // next_start_time requested
  std::vector<std::uint8_t> algo_dir;
  // batched engine, per slot: direction (0 fwd, 1 bwd) whose invalid value
  // fills every element of the slot's state that the validity bitmaps do
  // not list; -1 until a dense reset (first use, direction change, run ended
  // by an error)
  std::vector<std::int8_t> slot_dir;
  std::uint64_t* d_full_mask = nullptr;
  // configuration
  double admission_wait_ms = 2.0;
  std::uint32_t grid_waves = 8192U;
  // rounds replayed from captured hipGraphs (off: kernels launched one by one)
  bool use_graphs = true;
  // launch-shape knobs (nigiri_amd_engine_autotune chooses them)
  round_tune tune{1U, 16U, 64U, 64U, 17U};
  // serving: admit new queries only when no slot is busy (whole batches)
  bool admit_when_idle = false;
  // device state budget of the admitted queries (0: every slot's full state)
  std::uint64_t state_budget = 0U;
  round_graphs* graphs = round_graphs_create();
  // instrumentation
  bool profile = false;
  std::vector<hipEvent_t> events;  // 4 per round of one epoch
  nigiri_engine_stats stats{};

  ~nigiri_amd_engine() {
    // the legs pool's threads are joined before the stream and the device
    // buffers they may have used go away; then the stream drains
    pool.reset();
    if (stream != nullptr) (void)hipStreamSynchronize(stream);
    round_graphs_destroy(graphs);
    for (auto ev : events) (void)hipEventDestroy(ev);
    for (auto* p : grow_ptrs) (void)hipFree(p);
    if (stream != nullptr) (void)hipStreamDestroy(stream);
  }

  template <typename T>
  T* grow(T* old, std::size_t n) {
    if (old != nullptr) {
      grow_ptrs.erase(std::remove(grow_ptrs.begin(), grow_ptrs.end(), static_cast<void*>(old)), grow_ptrs.end());
      (void)hipFree(old);
    }
    void* p = nullptr;
    auto const err = hipMalloc(&p, std::max<std::size_t>(n, 1U) * sizeof(T));
    if (err != hipSuccess) throw status_error{NIGIRI_AMD_ENOMEM, "hipMalloc (scratch)"};
    grow_ptrs.push_back(p);
    return static_cast<T*>(p);
  }
};

namespace nigiri_amd {

namespace {

// Work order of the batched route phase: routes sorted by the Morton code of
// the mean position of their stops (normalised to the bounding box of the
// locations), so that neighbouring entries of the route scan's work list
// visit one neighbourhood and read the same state lines while they are in
// the caches.  Without coordinates routes keep their index order.  The order
// changes no result: the route phase's updates are order-independent.
std::vector<std::uint32_t> route_work_order(timetable const& tt) {
  auto const R = tt.n_routes();
  std::vector<std::uint32_t> order(R);
  for (std::uint32_t r = 0U; r != R; ++r) order[r] = r;
  if (tt.coord_x.size() != tt.n_locations || tt.coord_y.size() != tt.n_locations) return order;
  auto x0 = std::numeric_limits<double>::infinity(), y0 = x0, x1 = -x0, y1 = -x0;
  for (auto l = kNSpecialStations; l < tt.n_locations; ++l) {
    x0 = std::min(x0, static_cast<double>(tt.coord_x[l]));
    x1 = std::max(x1, static_cast<double>(tt.coord_x[l]));
    y0 = std::min(y0, static_cast<double>(tt.coord_y[l]));
    y1 = std::max(y1, static_cast<double>(tt.coord_y[l]));
  }
  if (!(x1 > x0) || !(y1 > y0)) return order;
  auto const cell = [](double v, double lo, double hi) {
    return static_cast<std::uint32_t>(std::clamp((v - lo) / (hi - lo), 0.0, 1.0) * 65535.0);
  };
  std::vector<std::uint64_t> key(R, 0U);
  for (std::uint32_t r = 0U; r != R; ++r) {
    auto const s0 = tt.route_stop_offsets[r], s1 = tt.route_stop_offsets[r + 1U];
    if (s1 == s0) continue;
    double mx = 0.0, my = 0.0;
    for (auto i = s0; i != s1; ++i) {
      auto const l = tt.route_stops[i] & kStopLocMask;
      mx += tt.coord_x[l];
      my += tt.coord_y[l];
    }
    auto const cx = cell(mx / (s1 - s0), x0, x1), cy = cell(my / (s1 - s0), y0, y1);
    for (int b = 0; b != 16; ++b) {
      key[r] |= static_cast<std::uint64_t>((cx >> b) & 1U) << (2 * b);
      key[r] |= static_cast<std::uint64_t>((cy >> b) & 1U) << (2 * b + 1);
    }
  }
  std::stable_sort(order.begin(), order.end(), [&](std::uint32_t a, std::uint32_t b) { return key[a] < key[b]; });
  return order;
}

void build_image(nigiri_amd_image& img, timetable const& tt, hipStream_t s) {
  auto& d = img.d;
  auto& m = img.mem;
  d.L = tt.n_locations;
  d.R = tt.n_routes();
  d.T = tt.n_transports();
  d.NB = tt.n_bitfields();
  d.route_stop_offsets = m.upload(tt.route_stop_offsets, s);
  d.route_stops = m.upload(tt.route_stops, s);
  d.route_transport_from = m.upload(tt.route_transport_from, s);
  d.route_stop_time_from = m.upload(tt.route_stop_time_from, s);
  d.event_times = m.upload(tt.event_times, s);
  d.transport_traffic_days = m.upload(tt.transport_traffic_days, s);
  d.route_traffic_days = m.upload(tt.route_traffic_days, s);
  d.bitfields = m.upload(tt.bitfields, s);
  d.location_route_offsets = m.upload(tt.location_route_offsets, s);
  d.location_routes = m.upload(tt.location_routes, s);
  {
    auto const order = route_work_order(tt);
    std::vector<std::uint32_t> rank(order.size()), lr_rank(tt.location_routes.size());
    for (std::uint32_t i = 0U; i != order.size(); ++i) rank[order[i]] = i;
    for (std::size_t j = 0U; j != lr_rank.size(); ++j) lr_rank[j] = rank[tt.location_routes[j]];
    d.rank_route = m.upload(order, s);
    d.lr_rank = m.upload(lr_rank, s);
    HIP_CHECK(hipStreamSynchronize(s));  // the host vectors go out of scope
  }
  d.transfer_time = m.upload(tt.transfer_time, s);
  d.fp_out_offsets = m.upload(tt.fp_out_offsets, s);
  d.fp_out = m.upload(tt.fp_out, s);
  d.fp_in_offsets = m.upload(tt.fp_in_offsets, s);
  d.fp_in = m.upload(tt.fp_in, s);
  // edge maps between the two CSR directions: footpaths_in is the exact
  // transpose of footpaths_out, so sorting both edge lists by (target,
  // source, duration) pairs every out-edge with its in-edge
  struct key {
    std::uint32_t t, src, dur, idx;
  };
  std::vector<key> outs, ins;
  outs.reserve(tt.fp_out.size());
  ins.reserve(tt.fp_in.size());
  for (auto l = 0U; l != tt.n_locations; ++l) {
    for (auto e = tt.fp_out_offsets[l]; e != tt.fp_out_offsets[l + 1U]; ++e) {
      outs.push_back({fp_target(tt.fp_out[e]), l, fp_duration(tt.fp_out[e]), e});
    }
    for (auto f = tt.fp_in_offsets[l]; f != tt.fp_in_offsets[l + 1U]; ++f) {
      ins.push_back({l, fp_target(tt.fp_in[f]), fp_duration(tt.fp_in[f]), f});
    }
  }
  auto const by_key = [](key const& a, key const& b) {
    return std::tie(a.t, a.src, a.dur, a.idx) < std::tie(b.t, b.src, b.dur, b.idx);
  };
  std::sort(outs.begin(), outs.end(), by_key);
  std::sort(ins.begin(), ins.end(), by_key);
  if (outs.size() != ins.size()) {
    throw status_error{NIGIRI_AMD_EINVAL, "footpaths_in is not the transpose of footpaths_out"};
  }
  std::vector<std::uint32_t> in_to_out(ins.size()), out_to_in(outs.size());
  for (std::size_t i = 0; i != outs.size(); ++i) {
    if (outs[i].t != ins[i].t || outs[i].src != ins[i].src || outs[i].dur != ins[i].dur) {
      throw status_error{NIGIRI_AMD_EINVAL, "footpaths_in is not the transpose of footpaths_out"};
    }
    out_to_in[outs[i].idx] = ins[i].idx;
    in_to_out[ins[i].idx] = outs[i].idx;
  }
  d.fp_in_to_out = m.upload(in_to_out, s);
  d.fp_out_to_in = m.upload(out_to_in, s);
  // (route, stop) visits of every location, in location_routes order
  {
    std::vector<std::uint32_t> off{0U}, route;
    std::vector<std::uint16_t> stop;
    for (auto l = 0U; l != tt.n_locations; ++l) {
      for (auto j = tt.location_route_offsets[l]; j != tt.location_route_offsets[l + 1U]; ++j) {
        auto const r = tt.location_routes[j];
        auto const s0 = tt.route_stop_offsets[r], n = tt.route_stop_offsets[r + 1U] - s0;
        if (n > 0xFFFFU) throw status_error{NIGIRI_AMD_EUNSUPPORTED, "route with more than 65535 stops"};
        for (auto i = 0U; i != n; ++i) {
          if ((tt.route_stops[s0 + i] & kStopLocMask) != l) continue;
          route.push_back(r);
          stop.push_back(static_cast<std::uint16_t>(i));
        }
      }
      off.push_back(static_cast<std::uint32_t>(route.size()));
    }
    d.lrs_offsets = m.upload(off, s);
    d.lrs_route = m.upload(route, s);
    d.lrs_stop = m.upload(stop, s);
    // the uploads are asynchronous: keep the host vectors until they are done
    HIP_CHECK(hipStreamSynchronize(s));
  }
  d.lb_fwd_offsets = m.upload(tt.lb_fwd_offsets, s);
  d.lb_fwd = m.upload(tt.lb_fwd, s);
  d.lb_bwd_offsets = m.upload(tt.lb_bwd_offsets, s);
  d.lb_bwd = m.upload(tt.lb_bwd, s);
  // PHAST sweeps of the lower-bound graph's contraction hierarchy
  for (int dir = 0; dir != 2; ++dir) {
    auto const& sw = dir == 0 ? tt.ch.sweep_fwd : tt.ch.sweep_bwd;
    auto& ds = d.sweep[dir];
    ds.n_nodes = static_cast<std::uint32_t>(sw.order.size());
    ds.order = m.upload(sw.order, s);
    ds.edge_off = m.upload(sw.edge_off, s);
    ds.src = m.upload(sw.src, s);
    ds.w = m.upload(sw.w, s);
    ds.level_off = m.upload(sw.level_off, s);
  }
  HIP_CHECK(hipStreamSynchronize(s));  // host vectors above go out of scope
}

std::int16_t clamp16(long long v) {
  return static_cast<std::int16_t>(std::clamp<long long>(v, SHRT_MIN, SHRT_MAX));
}

// ---------------------------------------------------------------------------
// batch scheduler

struct slot_info {
  std::int32_t q = -1;
  std::uint32_t epoch = 0U;
  bool running = false;
  bool is_new = false;
};

static_assert(kMaxQueryDests == kMaxDestBits && kMaxQueryEgress == kMaxDist &&
                  kMaxQueryDestEvents == kMaxDestEvents,
              "query limits of the host preparation and the device state agree");

// Runs the device journey reconstruction of `jobs` on the engine's stream;
// their records are in e.h_rc_out afterwards.
void run_reconstruct(nigiri_amd_engine& e, bool fwd, std::vector<rc_job> const& jobs,
                     std::vector<std::uint32_t> const& off_loc,
                     std::vector<std::int16_t> const& off_dur) {
  auto const stream = e.stream;
  for (auto const& j : jobs) {
    if (j.slot >= e.S || j.dest >= e.st.L || j.transfers + 2U > e.K ||
        static_cast<std::size_t>(j.off_begin) + j.n_off > off_loc.size()) {
      throw status_error{NIGIRI_AMD_EINVAL, "reconstruct: job out of range"};
    }
  }
  if (jobs.size() > e.rc_cap) {
    e.rc_cap = jobs.size() * 2U;
    e.d_rc_jobs = e.grow(e.d_rc_jobs, e.rc_cap);
    e.d_rc_out = e.grow(e.d_rc_out, e.rc_cap);
  }
  if (off_loc.size() > e.rc_off_cap) {
    e.rc_off_cap = off_loc.size() * 2U;
    e.d_rc_off_loc = e.grow(e.d_rc_off_loc, e.rc_off_cap);
    e.d_rc_off_dur = e.grow(e.d_rc_off_dur, e.rc_off_cap);
  }
  HIP_CHECK(hipMemcpyAsync(e.d_rc_jobs, jobs.data(), sizeof(rc_job) * jobs.size(), hipMemcpyHostToDevice, stream));
  if (!off_loc.empty()) {
    HIP_CHECK(hipMemcpyAsync(e.d_rc_off_loc, off_loc.data(), 4U * off_loc.size(), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(e.d_rc_off_dur, off_dur.data(), 2U * off_dur.size(), hipMemcpyHostToDevice, stream));
  }
  rc_args a{};
  a.im = e.img->d;
  a.st = e.st;
  a.jobs = e.d_rc_jobs;
  a.n_jobs = static_cast<std::uint32_t>(jobs.size());
  a.off_loc = e.d_rc_off_loc;
  a.off_dur = e.d_rc_off_dur;
  a.out = e.d_rc_out;
  launch_reconstruct(fwd, e.W, a, stream);
  HIP_CHECK(hipGetLastError());
  e.h_rc_out.resize(jobs.size());
  HIP_CHECK(hipMemcpyAsync(e.h_rc_out.data(), e.d_rc_out, sizeof(rc_out) * jobs.size(), hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
}

// Batched engine: lanes among `resetting` whose state elements outside the
// validity bitmaps do not hold this direction's invalid value (first use,
// the other direction before, or a run ended by an error) get a dense reset
// (per group: a lane mask); the plain engine's resets are always dense.
std::vector<std::uint64_t> dense_resets(nigiri_amd_engine& e, std::vector<std::uint32_t> const& resetting, bool fwd) {
  std::vector<std::uint64_t> full(e.G, 0U);
  if (e.W != 64U) return full;
  auto const dir = static_cast<std::int8_t>(fwd ? 0 : 1);
  for (auto const s : resetting) {
    if (e.slot_dir[s] != dir) {
      full[s / 64U] |= 1ULL << (s % 64U);
      e.slot_dir[s] = dir;
    }
  }
  return full;
}

struct runner {
  nigiri_amd_engine& e;
  timetable const& tt;
  bool fwd;
  std::vector<prepared_query>& pq;
  std::function<std::int32_t()> next;            // next query id or -1
  std::function<void(std::int32_t)> on_done;      // query finished
  std::function<void(std::int32_t)> on_admit;     // query takes a slot
  // query refused (status code, message); it is not handed to on_done
  std::function<void(std::int32_t, int, std::string const&)> on_fail;
  // admission by device state memory: what the admitted queries' rows of
  // round_times and their best / tmp / lb / mark columns take, against the
  // budget; a query that does not fit now waits in `held`
  std::uint64_t used_bytes = 0U;
  std::int32_t held = -1;
  std::vector<slot_info> slots;
  std::uint64_t rounds = 0U, epochs = 0U;
  // legs being written on the host pool, and the queries waiting for them
  bool legs_pending = false;
  std::vector<std::pair<std::int32_t, std::size_t>> rc_owner;
  std::vector<std::int32_t> completing;

  runner(nigiri_amd_engine& eng, timetable const& t, bool f, std::vector<prepared_query>& p)
      : e{eng}, tt{t}, fwd{f}, pq{p}, slots(eng.S) {}
  runner(runner const&) = delete;
  runner& operator=(runner const&) = delete;
  ~runner() {
    // the pool's job refers to this runner: join it (errors were reported
    // through the exception that ends the runner early)
    try {
      e.pool->wait();
    } catch (...) {
    }
    // a run ended by an error may leave marks that were never folded into
    // the validity bitmaps: the next use of every group starts dense
    if (std::uncaught_exceptions() > 0) std::fill(e.slot_dir.begin(), e.slot_dir.end(), std::int8_t{-1});
  }

  std::uint64_t budget() const {
    return e.state_budget != 0U ? e.state_budget
                                : static_cast<std::uint64_t>(e.S) * query_state_bytes(e.st.L, e.K);
  }

  void fail(std::int32_t q, int code, std::string const& msg) {
    if (!on_fail) throw status_error{code, msg};
    on_fail(q, code, msg);
  }

  round_args rargs() const {
    round_args a{};
    a.im = e.img->d;
    a.st = e.st;
    a.tune = e.tune;
    a.invalid = fwd ? SHRT_MAX : SHRT_MIN;
    if (e.profile && e.events.size() >= 4U) {
      for (int i = 0; i != 4; ++i) a.ev[i] = e.events[static_cast<std::size_t>(i)];
    }
    return a;
  }

  void collect_timings() {
    if (!e.profile || e.events.size() < 4U) return;
    float ms = 0.0F;
    if (hipEventElapsedTime(&ms, e.events[0], e.events[1]) == hipSuccess) {
      e.stats.route_scan_ms += ms;
      ++e.stats.route_scan_launches;
    }
    if (hipEventElapsedTime(&ms, e.events[2], e.events[3]) == hipSuccess) {
      e.stats.transfer_ms += ms;
      ++e.stats.transfer_launches;
    }
  }

  // Epoch begin (raptor.h:174-203) for the listed slots only: lanes of
  // other slots keep running.
  void begin_epochs(std::vector<std::uint32_t> const& starting) {
    auto const S = e.S, W = e.W, G = e.G, K = e.K;
    std::vector<std::uint64_t> rt_mask(G, 0U), bt_mask(G, 0U);
    std::vector<std::uint8_t> flags(S, 0U);
    std::vector<std::int16_t> worst(S, 0);
    std::vector<std::uint32_t> st_slot, st_loc, new_slots;
    std::vector<std::int16_t> st_delta;
    std::vector<std::uint64_t> dest_ops;
    std::vector<std::uint8_t> dest_set;
    for (auto const s : starting) {
      auto& si = slots[s];
      auto& p = pq[static_cast<std::size_t>(si.q)];
      auto const g = s / W, lane = s % W;
      bt_mask[g] |= 1ULL << lane;
      // reset_arrivals of an interval extension (search.h:299-311)
      auto const reset = !si.is_new && p.epochs_[si.epoch].reset_;
      if (reset) rt_mask[g] |= 1ULL << lane;
      if (si.is_new) {
        rt_mask[g] |= 1ULL << lane;
        new_slots.push_back(s);
        mark_execute_start(p);
        // slot parameters
        auto& sp = e.h_param[s];
        sp.tts_min = p.tts_.min_transfer_time_;
        sp.tts_add = p.tts_.additional_time_;
        sp.tts_factor = p.tts_.factor_;
        sp.tts_default = p.tts_.default_ ? 1U : 0U;
        sp.base = p.base_;
        sp.end_k = p.end_k_;  // <= K: checked at admission
        sp.intermodal = p.intermodal_dest_ ? 1U : 0U;
        sp.bounds_last_k = 0U;
        if (p.dests_.size() > kMaxDestBits || p.dist_loc_.size() > kMaxDist) {
          throw status_error{NIGIRI_AMD_EUNSUPPORTED, "too many destination offsets"};
        }
        sp.n_dest = static_cast<std::uint16_t>(p.dests_.size());
        sp.n_dist = static_cast<std::uint16_t>(p.dist_loc_.size());
        for (std::size_t j = 0; j != p.dests_.size(); ++j) e.h_dest[s * kMaxDestBits + j] = p.dests_[j];
        for (std::size_t j = 0; j != p.dist_loc_.size(); ++j) {
          e.h_dist_loc[s * kMaxDist + j] = p.dist_loc_[j];
          e.h_dist_min[s * kMaxDist + j] = p.dist_min_[j];
        }
        for (auto const l : e.slot_dest_bits[s]) {
          dest_ops.push_back((static_cast<std::uint64_t>(s) << 32) | l);
          dest_set.push_back(0U);
        }
        e.slot_dest_bits[s].clear();
        if (!p.intermodal_dest_) {
          for (auto const l : p.dests_) {
            dest_ops.push_back((static_cast<std::uint64_t>(s) << 32) | l);
            dest_set.push_back(1U);
            e.slot_dest_bits[s].push_back(l);
          }
        }
      }
      auto const& ep = p.epochs_[si.epoch];
      auto const base = p.base_minutes(tt);
      flags[s] = static_cast<std::uint8_t>(1U | (si.is_new || reset ? 2U : 0U));
      worst[s] = clamp16(static_cast<long long>(ep.worst_time_at_dest_) - base);
      for (std::size_t j = 0; j != ep.loc_.size(); ++j) {
        st_slot.push_back(s);
        st_loc.push_back(ep.loc_[j]);
        st_delta.push_back(clamp16(static_cast<long long>(ep.time_[j]) - base));
      }
      si.running = true;
      ++e.stats.epochs;
    }
    auto const stream = e.stream;
    auto const& tr = e.st;
    HIP_CHECK(hipMemcpyAsync(tr.param, e.h_param.data(), sizeof(slot_param) * S, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(tr.dest, e.h_dest.data(), sizeof(std::uint32_t) * e.h_dest.size(), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(tr.dist_loc, e.h_dist_loc.data(), sizeof(std::uint32_t) * e.h_dist_loc.size(), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(tr.dist_min, e.h_dist_min.data(), sizeof(std::uint16_t) * e.h_dist_min.size(), hipMemcpyHostToDevice, stream));
    epoch_args a{};
    a.st = e.st;
    a.invalid = fwd ? SHRT_MAX : SHRT_MIN;
    std::vector<std::uint32_t> reset_groups;
    for (std::uint32_t g = 0; g != G; ++g) {
      if (bt_mask[g] != 0U) reset_groups.push_back(g);
    }
    auto const full_mask = dense_resets(e, starting, fwd);
    HIP_CHECK(hipMemcpyAsync(e.d_full_mask, full_mask.data(), sizeof(std::uint64_t) * G, hipMemcpyHostToDevice, stream));
    a.reset_full_mask = e.d_full_mask;
    HIP_CHECK(hipMemcpyAsync(e.d_reset_groups, reset_groups.data(), sizeof(std::uint32_t) * reset_groups.size(), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(e.d_rt_mask, rt_mask.data(), sizeof(std::uint64_t) * G, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(e.d_bt_mask, bt_mask.data(), sizeof(std::uint64_t) * G, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(e.d_flags, flags.data(), S, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(e.d_worst, worst.data(), sizeof(std::int16_t) * S, hipMemcpyHostToDevice, stream));
    a.reset_groups = e.d_reset_groups;
    a.n_reset_groups = static_cast<std::uint32_t>(reset_groups.size());
    a.reset_rt_mask = e.d_rt_mask;
    a.reset_bt_mask = e.d_bt_mask;
    if (st_slot.size() > e.start_cap) {
      e.start_cap = st_slot.size() * 2U;
      e.d_start_slot = e.grow(e.d_start_slot, e.start_cap);
      e.d_start_loc = e.grow(e.d_start_loc, e.start_cap);
      e.d_start_delta = e.grow(e.d_start_delta, e.start_cap);
    }
    if (dest_ops.size() > e.dest_ops_cap) {
      e.dest_ops_cap = dest_ops.size() * 2U;
      e.d_dest_ops = e.grow(e.d_dest_ops, e.dest_ops_cap);
      e.d_dest_set = e.grow(e.d_dest_set, e.dest_ops_cap);
    }
    if (!st_slot.empty()) {
      HIP_CHECK(hipMemcpyAsync(e.d_start_slot, st_slot.data(), 4U * st_slot.size(), hipMemcpyHostToDevice, stream));
      HIP_CHECK(hipMemcpyAsync(e.d_start_loc, st_loc.data(), 4U * st_loc.size(), hipMemcpyHostToDevice, stream));
      HIP_CHECK(hipMemcpyAsync(e.d_start_delta, st_delta.data(), 2U * st_delta.size(), hipMemcpyHostToDevice, stream));
    }
    if (!dest_ops.empty()) {
      HIP_CHECK(hipMemcpyAsync(e.d_dest_ops, dest_ops.data(), 8U * dest_ops.size(), hipMemcpyHostToDevice, stream));
      HIP_CHECK(hipMemcpyAsync(e.d_dest_set, dest_set.data(), dest_set.size(), hipMemcpyHostToDevice, stream));
    }
    a.start_slot = e.d_start_slot;
    a.start_loc = e.d_start_loc;
    a.start_delta = e.d_start_delta;
    a.n_starts = static_cast<std::uint32_t>(st_slot.size());
    a.dest_ops = e.d_dest_ops;
    a.dest_set = e.d_dest_set;
    a.n_dest_ops = static_cast<std::uint32_t>(dest_ops.size());
    a.slot_flags = e.d_flags;
    a.slot_worst = e.d_worst;
    a.n_slots = S;
    // lower bounds of the new slots on the device (PHAST sweep from the
    // upward searches of their preparation)
    phast_lbs(new_slots);
    launch_epoch_begin(fwd, W, a, stream);
    HIP_CHECK(hipGetLastError());
    // the host vectors above go out of scope: wait for the copies
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  // One step: idle slots start the next epoch of their range query or take
  // a new query, then every running slot advances one round (each at its
  // own round index); slots whose epoch ended hand their destination rows
  // to the host.  Returns false when no slot runs.
  bool step() {
    auto const S = e.S, W = e.W, G = e.G, K = e.K;
    // continuing range queries start their next epoch; free slots take
    // queued queries, ordered by locality so that neighbouring lanes start
    // close to each other
    std::vector<std::uint32_t> free_list, starting;
    for (std::uint32_t s = 0; s != S; ++s) {
      auto& si = slots[s];
      si.is_new = false;
      if (si.running) continue;
      if (si.q >= 0 && si.epoch + 1U < pq[static_cast<std::size_t>(si.q)].epochs_.size()) {
        ++si.epoch;
        starting.push_back(s);
      } else {
        si.q = -1;
        free_list.push_back(s);
      }
    }
    std::vector<std::int32_t> admitted;
    // whole-batch admission: new queries only start together, once no slot
    // is busy (their preparation and lower bounds overlap the rounds before)
    bool const gate_closed = e.admit_when_idle && (!starting.empty() || any_running());
    while (!gate_closed && admitted.size() < free_list.size()) {
      std::int32_t q = held;
      held = -1;
      if (q < 0) q = next();
      if (q < 0) break;
      auto& p = pq[static_cast<std::size_t>(q)];
      if (p.epochs_.empty()) {  // overlap or no start labels
        finish(p);
        on_done(q);
        continue;
      }
      // admission by the device memory the query's state takes: its
      // end_k rows of round_times over all locations, plus best, tmp, lb and
      // marks; refused alone when it can never fit
      auto const need = query_state_bytes(e.st.L, p.end_k_);
      if (p.end_k_ > K || need > budget()) {
        fail(q, NIGIRI_AMD_ECAPACITY,
             "query needs " + std::to_string(p.end_k_) + " rows of round_times (" + std::to_string(need) +
                 " B of state); the engine holds " + std::to_string(K) + " rows per slot, budget " +
                 std::to_string(budget()) + " B");
        continue;
      }
      if (used_bytes + need > budget()) {
        held = q;  // admitted once running queries free their state
        break;
      }
      used_bytes += need;
      p.state_bytes_ = need;
      admitted.push_back(q);
    }
    std::stable_sort(admitted.begin(), admitted.end(), [&](std::int32_t a, std::int32_t b) {
      return pq[static_cast<std::size_t>(a)].locality_key_ < pq[static_cast<std::size_t>(b)].locality_key_;
    });
    for (std::size_t j = 0; j != admitted.size(); ++j) {
      auto const s = free_list[j];
      auto& si = slots[s];
      si.q = admitted[j];
      pq[static_cast<std::size_t>(admitted[j])].slot_ = static_cast<std::int32_t>(s);
      si.epoch = 0U;
      si.is_new = true;
      starting.push_back(s);
      if (on_admit) on_admit(admitted[j]);
    }
    if (!starting.empty()) begin_epochs(starting);

    bool any = false;
    std::vector<std::uint8_t> sweep(G, 0U);
    for (std::uint32_t s = 0; s != S; ++s) {
      if (!slots[s].running) continue;
      any = true;
      ++pq[static_cast<std::size_t>(slots[s].q)].rounds_;
      // after its first epoch a slot sees earlier rows in round_times
      if (slots[s].epoch != 0U) sweep[s / W] = 1U;
    }
    if (!any) return settle();
    auto const stream = e.stream;
    auto const& tr = e.st;
    // the sweep list in device memory: [count, groups...]
    std::vector<std::uint32_t> sweep_groups{0U};
    for (std::uint32_t g = 0; g != G; ++g) {
      if (sweep[g] != 0U) sweep_groups.push_back(g);
    }
    sweep_groups[0] = static_cast<std::uint32_t>(sweep_groups.size() - 1U);
    bool const has_sweep = sweep_groups[0] != 0U;
    if (has_sweep) {
      HIP_CHECK(hipMemcpyAsync(e.d_sweep_groups, sweep_groups.data(), 4U * sweep_groups.size(), hipMemcpyHostToDevice, stream));
    }
    if (e.use_graphs) {
      HIP_CHECK(launch_round_graph(*e.graphs, fwd, W, rargs(), e.d_sweep_groups, has_sweep, S,
                                   static_cast<int>(e.grid_waves), stream));
      ++e.stats.graph_rounds;
    } else {
      launch_round(fwd, W, rargs(), e.d_sweep_groups, has_sweep, S, static_cast<int>(e.grid_waves), stream);
    }
    HIP_CHECK(hipGetLastError());
    settle();  // while the round runs on the device
    ++rounds;
    ++e.stats.rounds;
    HIP_CHECK(hipMemcpyAsync(e.h_active.p, tr.grp_active, 8U * G, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipMemcpyAsync(e.h_counts.p, tr.counts, 4U * 8U, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if (e.h_counts.p[4] != 0U) {
      auto const flags = e.h_counts.p[4];
      std::uint32_t zero = 0U;
      HIP_CHECK(hipMemcpy(tr.counts + 4, &zero, 4U, hipMemcpyHostToDevice));
      throw status_error{NIGIRI_AMD_ECAPACITY,
                         "device work list overflow (flags " + std::to_string(flags) + ")"};
    }
    collect_timings();

    // slots whose epoch ended in this round
    std::vector<std::uint32_t> finished;
    for (std::uint32_t s = 0; s != S; ++s) {
      if (slots[s].running && ((e.h_active.p[s / W] >> (s % W)) & 1ULL) == 0ULL) finished.push_back(s);
    }
    if (finished.empty()) return true;
    HIP_CHECK(hipMemcpyAsync(e.d_slot_list, finished.data(), 4U * finished.size(), hipMemcpyHostToDevice, stream));
    launch_extract(W, tr, e.d_slot_list, static_cast<std::uint32_t>(finished.size()), stream);
    HIP_CHECK(hipGetLastError());
    auto const per_slot = static_cast<std::size_t>(kMaxDestBits) * K;
    HIP_CHECK(hipMemcpyAsync(e.h_rows.p, tr.dest_rows, 2U * per_slot * finished.size(), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    auto const t0 = std::chrono::steady_clock::now();
    for (std::size_t j = 0; j != finished.size(); ++j) {
      auto& si = slots[finished[j]];
      si.running = false;
      add_epoch_results(tt, pq[static_cast<std::size_t>(si.q)], si.epoch, e.h_rows.p + j * per_slot, K);
    }
    auto const t1 = std::chrono::steady_clock::now();
    reconstruct_new_journeys(finished);
    e.stats.results_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    e.stats.reconstruct_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    // after each epoch the search may stop early or, after its last one,
    // widen the interval and go on (search.h:237-317, 477-491); queries whose
    // last epoch ended complete once their legs are written
    for (auto const s : finished) {
      auto const& si = slots[s];
      auto& p = pq[static_cast<std::size_t>(si.q)];
      early_termination(p, si.epoch);
      if (si.epoch + 1U != p.epochs_.size()) continue;
      if (wants_extension(p)) {
        // extending may remove results the leg writers still refer to
        join_legs();
        if (extend_interval(tt, p)) continue;  // the slot runs the new epochs
      }
      // the query's device state is free once its last epoch's journeys
      // are reconstructed (the legs are written on the host from records)
      used_bytes -= p.state_bytes_;
      p.state_bytes_ = 0U;
      completing.push_back(si.q);
    }
    if (!legs_pending) settle();
    return true;
  }

  void join_legs() {
    if (!legs_pending) return;
    auto const t0 = std::chrono::steady_clock::now();
    e.pool->wait();
    legs_pending = false;
    e.stats.legs_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  // Joins the leg writing started by the last step and completes the queries
  // that waited for it.  Returns whether any query completed.
  bool settle() {
    join_legs();
    auto const any = !completing.empty();
    for (auto const q : completing) {
      finish(pq[static_cast<std::size_t>(q)]);
      on_done(q);
    }
    completing.clear();
    return any;
  }

  // Reconstructs the journeys that the epochs just ended added to their
  // queries' result sets (search.h:457-475), against the state those epochs
  // left: before any of these slots starts its next epoch.
  void reconstruct_new_journeys(std::vector<std::uint32_t> const& finished) {
    std::vector<rc_job> jobs;
    std::vector<std::pair<std::int32_t, std::size_t>> owner;  // (query, journey)
    std::vector<std::uint32_t> off_loc;
    std::vector<std::int16_t> off_dur;
    std::vector<std::size_t> cand;
    for (auto const s : finished) {
      auto const q = slots[s].q;
      auto const& p = pq[static_cast<std::size_t>(q)];
      reconstruction_candidates(p, cand);
      if (cand.empty()) continue;
      auto const off_begin = static_cast<std::uint32_t>(off_loc.size());
      std::uint16_t n_off = 0U;
      if (p.intermodal_dest_) {
        for (auto const& [loc, dur] : p.dest_offsets_) {
          off_loc.push_back(loc);
          off_dur.push_back(dur);
        }
        n_off = static_cast<std::uint16_t>(p.dest_offsets_.size());
      }
      for (auto const i : cand) {
        auto const& j = p.results_[i];
        jobs.push_back(rc_job{s, j.dest_, off_begin, n_off, j.transfers_,
                              static_cast<std::uint8_t>(p.intermodal_dest_ ? 1U : 0U)});
        owner.emplace_back(q, i);
      }
    }
    if (jobs.empty()) return;
    run_reconstruct(e, fwd, jobs, off_loc, off_dur);
    // the legs are written by the pool while the next round runs on the
    // device (settle() joins): journeys are independent, each writes only
    // its own result element, the timetable is read-only, and neither these
    // queries' results nor e.h_rc_out change before the join
    rc_owner = std::move(owner);
    e.pool->start(jobs.size(), [this](std::size_t x) {
      auto& p = pq[static_cast<std::size_t>(rc_owner[x].first)];
      assemble_legs(tt, p, e.h_rc_out[x], p.results_[rc_owner[x].second]);
    });
    legs_pending = true;
    e.stats.reconstructed += jobs.size();
  }

  bool any_running() const {
    return std::any_of(slots.begin(), slots.end(), [](slot_info const& si) { return si.running; });
  }

  // a slot runs an epoch or has further epochs of its range query to run
  bool any_busy() const {
    for (auto const& si : slots) {
      if (si.running || (si.q >= 0 && si.epoch + 1U < pq[static_cast<std::size_t>(si.q)].epochs_.size())) {
        return true;
      }
    }
    return false;
  }

  // idle slots that take a new query at the next step
  std::uint32_t free_slots() const {
    std::uint32_t n = 0U;
    for (auto const& si : slots) {
      if (!si.running && (si.q < 0 || si.epoch + 1U >= pq[static_cast<std::size_t>(si.q)].epochs_.size())) ++n;
    }
    return n;
  }

  // Lower bounds of the new slots (search.h:114-131): PHAST on the device,
  // seeded with the upward searches (and the host levels of the sweep) of
  // their host preparation.  Plain engine: each slot's lower bounds are a
  // row [L] of lb, swept in place.  Batched engine: a group with at least
  // kPhastDenseLanes new lanes is swept in place, a wavefront per (location,
  // group) whose lanes share each 128 B row (a whole batch admitted at once);
  // the new slots of the other groups (the one or two lanes a group gets in
  // a serving step) are swept as rows of the engine's PHAST buffer, a thread
  // per (location, slot), up to kPhastRows at a time, then copied into their
  // lanes.
  void phast_lbs(std::vector<std::uint32_t> const& new_slots) {
    if (new_slots.empty()) return;
    auto const L = e.st.L, W = e.W;
    std::vector<std::uint32_t> dense_sel, rows;
    std::vector<std::uint64_t> dense_mask;
    if (W == 64U) {
      std::vector<std::uint64_t> fresh(e.G, 0ULL);
      for (auto const s : new_slots) fresh[s / 64U] |= 1ULL << (s % 64U);
      for (std::uint32_t g = 0; g != e.G; ++g) {
        if (static_cast<std::uint32_t>(__builtin_popcountll(fresh[g])) >= kPhastDenseLanes) {
          dense_sel.push_back(g);
          dense_mask.push_back(fresh[g]);
        }
      }
      for (auto const s : new_slots) {
        if ((static_cast<std::uint32_t>(__builtin_popcountll(fresh[s / 64U]))) < kPhastDenseLanes) rows.push_back(s);
      }
    } else {
      rows = new_slots;
    }
    // seeds: the dense groups' (in the [G][L][64] layout of lb), then the
    // rows' in passes of rows_per ([row][L] of the buffer, or of lb)
    auto const rows_per = W == 64U ? std::min<std::size_t>(kPhastRows, std::max<std::size_t>(1U, rows.size()))
                                   : std::max<std::size_t>(1U, rows.size());
    std::vector<std::uint32_t> sel(dense_sel), seed_idx, seed_off{0U};
    std::vector<std::uint16_t> seed_min;
    auto const add_seeds = [&](std::uint32_t s, std::uint32_t base, std::uint32_t stride, std::uint32_t lane) {
      for (auto const& [l, d] : pq[static_cast<std::size_t>(slots[s].q)].lb_up_) {
        seed_idx.push_back((base + l) * stride + lane);
        seed_min.push_back(d);
      }
    };
    for (auto const s : new_slots) {
      if (W == 64U && std::binary_search(dense_sel.begin(), dense_sel.end(), s / 64U)) {
        add_seeds(s, (s / 64U) * L, 64U, s % 64U);
      }
    }
    seed_off.push_back(static_cast<std::uint32_t>(seed_idx.size()));
    for (std::size_t i = 0; i != rows.size(); ++i) {
      auto const row = W == 64U ? static_cast<std::uint32_t>(i % kPhastRows) : rows[i];
      if (i < rows_per) sel.push_back(row);
      add_seeds(rows[i], row * L, 1U, 0U);
      if ((i + 1U) % rows_per == 0U || i + 1U == rows.size()) {
        seed_off.push_back(static_cast<std::uint32_t>(seed_idx.size()));
      }
    }
    auto const stream = e.stream;
    if (seed_idx.size() > e.ph_seed_cap) {
      e.ph_seed_cap = seed_idx.size() * 2U;
      e.d_ph_seed_idx = e.grow(e.d_ph_seed_idx, e.ph_seed_cap);
      e.d_ph_seed_min = e.grow(e.d_ph_seed_min, e.ph_seed_cap);
    }
    HIP_CHECK(hipMemcpyAsync(e.d_ph_sel, sel.data(), 4U * sel.size(), hipMemcpyHostToDevice, stream));
    if (!dense_mask.empty()) {
      HIP_CHECK(hipMemcpyAsync(e.d_ph_masks, dense_mask.data(), 8U * dense_mask.size(), hipMemcpyHostToDevice,
                               stream));
    }
    if (W == 64U && !rows.empty()) {
      HIP_CHECK(hipMemcpyAsync(e.d_ph_slots, rows.data(), 4U * rows.size(), hipMemcpyHostToDevice, stream));
    }
    if (!seed_idx.empty()) {
      HIP_CHECK(hipMemcpyAsync(e.d_ph_seed_idx, seed_idx.data(), 4U * seed_idx.size(), hipMemcpyHostToDevice,
                               stream));
      HIP_CHECK(hipMemcpyAsync(e.d_ph_seed_min, seed_min.data(), 2U * seed_min.size(), hipMemcpyHostToDevice,
                               stream));
    }
    auto const& sw = fwd ? tt.ch.sweep_fwd : tt.ch.sweep_bwd;
    auto const n_levels = static_cast<std::uint32_t>(sw.level_off.size() - 1U);
    lb_phast_args a{};
    a.sw = e.img->d.sweep[fwd ? 0 : 1];
    a.L = L;
    if (!dense_sel.empty()) {
      a.dist = e.st.lb;
      a.sel = e.d_ph_sel;
      a.masks = e.d_ph_masks;
      a.n_sel = static_cast<std::uint32_t>(dense_sel.size());
      a.seed_idx = e.d_ph_seed_idx;
      a.seed_min = e.d_ph_seed_min;
      a.n_seeds = seed_off[1];
      launch_lb_phast(64U, a, sw.level_off.data(), sw.host_levels, n_levels, stream);
      HIP_CHECK(hipGetLastError());
    }
    for (std::size_t b = 1; b + 1U < seed_off.size(); ++b) {
      auto const n_rows = static_cast<std::uint32_t>(std::min(rows_per, rows.size() - (b - 1U) * rows_per));
      a.dist = W == 64U ? e.d_ph_rows : e.st.lb;
      a.sel = e.d_ph_sel + dense_sel.size();
      a.masks = nullptr;
      a.n_sel = n_rows;
      a.seed_idx = e.d_ph_seed_idx + seed_off[b];
      a.seed_min = e.d_ph_seed_min + seed_off[b];
      a.n_seeds = seed_off[b + 1U] - seed_off[b];
      launch_lb_phast(1U, a, sw.level_off.data(), sw.host_levels, n_levels, stream);
      if (W == 64U) launch_lb_scatter64(e.st, e.d_ph_rows, e.d_ph_slots + (b - 1U) * rows_per, n_rows, stream);
      HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipStreamSynchronize(stream));  // host vectors go out of scope
  }
};

// prepare queries on host threads
void prepare_all(timetable const& tt, nigiri_query const* qs, std::uint32_t n,
                 std::vector<prepared_query>& pq, std::vector<int>& status,
                 std::vector<std::string>& errors, bool host_lb,
                 std::vector<std::int16_t> const* fd = nullptr, std::vector<std::uint8_t> const* fd_have = nullptr) {
  pq.resize(n);
  status.assign(n, NIGIRI_AMD_OK);
  errors.assign(n, {});
  auto const hw = std::max(1U, std::thread::hardware_concurrency());
  auto const threads = std::min<std::uint32_t>(std::min<std::uint32_t>(hw, 64U), std::max(1U, n));
  std::atomic<std::uint32_t> next{0U};
  std::vector<std::thread> pool;
  for (std::uint32_t t = 0; t != threads; ++t) {
    pool.emplace_back([&]() {
      host_scratch scratch;
      while (true) {
        auto const i = next.fetch_add(1U);
        if (i >= n) break;
        try {
          if (fd_have != nullptr && (*fd_have)[i] == 2U) {
            throw status_error{NIGIRI_AMD_ECAPACITY, "more than " + std::to_string(kFdMaxNodes) +
                                                         " start and destination locations"};
          }
          auto const* given = fd != nullptr && (*fd_have)[i] == 1U ? &(*fd)[i] : nullptr;
          prepare(tt, qs[i], pq[i], scratch, host_lb, given);
        } catch (status_error const& e) {
          status[i] = e.code_;
          errors[i] = e.what();
        } catch (std::exception const& e) {
          status[i] = NIGIRI_AMD_EINVAL;
          errors[i] = e.what();
        }
      }
    });
  }
  for (auto& t : pool) t.join();
}

std::uint32_t write_results(std::vector<prepared_query> const& pq, std::vector<int> const& status,
                            nigiri_result* res, nigiri_journey* journeys, std::uint32_t cap,
                            nigiri_leg* legs, std::uint32_t leg_cap,
                            std::vector<std::uint32_t> const* epochs,
                            std::vector<double> const* latency, std::vector<double> const* queue = nullptr) {
  std::uint32_t pos = 0U, leg_pos = 0U;
  int overflow = 0;
  for (std::size_t i = 0; i != pq.size(); ++i) {
    auto& r = res[i];
    r = nigiri_result{};
    r.status = status[i];
    r.journey_offset = pos;
    r.slot = -1;
    if (status[i] != NIGIRI_AMD_OK) continue;
    auto const& p = pq[i];
    r.fastest_direct = p.fastest_direct_;
    r.base_day = static_cast<std::int16_t>(p.base_);
    r.interval_from = p.iv_from_;
    r.interval_to = p.iv_to_;
    r.n_epochs = static_cast<std::uint32_t>(p.epochs_.size());
    r.slot = p.slot_;
    r.n_rounds = epochs != nullptr ? (*epochs)[i] : p.rounds_;
    if (latency != nullptr) r.latency_ms = (*latency)[i];
    if (queue != nullptr) r.queue_ms = (*queue)[i];
    for (auto const& j : p.results_) {
      if (pos < cap && journeys != nullptr) {
        auto& out = journeys[pos];
        out = nigiri_journey{j.start_time_, j.dest_time_, j.dest_, j.transfers_, {}, leg_pos, 0U};
        if (legs != nullptr) {
          if (leg_pos + j.legs_.size() <= leg_cap) {
            std::copy(j.legs_.begin(), j.legs_.end(), legs + leg_pos);
            out.n_legs = static_cast<std::uint32_t>(j.legs_.size());
            leg_pos += out.n_legs;
          } else {
            overflow = 1;
          }
        }
        ++pos;
        ++r.n_journeys;
      } else {
        overflow = 1;
      }
    }
  }
  return overflow != 0 ? ~0U : pos;
}

}  // namespace

// used by the serving loop (serve.cc)
void engine_run_dir(nigiri_amd_engine& e, bool fwd, std::vector<prepared_query>& pq,
                    std::function<std::int32_t()> next, std::function<void(std::int32_t)> done,
                    std::function<void(std::int32_t, int, std::string const&)> fail) {
  runner r{e, *e.img->tt, fwd, pq};
  r.next = std::move(next);
  r.on_done = std::move(done);
  r.on_fail = std::move(fail);
  while (r.step() || r.held >= 0) {
  }
}

}  // namespace nigiri_amd

namespace nigiri_amd {
timetable const& engine_timetable(nigiri_amd_engine const* e) { return *e->img->tt; }
std::uint32_t engine_lanes(nigiri_amd_engine const* e) { return e->W; }
std::uint32_t engine_slots(nigiri_amd_engine const* e) { return e->S; }
std::uint32_t engine_max_k(nigiri_amd_engine const* e) { return e->K; }
}  // namespace nigiri_amd

extern "C" {

int nigiri_amd_image_create(const nigiri_amd_tt* tt, int device, nigiri_amd_image** out) {
  if (tt == nullptr || out == nullptr) {
    set_error("image_create: null argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(device));
    auto img = std::make_unique<nigiri_amd_image>();
    img->tt = &tt->tt;
    img->device = device;
    hipStream_t s = nullptr;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    try {
      build_image(*img, tt->tt, s);
      HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipStreamDestroy(s);
      throw;
    }
    HIP_CHECK(hipStreamDestroy(s));
    *out = img.release();
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_image_bytes(const nigiri_amd_image* img, uint64_t* bytes) {
  if (img == nullptr || bytes == nullptr) return NIGIRI_AMD_EINVAL;
  *bytes = img->mem.bytes;
  return NIGIRI_AMD_OK;
}

void nigiri_amd_image_destroy(nigiri_amd_image* img) { delete img; }

namespace nigiri_amd {
namespace {

// The upward search and the host levels of the sweep (host) of a
// lower-bound query; throws on bad seeds.
void upward_labels(nigiri_amd_image const& img, int32_t dir, uint32_t n_seeds, const uint32_t* seed_loc,
                   const uint16_t* seed_min, std::vector<std::pair<std::uint32_t, std::uint16_t>>& up) {
  thread_local ch_search_scratch s;
  thread_local std::vector<std::pair<std::uint32_t, std::uint16_t>> seeds;
  seeds.clear();
  for (uint32_t i = 0; i != n_seeds; ++i) {
    if (seed_loc[i] >= img.d.L) throw status_error{NIGIRI_AMD_EINVAL, "lower_bounds: seed out of range"};
    seeds.emplace_back(seed_loc[i], seed_min[i]);
  }
  ch_sweep_seeds(img.tt->ch, dir == NIGIRI_AMD_DIR_FWD, seeds, s, up);
}

// Stages the labels (host pinned-free copies into sel / seed arrays at the
// given device addresses) and runs the sweep into the row lb of [L].
void lb_row_sweep(nigiri_amd_image const& img, int32_t dir, std::vector<std::pair<std::uint32_t, std::uint16_t>> const& up,
                  std::uint32_t* d_sel, std::uint32_t* d_idx, std::uint16_t* d_min, std::uint16_t* lb) {
  thread_local std::vector<std::uint32_t> idx;
  thread_local std::vector<std::uint16_t> mins;
  idx.clear();
  mins.clear();
  for (auto const& [l, d] : up) {
    idx.push_back(l);
    mins.push_back(d);
  }
  std::uint32_t const zero = 0U;
  HIP_CHECK(hipMemcpy(d_sel, &zero, 4U, hipMemcpyHostToDevice));
  if (!idx.empty()) {
    HIP_CHECK(hipMemcpy(d_idx, idx.data(), 4U * idx.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(d_min, mins.data(), 2U * mins.size(), hipMemcpyHostToDevice));
  }
  lb_phast_args a{};
  a.sw = img.d.sweep[dir == NIGIRI_AMD_DIR_FWD ? 0 : 1];
  a.dist = lb;
  a.L = img.d.L;
  a.sel = d_sel;
  a.masks = nullptr;
  a.n_sel = 1U;
  a.seed_idx = d_idx;
  a.seed_min = d_min;
  a.n_seeds = static_cast<std::uint32_t>(idx.size());
  auto const& sw = dir == NIGIRI_AMD_DIR_FWD ? img.tt->ch.sweep_fwd : img.tt->ch.sweep_bwd;
  launch_lb_phast(1U, a, sw.level_off.data(), sw.host_levels, static_cast<std::uint32_t>(sw.level_off.size() - 1U),
                  nullptr);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(nullptr));
}

bool lb_args_ok(nigiri_amd_image const* img, int32_t dir, uint32_t n_seeds, const uint32_t* seed_loc,
                const uint16_t* seed_min, uint16_t const* lb) {
  return img != nullptr && lb != nullptr && (n_seeds == 0U || (seed_loc != nullptr && seed_min != nullptr)) &&
         (dir == NIGIRI_AMD_DIR_FWD || dir == NIGIRI_AMD_DIR_BWD);
}

}  // namespace
}  // namespace nigiri_amd

// Scratch of nigiri_amd_lower_bounds_noalloc: the selected row index (first
// 256 bytes), then up to kNoAllocMaxLabels upward-search labels (location
// indices, then minutes).  Each GPU holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// gives on that device.
constexpr std::size_t kNoAllocMaxLabels = std::size_t{1} << 20;
constexpr std::size_t kNoAllocScratchBytes = 256U + kNoAllocMaxLabels * 6U;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace nigiri_amd {
namespace {

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;  // per device, set up on first use

unsigned char* scratch_on(int device) {
  std::lock_guard<std::mutex> lock{g_scratch_mu};
  if (g_scratch_ptr.size() <= static_cast<std::size_t>(device)) {
    g_scratch_ptr.resize(static_cast<std::size_t>(device) + 1U, nullptr);
  }
  auto& p = g_scratch_ptr[static_cast<std::size_t>(device)];
  if (p == nullptr) {  // the device is current (hipSetDevice by the caller)
    void* addr = nullptr;
    auto const err = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || addr == nullptr) {
      throw status_error{NIGIRI_AMD_EDEVICE, std::string{"hipGetSymbolAddress(scratch): "} + hipGetErrorString(err)};
    }
    p = static_cast<unsigned char*>(addr);
  }
  return p;
}

}  // namespace
}  // namespace nigiri_amd

int nigiri_amd_lower_bounds_noalloc(nigiri_amd_image* img, int32_t dir, uint32_t n_seeds, const uint32_t* seed_loc,
                                    const uint16_t* seed_min, uint16_t* lb) {
  if (!lb_args_ok(img, dir, n_seeds, seed_loc, seed_min, lb)) {
    set_error("lower_bounds_noalloc: bad argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(img->device));
    auto* base = scratch_on(img->device);
    thread_local std::vector<std::pair<std::uint32_t, std::uint16_t>> up;
    upward_labels(*img, dir, n_seeds, seed_loc, seed_min, up);
    if (up.size() > kNoAllocMaxLabels) {
      throw status_error{NIGIRI_AMD_ECAPACITY, "lower_bounds_noalloc: " + std::to_string(up.size()) +
                                                   " upward-search labels exceed the scratch capacity of " +
                                                   std::to_string(kNoAllocMaxLabels)};
    }
    auto* d_sel = reinterpret_cast<std::uint32_t*>(base);
    auto* d_idx = reinterpret_cast<std::uint32_t*>(base + 256U);
    auto* d_min = reinterpret_cast<std::uint16_t*>(base + 256U + 4U * kNoAllocMaxLabels);
    lb_row_sweep(*img, dir, up, d_sel, d_idx, d_min, lb);
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_lower_bounds(nigiri_amd_image* img, int32_t dir, uint32_t n_seeds, const uint32_t* seed_loc,
                            const uint16_t* seed_min, uint16_t* lb) {
  if (!lb_args_ok(img, dir, n_seeds, seed_loc, seed_min, lb)) {
    set_error("lower_bounds: bad argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(img->device));
    std::lock_guard<std::mutex> lock{img->lb_mu};
    std::vector<std::pair<std::uint32_t, std::uint16_t>> up;
    upward_labels(*img, dir, n_seeds, seed_loc, seed_min, up);
    if (img->d_lb_sel == nullptr) img->d_lb_sel = img->mem.alloc<std::uint32_t>(1U);
    if (up.size() > img->lb_seed_cap) {
      if (img->d_lb_seed_idx != nullptr) (void)hipFree(img->d_lb_seed_idx);
      if (img->d_lb_seed_min != nullptr) (void)hipFree(img->d_lb_seed_min);
      img->d_lb_seed_idx = nullptr;
      img->d_lb_seed_min = nullptr;
      img->lb_seed_cap = 0U;
      auto const cap = std::max<std::size_t>(up.size() * 2U, 1024U);
      HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&img->d_lb_seed_idx), 4U * cap));
      HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&img->d_lb_seed_min), 2U * cap));
      img->lb_seed_cap = cap;
    }
    lb_row_sweep(*img, dir, up, img->d_lb_sel, img->d_lb_seed_idx, img->d_lb_seed_min, lb);
    return NIGIRI_AMD_OK;
  });
}

namespace nigiri_amd {
namespace {

static_assert(sizeof(fd_node) == sizeof(fd_node_host), "fastest-direct node layouts agree");

// get_fastest_direct of queries qs[0 .. n) on the device.  have[i] = 1 where
// val[i] holds the result; 0 for a query whose node set cannot be built (a
// location out of range: its preparation reports that) and for one that
// brings its own value (has_fastest_direct); 2 for a query with more than
// kFdMaxNodes start and destination locations (not computed).
void device_fastest_direct(nigiri_amd_image& img, nigiri_query const* qs, std::uint32_t n,
                           std::vector<std::int16_t>& val, std::vector<std::uint8_t>& have) {
  val.assign(n, 0);
  have.assign(n, 0U);
  std::vector<fd_query> hq;
  std::vector<fd_node_host> hn, nodes;
  std::vector<std::uint32_t> idx;
  for (std::uint32_t i = 0; i != n; ++i) {
    auto const& q = qs[i];
    if (q.has_fastest_direct != 0) continue;
    try {
      fastest_direct_nodes(*img.tt, q, nodes);
    } catch (status_error const&) {
      continue;
    }
    if (nodes.size() > kFdMaxNodes) {
      have[i] = 2U;
      continue;
    }
    fd_query f{};
    f.node_begin = static_cast<std::uint32_t>(hn.size());
    hn.insert(hn.end(), nodes.begin(), nodes.end());
    f.node_end = static_cast<std::uint32_t>(hn.size());
    f.factor = q.fastest_direct_factor;
    f.tts_factor = q.tts_factor;
    f.tts_min = q.tts_min_transfer_time;
    f.tts_add = q.tts_additional_time;
    f.fwd = q.dir == NIGIRI_AMD_DIR_FWD ? 1U : 0U;
    f.tts_default = q.tts_default != 0 ? 1U : 0U;
    hq.push_back(f);
    idx.push_back(i);
  }
  if (hq.empty()) return;
  auto const bytes_q = sizeof(fd_query) * hq.size();
  auto const bytes_n = (sizeof(fd_node) * hn.size() + 255U) / 256U * 256U;
  auto const bytes_o = sizeof(std::int16_t) * hq.size();
  auto const need = bytes_q + bytes_n + bytes_o + 512U;
  std::lock_guard<std::mutex> lock{img.fd_mu};
  if (need > img.fd_cap) {
    if (img.d_fd != nullptr) (void)hipFree(img.d_fd);
    img.d_fd = nullptr;
    img.fd_cap = 0U;
    HIP_CHECK(hipMalloc(&img.d_fd, need * 2U));
    img.fd_cap = need * 2U;
  }
  auto* base = static_cast<unsigned char*>(img.d_fd);
  auto* d_nodes = reinterpret_cast<fd_node*>(base);
  auto* d_q = reinterpret_cast<fd_query*>(base + bytes_n);
  auto* d_out = reinterpret_cast<std::int16_t*>(base + bytes_n + (bytes_q + 255U) / 256U * 256U);
  if (!hn.empty()) HIP_CHECK(hipMemcpy(d_nodes, hn.data(), sizeof(fd_node) * hn.size(), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_q, hq.data(), bytes_q, hipMemcpyHostToDevice));
  fd_args a{};
  a.im = img.d;
  a.q = d_q;
  a.nodes = d_nodes;
  a.out = d_out;
  a.n = static_cast<std::uint32_t>(hq.size());
  launch_fastest_direct(a, nullptr);
  HIP_CHECK(hipGetLastError());
  std::vector<std::int16_t> out(hq.size());
  HIP_CHECK(hipMemcpy(out.data(), d_out, bytes_o, hipMemcpyDeviceToHost));
  for (std::size_t j = 0; j != idx.size(); ++j) {
    val[idx[j]] = out[j];
    have[idx[j]] = 1U;
  }
}

}  // namespace
}  // namespace nigiri_amd

int nigiri_amd_fastest_direct(nigiri_amd_image* img, const nigiri_query* qs, uint32_t n, int16_t* out) {
  if (img == nullptr || (n != 0U && (qs == nullptr || out == nullptr))) {
    set_error("fastest_direct: bad argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(img->device));
    std::vector<fd_node_host> nodes;
    for (uint32_t i = 0; i != n; ++i) {
      if (qs[i].has_fastest_direct == 0) fastest_direct_nodes(*img->tt, qs[i], nodes);  // throws EINVAL
    }
    std::vector<std::int16_t> val;
    std::vector<std::uint8_t> have;
    device_fastest_direct(*img, qs, n, val, have);
    if (std::find(have.begin(), have.end(), std::uint8_t{2U}) != have.end()) {
      throw status_error{NIGIRI_AMD_ECAPACITY, "fastest_direct: more than " + std::to_string(kFdMaxNodes) +
                                                   " start and destination locations in one query"};
    }
    for (uint32_t i = 0; i != n; ++i) out[i] = qs[i].has_fastest_direct != 0 ? qs[i].fastest_direct : val[i];
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_engine_create(nigiri_amd_image* img, uint32_t lanes, uint32_t n_groups,
                             uint32_t max_k, nigiri_amd_engine** out) {
  if (img == nullptr || out == nullptr || (lanes != 64U && lanes != 1U) || n_groups == 0U ||
      (lanes == 64U && n_groups > 1024U) ||
      max_k < 2U || max_k > 16U) {
    set_error("engine_create: lanes must be 1 or 64, 0 < n_groups (<= 1024 for 64 lanes), "
              "2 <= max_k <= 16");
    return NIGIRI_AMD_EINVAL;
  }
  if (static_cast<std::uint64_t>(n_groups) * img->d.L * lanes >= (1ULL << 32)) {
    set_error("engine_create: n_groups * n_locations * lanes must stay below 2^32");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(img->device));
    auto e = std::make_unique<nigiri_amd_engine>();
    e->pool = std::make_unique<host_pool>(kLegWorkers);
    e->img = img;
    e->W = lanes;
    e->G = n_groups;
    e->K = max_k;
    e->S = lanes * n_groups;
    HIP_CHECK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    {
      int cus = 0;
      HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, img->device));
      e->grid_waves = static_cast<std::uint32_t>(std::max(1, cus) * 32);
    }
    auto& st = e->st;
    auto& m = e->mem;
    auto const L = img->d.L, R = img->d.R, G = e->G, W = e->W, K = e->K, S = e->S;
    st.W = W;
    st.G = G;
    st.K = K;
    st.L = L;
    st.R = R;
    st.LW = (L + 63U) / 64U;
    st.RW = (R + 63U) / 64U;
    auto const cols = static_cast<std::size_t>(G) * L * W;
    st.round_times = m.alloc<std::int16_t>(cols * K);
    st.bounds = W == 64U ? nullptr : m.alloc<std::int16_t>(cols * K);
    st.best = m.alloc<std::int16_t>(cols);
    st.tmp = m.alloc<std::int32_t>(cols);
    st.lb = m.alloc<std::uint16_t>(cols);
    auto const lwords = W == 64U ? static_cast<std::size_t>(G) * L : static_cast<std::size_t>(G) * st.LW;
    auto const rwords = W == 64U ? static_cast<std::size_t>(G) * R : static_cast<std::size_t>(G) * st.RW;
    st.mark_t = m.alloc<std::uint64_t>(lwords);
    st.mark_r = m.alloc<std::uint64_t>(lwords);
    st.dest_mask = m.alloc<std::uint64_t>(lwords);
    st.hz_mask = m.alloc<std::uint64_t>(lwords);
    auto const lflag = 8U * flag_words(static_cast<std::size_t>(G) * L);
    auto const rflag = 8U * flag_words(static_cast<std::size_t>(G) * R);
    st.flag_t = m.alloc<std::uint8_t>(lflag);
    st.flag_r = m.alloc<std::uint8_t>(lflag);
    st.flag_hz = m.alloc<std::uint8_t>(lflag);
    st.flag_route = m.alloc<std::uint8_t>(rflag);
    st.route_mark = W == 64U ? nullptr : m.alloc<std::uint64_t>(rwords);
    st.mark_rb = W == 64U ? m.alloc<std::uint8_t>(cols) : nullptr;
    st.rt_valid = W == 64U ? m.alloc<std::uint64_t>(static_cast<std::size_t>(G) * K * L) : nullptr;
    st.bt_valid = W == 64U ? m.alloc<std::uint64_t>(static_cast<std::size_t>(G) * L) : nullptr;
    st.grp_active = m.alloc<std::uint64_t>(G);
    st.any_route = m.alloc<std::uint8_t>(static_cast<std::size_t>(G) * W);
    st.any_tmp = m.alloc<std::uint8_t>(static_cast<std::size_t>(G) * W);
    st.cap_t = static_cast<std::size_t>(G) * L;
    st.cap_r = static_cast<std::size_t>(G) * L;
    st.cap_route = static_cast<std::size_t>(G) * R;
    st.cap_hz = static_cast<std::size_t>(G) * L;
    st.list_t = m.alloc<std::uint64_t>(st.cap_t);
    st.list_r = m.alloc<std::uint64_t>(st.cap_r);
    st.list_route = m.alloc<std::uint64_t>(st.cap_route);
    // packed wavefronts: at most one per route list entry (route scan) or
    // per station list entry (transfers)
    st.cap_pack = W == 64U ? std::max(st.cap_route, st.cap_r) : 0U;
    st.packed_info = W == 64U ? m.alloc<std::uint64_t>(st.cap_pack) : nullptr;
    st.packed_lanes = W == 64U ? m.alloc<std::uint16_t>(st.cap_pack * 64U) : nullptr;
    // loose route items: only packed wavefronts of fewer than 33 queries go
    // loose, and each of their list entries adds fewer than 33 items
    st.cap_loose = W == 64U ? 32U * st.cap_route : 0U;
    st.loose = W == 64U ? m.alloc<std::uint64_t>(st.cap_loose) : nullptr;
    st.list_hz = m.alloc<std::uint64_t>(st.cap_hz);
    st.counts = m.alloc<std::uint32_t>(16U);
    st.work = m.alloc<unsigned long long>(kWorkCount);
    st.param = m.alloc<slot_param>(S);
    st.kslot = m.alloc<std::uint8_t>(S);
    st.tad = m.alloc<std::int16_t>(static_cast<std::size_t>(S) * kRowsMax);
    st.dest_ev = m.alloc<dest_event>(static_cast<std::size_t>(S) * kMaxDestEvents);
    st.tad0 = m.alloc<std::int16_t>(S);
    st.n_chg = m.alloc<std::uint32_t>(S);
    st.chg_pos = m.alloc<std::uint32_t>(static_cast<std::size_t>(S) * kMaxChanges);
    st.chg_tad = m.alloc<std::int16_t>(static_cast<std::size_t>(S) * kMaxChanges);
    st.dest = m.alloc<std::uint32_t>(static_cast<std::size_t>(S) * kMaxDestBits);
    st.dist_loc = m.alloc<std::uint32_t>(static_cast<std::size_t>(S) * kMaxDist);
    st.dist_min = m.alloc<std::uint16_t>(static_cast<std::size_t>(S) * kMaxDist);
    st.dest_rows = m.alloc<std::int16_t>(static_cast<std::size_t>(S) * kMaxDestBits * K);
    // zero the marks, masks, counters (one-time set-up, not on a hot path)
    HIP_CHECK(hipMemsetAsync(st.mark_t, 0, 8U * lwords, e->stream));
    HIP_CHECK(hipMemsetAsync(st.mark_r, 0, 8U * lwords, e->stream));
    HIP_CHECK(hipMemsetAsync(st.dest_mask, 0, 8U * lwords, e->stream));
    HIP_CHECK(hipMemsetAsync(st.hz_mask, 0, 8U * lwords, e->stream));
    HIP_CHECK(hipMemsetAsync(st.flag_t, 0, lflag, e->stream));
    HIP_CHECK(hipMemsetAsync(st.flag_r, 0, lflag, e->stream));
    HIP_CHECK(hipMemsetAsync(st.flag_hz, 0, lflag, e->stream));
    HIP_CHECK(hipMemsetAsync(st.flag_route, 0, rflag, e->stream));
    if (st.route_mark != nullptr) HIP_CHECK(hipMemsetAsync(st.route_mark, 0, 8U * rwords, e->stream));
    if (st.mark_rb != nullptr) HIP_CHECK(hipMemsetAsync(st.mark_rb, 0, cols, e->stream));
    if (st.rt_valid != nullptr) {
      HIP_CHECK(hipMemsetAsync(st.rt_valid, 0, 8U * static_cast<std::size_t>(G) * K * L, e->stream));
      HIP_CHECK(hipMemsetAsync(st.bt_valid, 0, 8U * static_cast<std::size_t>(G) * L, e->stream));
    }
    HIP_CHECK(hipMemsetAsync(st.grp_active, 0, 8U * G, e->stream));
    HIP_CHECK(hipMemsetAsync(st.any_route, 0, static_cast<std::size_t>(G) * W, e->stream));
    HIP_CHECK(hipMemsetAsync(st.any_tmp, 0, static_cast<std::size_t>(G) * W, e->stream));
    HIP_CHECK(hipMemsetAsync(st.counts, 0, 64U, e->stream));
    HIP_CHECK(hipMemsetAsync(st.work, 0, 8U * kWorkCount, e->stream));
    HIP_CHECK(hipMemsetAsync(st.n_chg, 0, 4U * S, e->stream));
    HIP_CHECK(hipMemsetAsync(st.kslot, 0, S, e->stream));
    e->h_param.assign(S, slot_param{});
    e->h_dest.assign(static_cast<std::size_t>(S) * kMaxDestBits, 0U);
    e->h_dist_loc.assign(static_cast<std::size_t>(S) * kMaxDist, 0U);
    e->h_dist_min.assign(static_cast<std::size_t>(S) * kMaxDist, 0U);
    e->slot_dest_bits.assign(S, {});
    e->d_reset_groups = m.alloc<std::uint32_t>(G);
    e->d_rt_mask = m.alloc<std::uint64_t>(G);
    e->d_bt_mask = m.alloc<std::uint64_t>(G);
    e->d_flags = m.alloc<std::uint8_t>(S);
    e->d_worst = m.alloc<std::int16_t>(S);
    e->d_sweep_groups = m.alloc<std::uint32_t>(G + 1U);
    e->d_slot_list = m.alloc<std::uint32_t>(S);
    e->d_lb_row = m.alloc<std::uint16_t>(L);
    e->d_ph_sel = m.alloc<std::uint32_t>(static_cast<std::size_t>(S) + G);
    if (lanes == 64U) {
      e->d_ph_slots = m.alloc<std::uint32_t>(S);
      e->d_ph_masks = m.alloc<std::uint64_t>(G);
      e->d_ph_rows = m.alloc<std::uint16_t>(static_cast<std::size_t>(std::min<std::uint32_t>(S, kPhastRows)) * L);
    }
    e->h_active.reserve(G);
    e->h_rows.reserve(static_cast<std::size_t>(S) * kMaxDestBits * K);
    e->h_counts.reserve(8U);
    e->pending_starts.assign(S, {});
    e->algo_new.assign(S, 0U);
    e->algo_fresh.assign(S, 0U);
    e->algo_dir.assign(S, 0U);
    e->slot_dir.assign(S, std::int8_t{-1});
    e->d_full_mask = m.alloc<std::uint64_t>(G);
    HIP_CHECK(hipStreamSynchronize(e->stream));
    *out = e.release();
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_engine_state_bytes(const nigiri_amd_engine* e, uint64_t* b) {
  if (e == nullptr || b == nullptr) return NIGIRI_AMD_EINVAL;
  *b = e->mem.bytes;
  return NIGIRI_AMD_OK;
}

void nigiri_amd_engine_destroy(nigiri_amd_engine* e) { delete e; }

int nigiri_amd_engine_configure(nigiri_amd_engine* e, const nigiri_engine_config* c) {
  if (e == nullptr || c == nullptr || c->admission_wait_ms < 0.0) return NIGIRI_AMD_EINVAL;
  e->admission_wait_ms = c->admission_wait_ms;
  if (c->grid_waves != 0U) e->grid_waves = c->grid_waves;
  e->use_graphs = c->eager_rounds == 0U;
  if (c->route_claim != 0U) e->tune.route_claim = c->route_claim;
  if (c->transfer_claim != 0U) e->tune.transfer_claim = c->transfer_claim;
  if (c->pack_max_entries != 0U) e->tune.pack_max_entries = std::min<std::uint32_t>(c->pack_max_entries, 64U);
  if (c->transfer_pack_max_entries != 0U) {
    e->tune.transfer_pack_max_entries = std::min<std::uint32_t>(c->transfer_pack_max_entries, 64U);
  }
  if (c->loose_lanes != 0U) e->tune.loose_lanes = std::min<std::uint32_t>(c->loose_lanes, 33U);
  e->admit_when_idle = c->admit_when_idle != 0U;
  e->state_budget = c->state_budget_bytes;
  return NIGIRI_AMD_OK;
}

namespace nigiri_amd {
namespace {

// search::execute for every query (search.h:250-340): host preparation,
// then the engine runs the forward queries and then the backward ones.
// Returns the wall time of the engine part in milliseconds.
double run_search(nigiri_amd_engine& e, const nigiri_query* qs, std::uint32_t n, std::vector<prepared_query>& pq,
                  std::vector<int>& status, std::vector<std::string>& errors) {
  auto const& tt = *e.img->tt;
  // fastest direct connections on the device; lower bounds: upward searches
  // on the host, sweeps on the device
  std::vector<std::int16_t> fd;
  std::vector<std::uint8_t> fd_have;
  device_fastest_direct(*e.img, qs, n, fd, fd_have);
  prepare_all(tt, qs, n, pq, status, errors, false, &fd, &fd_have);
  auto const t0 = std::chrono::steady_clock::now();
  for (int dir = 0; dir != 2; ++dir) {
    std::vector<std::int32_t> queue;
    for (std::uint32_t i = 0; i != n; ++i) {
      if (status[i] == NIGIRI_AMD_OK && (qs[i].dir == NIGIRI_AMD_DIR_FWD) == (dir == 0)) {
        queue.push_back(static_cast<std::int32_t>(i));
      }
    }
    if (queue.empty()) continue;
    std::size_t head = 0U;
    engine_run_dir(
        e, dir == 0, pq, [&]() { return head < queue.size() ? queue[head++] : -1; }, [](std::int32_t) {},
        [&](std::int32_t q, int code, std::string const& msg) {
          status[static_cast<std::size_t>(q)] = code;
          errors[static_cast<std::size_t>(q)] = msg;
        });
  }
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace
}  // namespace nigiri_amd

int nigiri_amd_engine_get_config(const nigiri_amd_engine* e, nigiri_engine_config* out) {
  if (e == nullptr || out == nullptr) return NIGIRI_AMD_EINVAL;
  *out = nigiri_engine_config{};
  out->admission_wait_ms = e->admission_wait_ms;
  out->grid_waves = e->grid_waves;
  out->eager_rounds = e->use_graphs ? 0U : 1U;
  out->route_claim = e->tune.route_claim;
  out->transfer_claim = e->tune.transfer_claim;
  out->pack_max_entries = e->tune.pack_max_entries;
  out->transfer_pack_max_entries = e->tune.transfer_pack_max_entries;
  out->loose_lanes = e->tune.loose_lanes;
  out->admit_when_idle = e->admit_when_idle ? 1U : 0U;
  out->state_budget_bytes = e->state_budget;
  return NIGIRI_AMD_OK;
}

int nigiri_amd_engine_autotune(nigiri_amd_engine* e, const nigiri_query* qs, uint32_t n, uint32_t repeats,
                               nigiri_engine_config* out, nigiri_tune_trial* trials, uint32_t trial_capacity,
                               uint32_t* n_trials) {
  if (e == nullptr || qs == nullptr || n == 0U || repeats == 0U || (trial_capacity != 0U && trials == nullptr)) {
    set_error("autotune: invalid argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(e->img->device));
    int cus = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->img->device));
    auto const c = static_cast<std::uint32_t>(std::max(1, cus));
    struct knob {
      std::uint32_t id;
      std::uint32_t* field;
      std::vector<std::uint32_t> values;
    };
    // knob ids as in nigiri_tune_trial: 0 grid_waves, 1 route_claim,
    // 2 transfer_claim, 3 pack_max_entries, 4 transfer_pack_max_entries,
    // 5 loose_lanes
    std::vector<knob> knobs{{0U, &e->grid_waves, {8U * c, 16U * c, 32U * c, 64U * c}},
                            {1U, &e->tune.route_claim, {1U, 2U, 4U}},
                            {2U, &e->tune.transfer_claim, {4U, 16U, 64U}}};
    if (e->W == 64U) {  // packing and device lower bounds: batched engine only
      knobs.push_back({3U, &e->tune.pack_max_entries, {1U, 2U, 4U, 8U, 64U}});
      knobs.push_back({4U, &e->tune.transfer_pack_max_entries, {1U, 2U, 4U, 8U, 64U}});
      knobs.push_back({5U, &e->tune.loose_lanes, {1U, 9U, 17U, 33U}});
    }
    std::uint32_t nt = 0U;
    std::vector<prepared_query> pq;
    std::vector<int> status;
    std::vector<std::string> errors;
    // one untimed run: graphs captured, caches and allocations warm
    (void)run_search(*e, qs, n, pq, status, errors);
    // coordinate descent, one pass: each knob takes its fastest value with
    // the knobs before it already set (median of `repeats` runs)
    for (auto& k : knobs) {
      auto best_v = *k.field;
      double best_ms = 0.0;
      bool first = true;
      for (auto const v : k.values) {
        *k.field = v;
        std::vector<double> ms;
        for (std::uint32_t r = 0; r != repeats; ++r) ms.push_back(run_search(*e, qs, n, pq, status, errors));
        std::sort(ms.begin(), ms.end());
        auto const med = ms[ms.size() / 2U];
        if (nt < trial_capacity) trials[nt] = nigiri_tune_trial{k.id, v, med};
        ++nt;
        if (first || med < best_ms) {
          best_ms = med;
          best_v = v;
          first = false;
        }
      }
      *k.field = best_v;
    }
    if (n_trials != nullptr) *n_trials = nt;
    if (out != nullptr) (void)nigiri_amd_engine_get_config(e, out);
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_engine_profile(nigiri_amd_engine* e, int enable) {
  if (e == nullptr) return NIGIRI_AMD_EINVAL;
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(e->img->device));
    if (enable != 0 && e->events.empty()) {
      e->events.resize(4U * (kRowsMax + 1U));
      for (auto& ev : e->events) HIP_CHECK(hipEventCreate(&ev));
    }
    e->profile = enable != 0;
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_engine_stats(nigiri_amd_engine* e, nigiri_engine_stats* out, int reset) {
  if (e == nullptr || out == nullptr) return NIGIRI_AMD_EINVAL;
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(e->img->device));
    HIP_CHECK(hipStreamSynchronize(e->stream));
    unsigned long long w[kWorkCount];
    HIP_CHECK(hipMemcpy(w, e->st.work, sizeof(w), hipMemcpyDeviceToHost));
    *out = e->stats;
    out->graph_captures = round_graphs_captures(e->graphs);
    for (int i = 0; i != kWorkCount; ++i) out->work[i] = w[i];
    if (reset != 0) {
      e->stats = nigiri_engine_stats{};
      HIP_CHECK(hipMemset(e->st.work, 0, sizeof(w)));
    }
    return NIGIRI_AMD_OK;
  });
}


int nigiri_amd_search(nigiri_amd_engine* e, const nigiri_query* qs, uint32_t n, nigiri_result* res,
                      nigiri_journey* journeys, uint32_t journey_capacity, nigiri_leg* legs,
                      uint32_t leg_capacity) {
  if (e == nullptr || (n != 0U && (qs == nullptr || res == nullptr))) {
    set_error("search: null argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(e->img->device));
    std::vector<prepared_query> pq;
    std::vector<int> status;
    std::vector<std::string> errors;
    run_search(*e, qs, n, pq, status, errors);
    auto const w = write_results(pq, status, res, journeys, journey_capacity, legs, leg_capacity,
                                 nullptr, nullptr);
    if (w == ~0U) {
      set_error("search: journey_capacity or leg_capacity too small");
      return NIGIRI_AMD_ECAPACITY;
    }
    for (std::uint32_t i = 0; i != n; ++i) {
      if (status[i] != NIGIRI_AMD_OK) {
        set_error(errors[i]);
      }
    }
    return NIGIRI_AMD_OK;
  });
}

// ---- serving loop -----------------------------------------------------------
// Queries become visible at their arrival time; a pool of host workers
// prepares them (destinations, lower bounds, start labels) as they arrive;
// the GPU loop admits prepared queries into free slots at every epoch
// boundary while other slots keep running (continuous batching).
int nigiri_amd_serve(nigiri_amd_engine* e, const nigiri_query* qs, const double* arrival_ms, uint32_t n,
                     nigiri_result* res, nigiri_journey* journeys, uint32_t journey_capacity,
                     nigiri_leg* legs, uint32_t leg_capacity, double* wall_ms) {
  if (e == nullptr || (n != 0U && (qs == nullptr || arrival_ms == nullptr || res == nullptr))) {
    set_error("serve: null argument");
    return NIGIRI_AMD_EINVAL;
  }
  for (uint32_t i = 0; i != n; ++i) {
    if (qs[i].dir != qs[0].dir) {
      set_error("serve: all queries of one serving run must share a direction");
      return NIGIRI_AMD_EUNSUPPORTED;
    }
    if (i != 0U && arrival_ms[i] < arrival_ms[i - 1U]) {
      set_error("serve: arrival times must be non-decreasing");
      return NIGIRI_AMD_EINVAL;
    }
  }
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(e->img->device));
    auto const& tt = *e->img->tt;
    using clock = std::chrono::steady_clock;
    std::vector<prepared_query> pq(n);
    std::vector<int> status(n, NIGIRI_AMD_OK);
    std::vector<double> latency(n, 0.0), queue(n, 0.0);
    std::mutex mu;
    std::deque<std::int32_t> ready;
    std::atomic<std::uint32_t> next_prep{0U};
    std::atomic<std::uint32_t> n_done{0U};
    std::atomic<std::uint32_t> n_prepared_ctr{0U};
    std::atomic<std::uint32_t> n_failed{0U};
    bool const fwd = qs[0].dir == NIGIRI_AMD_DIR_FWD;
    auto const t0 = clock::now();
    auto const ms_since = [&](clock::time_point t) {
      return std::chrono::duration<double, std::milli>(t - t0).count();
    };
    auto const hw = std::max(1U, std::thread::hardware_concurrency());
    auto const n_workers = std::min<std::uint32_t>(std::max(1U, hw - 1U), 32U);
    std::vector<std::thread> workers;
    for (std::uint32_t w = 0; w != n_workers; ++w) {
      workers.emplace_back([&]() {
        host_scratch scratch;
        while (true) {
          auto const i = next_prep.fetch_add(1U);
          if (i >= n) break;
          auto const due = t0 + std::chrono::duration_cast<clock::duration>(
                                    std::chrono::duration<double, std::milli>(arrival_ms[i]));
          std::this_thread::sleep_until(due);
          try {
            prepare(tt, qs[i], pq[i], scratch, false);
          } catch (status_error const& ex) {
            status[i] = ex.code_;
          } catch (std::exception const&) {
            status[i] = NIGIRI_AMD_EINVAL;
          }
          if (status[i] != NIGIRI_AMD_OK) {
            latency[i] = ms_since(clock::now()) - arrival_ms[i];
            n_done.fetch_add(1U);
            n_failed.fetch_add(1U);
            n_prepared_ctr.fetch_add(1U);
            continue;
          }
          {
            std::lock_guard<std::mutex> lock{mu};
            ready.push_back(static_cast<std::int32_t>(i));
          }
          n_prepared_ctr.fetch_add(1U);
        }
      });
    }
    runner r{*e, tt, fwd, pq};
    r.next = [&]() -> std::int32_t {
      std::lock_guard<std::mutex> lock{mu};
      if (ready.empty()) return -1;
      auto const q = ready.front();
      ready.pop_front();
      return q;
    };
    r.on_done = [&](std::int32_t q) {
      latency[static_cast<std::size_t>(q)] = ms_since(clock::now()) - arrival_ms[q];
      n_done.fetch_add(1U);
    };
    r.on_fail = [&](std::int32_t q, int code, std::string const&) {
      status[static_cast<std::size_t>(q)] = code;
      latency[static_cast<std::size_t>(q)] = ms_since(clock::now()) - arrival_ms[q];
      n_done.fetch_add(1U);
    };
    r.on_admit = [&](std::int32_t q) {
      queue[static_cast<std::size_t>(q)] = ms_since(clock::now()) - arrival_ms[q];
    };
    std::atomic<std::uint32_t>& n_prepared = n_prepared_ctr;
    try {
      while (n_done.load() < n) {
        // admission window: let arrived queries finish preparation so that
        // free slots start the next epoch filled
        auto const free = r.free_slots();
        auto const deadline = clock::now() + std::chrono::duration_cast<clock::duration>(
                                                  std::chrono::duration<double, std::milli>(e->admission_wait_ms));
        // (whole-batch admission: nothing is admitted while slots are busy)
        while (!(e->admit_when_idle && r.any_busy()) && clock::now() < deadline) {
          std::size_t n_ready = 0U;
          std::uint32_t n_settled = 0U;  // arrived queries that cannot get more ready
          {
            std::lock_guard<std::mutex> lock{mu};
            n_ready = ready.size();
            n_settled = n_prepared.load();
          }
          auto const now_ms = ms_since(clock::now());
          auto const arrived = static_cast<std::uint32_t>(
              std::upper_bound(arrival_ms, arrival_ms + n, now_ms) - arrival_ms);
          if (n_ready >= free || n_settled >= arrived) break;
          std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        if (!r.step()) {
          std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
      }
    } catch (...) {
      next_prep.store(n);
      for (auto& t : workers) t.join();
      throw;
    }
    for (auto& t : workers) t.join();
    if (wall_ms != nullptr) *wall_ms = ms_since(clock::now());
    auto const w = write_results(pq, status, res, journeys, journey_capacity, legs, leg_capacity,
                                 nullptr, &latency, &queue);
    if (w == ~0U) {
      set_error("serve: journey_capacity too small");
      return NIGIRI_AMD_ECAPACITY;
    }
    return NIGIRI_AMD_OK;
  });
}

// ---- Algo-level entry points (one slot) -----------------------------------

int nigiri_amd_algo_init(nigiri_amd_engine* e, uint32_t slot, const nigiri_algo_init* init) {
  if (e == nullptr || init == nullptr || slot >= e->S || init->lb == nullptr) {
    set_error("algo_init: bad argument");
    return NIGIRI_AMD_EINVAL;
  }
  if (init->n_dest > kMaxDestBits || init->n_dist > kMaxDist ||
      (init->n_dest != 0U && init->dest == nullptr) || (init->n_dist != 0U && init->dist_loc == nullptr)) {
    set_error("algo_init: too many destinations");
    return NIGIRI_AMD_EUNSUPPORTED;
  }
  {
    std::vector<std::uint32_t> d(init->intermodal != 0 ? init->dist_loc : init->dest,
                                 (init->intermodal != 0 ? init->dist_loc : init->dest) +
                                     (init->intermodal != 0 ? init->n_dist : init->n_dest));
    auto const L = e->img->d.L;
    if (std::any_of(init->dest, init->dest + init->n_dest, [&](uint32_t l) { return l >= L; }) ||
        std::any_of(init->dist_loc, init->dist_loc + init->n_dist, [&](uint32_t l) { return l >= L; })) {
      set_error("algo_init: destination out of range");
      return NIGIRI_AMD_EINVAL;
    }
    if (dest_relaxations(*e->img->tt, init->dir == NIGIRI_AMD_DIR_FWD, init->intermodal != 0, d) >
        kMaxDestEvents) {
      set_error("algo_init: destinations reachable by too many transfers and footpaths in one round");
      return NIGIRI_AMD_EUNSUPPORTED;
    }
  }
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(e->img->device));
    auto& sp = e->h_param[slot];
    sp.base = static_cast<std::uint16_t>(init->base_day);
    sp.intermodal = init->intermodal != 0 ? 1U : 0U;
    sp.n_dest = static_cast<std::uint16_t>(init->n_dest);
    sp.n_dist = static_cast<std::uint16_t>(init->n_dist);
    sp.tts_default = init->tts_default != 0 ? 1U : 0U;
    sp.tts_min = init->tts_min_transfer_time;
    sp.tts_add = init->tts_additional_time;
    sp.tts_factor = init->tts_factor;
    sp.end_k = static_cast<std::uint8_t>(e->K);
    sp.bounds_last_k = 0U;
    for (std::uint32_t j = 0; j != init->n_dest; ++j) e->h_dest[slot * kMaxDestBits + j] = init->dest[j];
    for (std::uint32_t j = 0; j != init->n_dist; ++j) {
      e->h_dist_loc[slot * kMaxDist + j] = init->dist_loc[j];
      e->h_dist_min[slot * kMaxDist + j] = init->dist_min[j];
    }
    e->algo_dir[slot] = static_cast<std::uint8_t>(init->dir);
    // destination masks and lower bounds
    epoch_args a{};
    a.st = e->st;
    std::vector<std::uint64_t> ops;
    std::vector<std::uint8_t> set;
    for (auto const l : e->slot_dest_bits[slot]) {
      ops.push_back((static_cast<std::uint64_t>(slot) << 32) | l);
      set.push_back(0U);
    }
    e->slot_dest_bits[slot].clear();
    if (init->intermodal == 0) {
      for (std::uint32_t j = 0; j != init->n_dest; ++j) {
        ops.push_back((static_cast<std::uint64_t>(slot) << 32) | init->dest[j]);
        set.push_back(1U);
        e->slot_dest_bits[slot].push_back(init->dest[j]);
      }
    }
    if (ops.size() > e->dest_ops_cap) {
      e->dest_ops_cap = ops.size() * 2U;
      e->d_dest_ops = e->grow(e->d_dest_ops, e->dest_ops_cap);
      e->d_dest_set = e->grow(e->d_dest_set, e->dest_ops_cap);
    }
    // the caller's lower bounds (search.h:114-131 computes them) into the
    // slot's column
    HIP_CHECK(hipMemcpyAsync(e->d_lb_row, init->lb, 2U * e->st.L, hipMemcpyHostToDevice, e->stream));
    launch_lb_put(e->W, e->st, slot, e->d_lb_row, e->stream);
    if (!ops.empty()) {
      HIP_CHECK(hipMemcpyAsync(e->d_dest_ops, ops.data(), 8U * ops.size(), hipMemcpyHostToDevice, e->stream));
      HIP_CHECK(hipMemcpyAsync(e->d_dest_set, set.data(), set.size(), hipMemcpyHostToDevice, e->stream));
    }
    a.dest_ops = e->d_dest_ops;
    a.dest_set = e->d_dest_set;
    a.n_dest_ops = static_cast<std::uint32_t>(ops.size());
    a.n_slots = 0U;  // no slot init
    launch_epoch_begin(init->dir == NIGIRI_AMD_DIR_FWD, e->W, a, e->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(e->stream));
    e->algo_new[slot] = 1U;
    e->algo_fresh[slot] = 1U;
    e->pending_starts[slot].clear();
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_algo_reset_arrivals(nigiri_amd_engine* e, uint32_t slot) {
  if (e == nullptr || slot >= e->S) return NIGIRI_AMD_EINVAL;
  e->algo_new[slot] = 1U;
  return NIGIRI_AMD_OK;
}

int nigiri_amd_algo_next_start_time(nigiri_amd_engine* e, uint32_t slot) {
  if (e == nullptr || slot >= e->S) return NIGIRI_AMD_EINVAL;
  e->algo_fresh[slot] = 1U;
  e->pending_starts[slot].clear();
  return NIGIRI_AMD_OK;
}

int nigiri_amd_algo_add_start(nigiri_amd_engine* e, uint32_t slot, uint32_t location, int32_t t) {
  if (e == nullptr || slot >= e->S || location >= e->st.L) return NIGIRI_AMD_EINVAL;
  e->pending_starts[slot].emplace_back(location, t);
  return NIGIRI_AMD_OK;
}

int nigiri_amd_algo_fill_bounds(nigiri_amd_engine* e, uint32_t src_slot, uint32_t dst_slot,
                                uint32_t n_rows) {
  if (e == nullptr || src_slot >= e->S || dst_slot >= e->S || n_rows == 0U || n_rows > e->K) {
    set_error("algo_fill_bounds: bad argument");
    return NIGIRI_AMD_EINVAL;
  }
  if (e->W != 1U) {
    set_error("algo_fill_bounds: bounded searches run on the plain engine (lanes == 1)");
    return NIGIRI_AMD_EUNSUPPORTED;
  }
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(e->img->device));
    launch_fill_bounds(e->algo_dir[src_slot] == NIGIRI_AMD_DIR_FWD, e->st, src_slot, dst_slot, n_rows,
                       e->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(e->stream));
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_algo_set_bounds(nigiri_amd_engine* e, uint32_t slot, uint32_t last_round) {
  if (e == nullptr || slot >= e->S || last_round >= e->K) {
    set_error("algo_set_bounds: bad argument");
    return NIGIRI_AMD_EINVAL;
  }
  if (e->W != 1U && last_round != 0U) {
    set_error("algo_set_bounds: bounded searches run on the plain engine (lanes == 1)");
    return NIGIRI_AMD_EUNSUPPORTED;
  }
  e->h_param[slot].bounds_last_k = static_cast<std::uint8_t>(last_round);
  return NIGIRI_AMD_OK;
}

int nigiri_amd_algo_execute(nigiri_amd_engine* e, uint32_t slot, int32_t start_time,
                            uint8_t max_transfers, int32_t worst_time_at_dest, int16_t* dest_times,
                            uint32_t* n_rounds) {
  if (e == nullptr || slot >= e->S) return NIGIRI_AMD_EINVAL;
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(e->img->device));
    auto const& tt = *e->img->tt;
    auto const fwd = e->algo_dir[slot] == NIGIRI_AMD_DIR_FWD;
    auto& sp = e->h_param[slot];
    auto const end_k = std::min<std::uint32_t>(std::min<std::uint32_t>(max_transfers, 14U) + 2U, e->K);
    if (sp.bounds_last_k != 0U && end_k - 1U > sp.bounds_last_k) {
      // within_bounds reads bounds row last_k - k (raptor.h:466-469)
      throw status_error{NIGIRI_AMD_EINVAL, "algo_execute: rounds beyond the bounds' last round"};
    }
    sp.end_k = static_cast<std::uint8_t>(end_k);
    auto const base = tt.internal_from + static_cast<std::int32_t>(sp.base) * 1440;
    (void)start_time;
    auto const S = e->S, G = e->G;
    std::vector<std::uint64_t> rt_mask(G, 0U), bt_mask(G, 0U);
    std::vector<std::uint8_t> flags(S, 0U);
    std::vector<std::int16_t> worst(S, 0);
    auto const g = slot / e->W, lane = slot % e->W;
    if (e->algo_new[slot] != 0U) rt_mask[g] |= 1ULL << lane;
    if (e->algo_fresh[slot] != 0U) bt_mask[g] |= 1ULL << lane;
    flags[slot] = static_cast<std::uint8_t>(1U | (e->algo_new[slot] != 0U ? 2U : 0U));
    worst[slot] = clamp16(static_cast<long long>(worst_time_at_dest) - base);
    std::vector<std::uint32_t> st_slot, st_loc;
    std::vector<std::int16_t> st_delta;
    for (auto const& [l, t] : e->pending_starts[slot]) {
      st_slot.push_back(slot);
      st_loc.push_back(l);
      st_delta.push_back(clamp16(static_cast<long long>(t) - base));
    }
    if (st_slot.size() > e->start_cap) {
      e->start_cap = st_slot.size() * 2U;
      e->d_start_slot = e->grow(e->d_start_slot, e->start_cap);
      e->d_start_loc = e->grow(e->d_start_loc, e->start_cap);
      e->d_start_delta = e->grow(e->d_start_delta, e->start_cap);
    }
    auto const stream = e->stream;
    auto const& tr = e->st;
    HIP_CHECK(hipMemcpyAsync(tr.param, e->h_param.data(), sizeof(slot_param) * S, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(tr.dest, e->h_dest.data(), 4U * e->h_dest.size(), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(tr.dist_loc, e->h_dist_loc.data(), 4U * e->h_dist_loc.size(), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(tr.dist_min, e->h_dist_min.data(), 2U * e->h_dist_min.size(), hipMemcpyHostToDevice, stream));
    std::uint32_t const gg = g;
    HIP_CHECK(hipMemcpyAsync(e->d_reset_groups, &gg, 4U, hipMemcpyHostToDevice, stream));
    auto const full_mask = dense_resets(*e, {slot}, fwd);
    HIP_CHECK(hipMemcpyAsync(e->d_full_mask, full_mask.data(), 8U * G, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(e->d_rt_mask, rt_mask.data(), 8U * G, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(e->d_bt_mask, bt_mask.data(), 8U * G, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(e->d_flags, flags.data(), S, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(e->d_worst, worst.data(), 2U * S, hipMemcpyHostToDevice, stream));
    if (!st_slot.empty()) {
      HIP_CHECK(hipMemcpyAsync(e->d_start_slot, st_slot.data(), 4U * st_slot.size(), hipMemcpyHostToDevice, stream));
      HIP_CHECK(hipMemcpyAsync(e->d_start_loc, st_loc.data(), 4U * st_loc.size(), hipMemcpyHostToDevice, stream));
      HIP_CHECK(hipMemcpyAsync(e->d_start_delta, st_delta.data(), 2U * st_delta.size(), hipMemcpyHostToDevice, stream));
    }
    epoch_args a{};
    a.st = tr;
    a.invalid = fwd ? SHRT_MAX : SHRT_MIN;
    a.reset_groups = e->d_reset_groups;
    a.n_reset_groups = (rt_mask[g] | bt_mask[g] | full_mask[g]) != 0U ? 1U : 0U;
    a.reset_rt_mask = e->d_rt_mask;
    a.reset_bt_mask = e->d_bt_mask;
    a.reset_full_mask = e->d_full_mask;
    a.start_slot = e->d_start_slot;
    a.start_loc = e->d_start_loc;
    a.start_delta = e->d_start_delta;
    a.n_starts = static_cast<std::uint32_t>(st_slot.size());
    a.slot_flags = e->d_flags;
    a.slot_worst = e->d_worst;
    a.n_slots = S;
    launch_epoch_begin(fwd, e->W, a, stream);
    HIP_CHECK(hipGetLastError());
    // rounds until the slot's epoch ends (at most end_k - 1)
    std::uint32_t rounds = 0U;
    round_args ra{};
    ra.im = e->img->d;
    ra.st = tr;
    ra.tune = e->tune;
    ra.invalid = a.invalid;
    // a slot after its first epoch sees earlier rows in round_times
    bool const has_sweep = e->algo_new[slot] == 0U;
    std::uint32_t const sw[2] = {1U, g};
    if (has_sweep) {
      HIP_CHECK(hipMemcpyAsync(e->d_sweep_groups, sw, sizeof(sw), hipMemcpyHostToDevice, stream));
    }
    while (true) {
      launch_round(fwd, e->W, ra, e->d_sweep_groups, has_sweep, S, 2048, stream);
      HIP_CHECK(hipGetLastError());
      ++rounds;
      HIP_CHECK(hipMemcpyAsync(e->h_active.p, tr.grp_active, 8U * G, hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      if ((e->h_active.p[g] & (1ULL << lane)) == 0U) break;
      if (rounds > e->K) throw status_error{NIGIRI_AMD_EDEVICE, "epoch did not end"};
    }
    std::uint32_t const one = slot;
    HIP_CHECK(hipMemcpyAsync(e->d_slot_list, &one, 4U, hipMemcpyHostToDevice, stream));
    launch_extract(e->W, tr, e->d_slot_list, 1U, stream);
    auto const rows = static_cast<std::size_t>(kMaxDestBits) * e->K;
    HIP_CHECK(hipMemcpyAsync(e->h_rows.p, tr.dest_rows, 2U * rows, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipMemcpyAsync(e->h_counts.p, tr.counts, 32U, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if (e->h_counts.p[4] != 0U) {
      std::uint32_t zero = 0U;
      HIP_CHECK(hipMemcpy(tr.counts + 4, &zero, 4U, hipMemcpyHostToDevice));
      throw status_error{NIGIRI_AMD_ECAPACITY, "device work list overflow"};
    }
    if (dest_times != nullptr) {  // the slot's destination rows: [n_dest_bits][K]
      auto const& sp = e->h_param[slot];
      auto const bits = static_cast<std::size_t>(sp.intermodal != 0U ? 1U : sp.n_dest);
      std::memcpy(dest_times, e->h_rows.p, 2U * bits * e->K);
    }
    if (n_rounds != nullptr) *n_rounds = rounds;
    e->algo_new[slot] = 0U;
    e->algo_fresh[slot] = 0U;
    e->pending_starts[slot].clear();
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_algo_reconstruct(nigiri_amd_engine* e, uint32_t slot, const nigiri_query* q,
                                nigiri_journey* j, nigiri_leg* legs, uint32_t leg_capacity) {
  if (e == nullptr || q == nullptr || j == nullptr || slot >= e->S || j->dest >= e->st.L ||
      j->transfers + 2U > e->K || (legs == nullptr && leg_capacity != 0U)) {
    set_error("algo_reconstruct: bad argument");
    return NIGIRI_AMD_EINVAL;
  }
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(e->img->device));
    auto const& tt = *e->img->tt;
    auto const& sp = e->h_param[slot];
    prepared_query p;
    p.fwd_ = q->dir == NIGIRI_AMD_DIR_FWD;
    p.ontrip_ = q->is_interval == 0;
    p.intermodal_dest_ = sp.intermodal != 0U;
    p.start_intermodal_ = q->start_match_mode == NIGIRI_AMD_MATCH_INTERMODAL;
    p.base_ = sp.base;
    p.tts_ = host_tts{sp.tts_default != 0U, sp.tts_min, sp.tts_add, sp.tts_factor};
    for (auto i = 0U; i != q->n_start; ++i) p.start_offsets_.emplace_back(q->start_loc[i], q->start_dur[i]);
    for (auto i = 0U; i != q->n_dest; ++i) p.dest_offsets_.emplace_back(q->dest_loc[i], q->dest_dur[i]);
    auto const by_dur = [](auto const& a, auto const& b) { return a.second < b.second; };
    std::stable_sort(p.start_offsets_.begin(), p.start_offsets_.end(), by_dur);
    std::stable_sort(p.dest_offsets_.begin(), p.dest_offsets_.end(), by_dur);
    std::vector<std::uint32_t> off_loc;
    std::vector<std::int16_t> off_dur;
    if (p.intermodal_dest_) {
      for (auto const& [loc, dur] : p.dest_offsets_) {
        off_loc.push_back(loc);
        off_dur.push_back(dur);
      }
    }
    std::vector<rc_job> jobs{rc_job{slot, j->dest, 0U, static_cast<std::uint16_t>(off_loc.size()),
                                    j->transfers, sp.intermodal}};
    run_reconstruct(*e, p.fwd_, jobs, off_loc, off_dur);
    host_journey hj{j->start_time, j->dest_time, j->dest, j->transfers};
    assemble_legs(tt, p, e->h_rc_out[0], hj);
    j->leg_offset = 0U;
    j->n_legs = 0U;
    if (hj.error_) {
      set_error("algo_reconstruct: reconstruction failed");
      return NIGIRI_AMD_ERECONSTRUCT;
    }
    if (hj.legs_.size() > leg_capacity) {
      set_error("algo_reconstruct: leg_capacity too small");
      return NIGIRI_AMD_ECAPACITY;
    }
    std::copy(hj.legs_.begin(), hj.legs_.end(), legs);
    j->n_legs = static_cast<std::uint32_t>(hj.legs_.size());
    return NIGIRI_AMD_OK;
  });
}

int nigiri_amd_algo_download(nigiri_amd_engine* e, uint32_t slot, int16_t* round_times, int16_t* best,
                             int16_t* tmp, int16_t* time_at_dest, uint16_t* lb) {
  if (e == nullptr || slot >= e->S) return NIGIRI_AMD_EINVAL;
  return guarded([&]() {
    HIP_CHECK(hipSetDevice(e->img->device));
    auto const L = e->st.L, K = e->K;
    std::int16_t* d = nullptr;
    auto const rows = static_cast<std::size_t>(K + 3U);
    HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&d), 2U * rows * L));
    std::vector<std::int16_t> h(rows * L);
    try {
      launch_gather_slot(e->W, e->st, slot, d, d + static_cast<std::size_t>(K) * L,
                         d + static_cast<std::size_t>(K + 1U) * L,
                         reinterpret_cast<std::uint16_t*>(d + static_cast<std::size_t>(K + 2U) * L),
                         e->stream);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpyAsync(h.data(), d, 2U * h.size(), hipMemcpyDeviceToHost, e->stream));
      HIP_CHECK(hipStreamSynchronize(e->stream));
    } catch (...) {
      (void)hipFree(d);
      throw;
    }
    HIP_CHECK(hipFree(d));
    if (round_times != nullptr) std::memcpy(round_times, h.data(), 2U * static_cast<std::size_t>(K) * L);
    if (best != nullptr) std::memcpy(best, h.data() + static_cast<std::size_t>(K) * L, 2U * L);
    if (tmp != nullptr) std::memcpy(tmp, h.data() + static_cast<std::size_t>(K + 1U) * L, 2U * L);
    if (lb != nullptr) std::memcpy(lb, h.data() + static_cast<std::size_t>(K + 2U) * L, 2U * L);
    if (time_at_dest != nullptr) {
      HIP_CHECK(hipMemcpy(time_at_dest, e->st.tad + static_cast<std::size_t>(slot) * kRowsMax,
                          2U * kRowsMax, hipMemcpyDeviceToHost));
    }
    return NIGIRI_AMD_OK;
  });
}

}  // extern "C"
