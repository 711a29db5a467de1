// server.hip — request-serving loop with continuous batching (no reference
// counterpart; the reference runs one instance per executor call,
// benchmark/smp_engine_exec.h:6-32).
//
// Requests are admitted in arrival order when the HBM arena has room for their
// state (int16 rank matrix, plus a copy of the lists for host-submitted
// requests).  Where each goes:
//   - n >= large_n: to large_workers threads, each with a single-instance
//     pipeline of its own (engine.hpp: chains over the whole GPU, then the host
//     tail), so a large congested request does not hold a lane for its whole
//     serial chain;
//   - smaller ones, with small_to_idle_worker, to a worker that is idle while no
//     request waits for the workers: at low load a small request then gets the
//     same pipeline as in plain mode, whose host tail walks the last chains
//     faster than one workgroup can;
//   - the rest to the launch lanes (streams): admitted requests are launched
//     together on whichever lane is free, so a request that arrives while
//     others are on the GPU starts as soon as a lane and memory free up, and its
//     kernels overlap with the ones already running.  A launch whose requests
//     all have n <= kFusedLaneN is one kernel, k_serve_fused (one workgroup per
//     request: shared-list check, rank rows, chains, matching); a larger one is
//     k_serve_prep, a flat k_rank_rows16 pass over all its women's rows (one
//     wave per row, so a large request is spread over the whole GPU) and
//     k_serve_chains (one workgroup per request).  Results, proposal counts,
//     error bits and a
//     completion flag go straight into pinned host memory; the server loop
//     polls the flag.
// The plain mode is the baseline: one request at a time through the plain
// single-instance pipeline.

#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "engine.hpp"
#include "wg_solver.hpp"

namespace pactae {

struct ServeReq {
  const int32_t* pm;
  const int32_t* pw;
  int16_t* rank;
  int32_t* match;  // pinned host memory (device-visible)
  int n;
  int row0;        // first flat row index of this request in the launch
};

// The launch's request descriptors, written by the host into pinned memory, are
// copied to device memory (the rank pass looks a row's request up by binary
// search over them), and the error words are cleared.  One small kernel in place
// of a copy and a fill: each copy on a lane's stream is a switch between the
// compute queue and a copy engine.
__global__ __launch_bounds__(256) void k_serve_prep(const ServeReq* __restrict__ host_reqs, ServeReq* reqs,
                                                    int count, int* errs) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) {
    reqs[i] = host_reqs[i];
    errs[i] = 0;
  }
}

// One workgroup per request.  The matching goes straight into pinned host memory
// (q.match), and so do the proposal count and the error bits (the rank pass's,
// from errs, with the chains').  Each workgroup releases its host writes at system
// scope; the last one to finish (counted in *done, which it resets) then writes
// seq into the host's flag word, so a host that sees seq sees every result.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 8) void k_serve_chains(const ServeReq* __restrict__ reqs, int count,
                                                        const int* __restrict__ errs,
                                                        unsigned long long* h_props, int* h_errs, int* done,
                                                        unsigned* h_flag, unsigned seq) {
  extern __shared__ unsigned long long partner[];
  __shared__ int cursor, sbad;
  __shared__ unsigned long long red[BLOCK / 64];
  const int i = blockIdx.x;
  if (threadIdx.x == 0) sbad = 0;
  const ServeReq q = reqs[i];
  InstanceRef ref{q.pm, q.pw, q.rank, q.match, q.n};
  unsigned long long p = 0;
  const int bad = wg_chains_g<BLOCK, 16, 8, false>(ref, partner, &cursor, red, &p);
  if (bad) atomicOr(&sbad, bad);
  __syncthreads();
  if (threadIdx.x == 0) {
    h_props[i] = p;
    h_errs[i] = sbad | __hip_atomic_load(&errs[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0 && atomicAdd(done, 1) == count - 1) {
    *done = 0;
    __hip_atomic_store(h_flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Launches whose requests all have at most kFusedLaneN men: each request in one
// workgroup, by the batched path's per-instance solve (wg_solve_instance: shared
// men's list check, women's rank rows staged in LDS and written where the
// workgroup's own chains read them, Linv for a shared list, chains of one wave
// each with the women's words in LDS, matching), one kernel instead of a flat
// rank pass and a chain kernel.  The descriptor is read from pinned host memory;
// the matching, proposal count and error bits go into pinned host memory, and
// the completion flag is written as in k_serve_chains.
constexpr int kFusedLaneN = 2048;
template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 8) void k_serve_fused(const ServeReq* __restrict__ host_reqs, int count,
                                                          unsigned long long* h_props, int* h_errs, int* done,
                                                          unsigned* h_flag, unsigned seq, int max_n) {
  static_assert(sizeof(ServeReq) % 4 == 0 && sizeof(ServeReq) / 4 <= 64, "descriptor words");
  constexpr int kDescWords = (int)(sizeof(ServeReq) / 4);
  constexpr int NW = BLOCK / 64;
  extern __shared__ unsigned long long smem64[];
  unsigned long long* partner = smem64;
  int32_t* linv = reinterpret_cast<int32_t*>(partner + max_n);
  uint16_t* stage = reinterpret_cast<uint16_t*>(linv + max_n);
  __shared__ int desc[kDescWords];
  __shared__ int ctl[3];
  __shared__ int sbad;
  __shared__ unsigned long long red[NW];
  const int i = blockIdx.x;
  if (threadIdx.x < kDescWords) desc[threadIdx.x] = reinterpret_cast<const int*>(host_reqs + i)[threadIdx.x];
  if (threadIdx.x == 0) sbad = 0;
  __syncthreads();
  const ServeReq q = *reinterpret_cast<const ServeReq*>(desc);
  int bad = 0;
  unsigned long long p = 0;
  if (q.n > 0 && q.n <= max_n && q.n <= BLOCK * fused_per(BLOCK)) {
    bad = wg_solve_instance<BLOCK>(q.pm, q.pw, q.rank, q.match, q.n, partner, linv, stage, ctl, &p);
  } else {
    bad = kErrEntryRange;  // the host sends only 1 <= n <= max_n here
  }
  for (int off = 32; off > 0; off >>= 1) p += __shfl_down(p, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = p;
  if (bad) atomicOr(&sbad, bad);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < NW; ++w) t += red[w];
    h_props[i] = t;
    h_errs[i] = sbad;
  }
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0 && atomicAdd(done, 1) == count - 1) {
    *done = 0;
    __hip_atomic_store(h_flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// First-fit arena over one device allocation, 256-byte granules.
class Arena {
 public:
  void init(size_t bytes) {
    size_ = bytes;
    free_.clear();
    free_[0] = bytes;
    used_ = peak_ = 0;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = (bytes + 255) & ~size_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes, start = it->first + bytes;
        free_.erase(it);
        if (rest) free_[start] = rest;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
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
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  size_t size() const { return size_; }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0, used_ = 0, peak_ = 0;
};

}  // namespace pactae

using pactae::Clock;

struct smp_server {
  enum State { kPending, kReady, kLaunched, kDone, kFailed };
  struct Request {
    uint64_t ticket = 0;
    int n = 0;
    bool on_device = false;
    std::vector<int32_t> host_pm, host_pw;
    const int32_t* d_pm = nullptr;
    const int32_t* d_pw = nullptr;
    size_t off = 0, bytes = 0;
    bool has_block = false;
    bool to_worker = false;  // solved by a worker's single-instance pipeline
    State state = kPending;
    int status = SMP_OK;
    std::string error;
    Clock::time_point t_submit, t_launch, t_done;
    std::vector<int32_t> result;
    int64_t proposals = 0;
    int batch_size = 0, lane = -1;
  };
  struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<Request*> reqs;
    pactae::ServeReq* h_reqs = nullptr;  // pinned
    pactae::ServeReq* d_reqs = nullptr;
    int32_t* h_match = nullptr;          // pinned results, max_batch * max_n
    int* d_errs = nullptr;               // the rank pass's error bits per request
    unsigned long long* h_props = nullptr;  // pinned, written by k_serve_chains
    int* h_errs = nullptr;                  // pinned, written by k_serve_chains
    unsigned* h_flag = nullptr;             // pinned: seq of the last finished launch
    int* d_done = nullptr;                  // workgroups of the launch finished
    // Device-visible addresses of the pinned buffers.
    pactae::ServeReq* h_reqs_dev = nullptr;
    int32_t* h_match_dev = nullptr;
    unsigned long long* h_props_dev = nullptr;
    int* h_errs_dev = nullptr;
    unsigned* h_flag_dev = nullptr;
    unsigned seq = 0;  // sequence number of the launch in flight
  };

  smp_server_opts opts{};
  pactae::DeviceInfo dev;
  pactae::Arena arena;
  char* d_arena = nullptr;
  std::vector<Lane> lanes;
  hipStream_t copy_stream = nullptr;
  // Single-instance pipelines: plain mode uses engines[0] for every request; the
  // batched mode gives one to each large-request worker.
  std::vector<std::unique_ptr<smp_engine>> engines;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending, ready, ready_large;
  std::condition_variable cv_large;
  std::vector<std::thread> large_workers;
  std::map<uint64_t, std::unique_ptr<Request>> table;
  uint64_t next_ticket = 1;
  bool stop = false, stop_large = false;
  bool fatal = false;
  std::string fatal_msg;
  std::thread worker;
  smp_server_stats stats{};
  int inflight = 0;
  unsigned lane_polls = 0;  // flag checks of in-flight lanes (every 256th also queries the event)
  // Work counters read without the lock by threads that spin briefly for new
  // work before they block (a wake-up through a condition variable costs tens
  // of microseconds, a visible share of a small request's latency).
  std::atomic<uint64_t> submits{0};
  std::atomic<int> large_queued{0};
  static constexpr auto kSpin = std::chrono::microseconds(300);
  template <typename Pred>
  static void spin_for(Pred ready) {
    const auto end = Clock::now() + kSpin;
    while (!ready() && Clock::now() < end) {
    }
  }

  bool is_large(const Request& r) const { return opts.mode == SMP_SERVE_BATCHED && r.n >= opts.large_n; }
  int busy_workers = 0;  // workers solving a request (with mu held)
  // Whether a request goes to the workers (with mu held): every large request,
  // and with small_to_idle_worker a small one while a worker is idle and no
  // request waits for the workers, so that at low load it gets the whole GPU's
  // pipeline (whose host tail walks the last chains far faster than one
  // workgroup does) and under load the lanes take it.
  bool route_to_worker(const Request& r) const {
    if (opts.mode != SMP_SERVE_BATCHED || large_workers.empty()) return false;
    if (is_large(r)) return true;
    return opts.small_to_idle_worker > 0 &&
           (int)large_workers.size() - busy_workers > (int)ready_large.size();
  }

  // Arena bytes of the int16 rank matrix (lane requests only: the
  // single-instance pipelines of the large-request workers and of plain mode
  // have their own) and of the list copies (host-submitted requests).
  size_t rank_bytes(const Request& r) const {
    return opts.mode != SMP_SERVE_BATCHED || r.to_worker
               ? 0
               : (((size_t)r.n * r.n * 2 + 255) & ~size_t(255));
  }
  size_t footprint(const Request& r) const {
    size_t b = rank_bytes(r);
    if (!r.on_device) b += 2 * (((size_t)r.n * r.n * 4 + 255) & ~size_t(255));
    return std::max<size_t>(b, 256);
  }

  int setup() {
    if (int st = pactae::query_device(opts.device, &dev)) return st;
    if (hipMalloc(&d_arena, (size_t)opts.hbm_budget_bytes) != hipSuccess) {
      (void)hipGetLastError();
      pactae::set_error("server: arena allocation of " + std::to_string(opts.hbm_budget_bytes) +
                        " bytes failed");
      return SMP_ERR_NOMEM;
    }
    arena.init((size_t)opts.hbm_budget_bytes);
    SMP_HIP_TRY(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
    const int nengines = opts.mode == SMP_SERVE_PLAIN ? 1 : opts.large_n <= opts.max_n ? opts.large_workers : 0;
    for (int i = 0; i < nengines; ++i) {
      smp_opts o;
      smp_default_opts(&o);
      o.device = opts.device;
      auto e = std::make_unique<smp_engine>();
      e->n = opts.max_n;
      e->opts = o;
      if (int st = e->setup()) return st;
      engines.push_back(std::move(e));
    }
    if (opts.mode == SMP_SERVE_BATCHED) {
      lanes.resize(opts.lanes);
      const size_t mb = (size_t)opts.max_batch;
      for (auto& L : lanes) {
        SMP_HIP_TRY(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
        SMP_HIP_TRY(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
        SMP_HIP_TRY(hipHostMalloc(&L.h_reqs, mb * sizeof(pactae::ServeReq), hipHostMallocDefault));
        SMP_HIP_TRY(hipHostMalloc(&L.h_match, mb * opts.max_n * sizeof(int32_t), hipHostMallocDefault));
        SMP_HIP_TRY(hipHostMalloc(&L.h_props, mb * sizeof(unsigned long long), hipHostMallocDefault));
        SMP_HIP_TRY(hipHostMalloc(&L.h_errs, mb * sizeof(int), hipHostMallocDefault));
        SMP_HIP_TRY(hipHostMalloc(&L.h_flag, sizeof(unsigned), hipHostMallocDefault));
        *L.h_flag = 0;
        SMP_HIP_TRY(hipMalloc(&L.d_reqs, mb * sizeof(pactae::ServeReq)));
        SMP_HIP_TRY(hipMalloc(&L.d_errs, mb * sizeof(int)));
        SMP_HIP_TRY(hipMalloc(&L.d_done, sizeof(int)));
        SMP_HIP_TRY(hipMemset(L.d_done, 0, sizeof(int)));
        SMP_HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&L.h_reqs_dev), L.h_reqs, 0));
        SMP_HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&L.h_match_dev), L.h_match, 0));
        SMP_HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&L.h_props_dev), L.h_props, 0));
        SMP_HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&L.h_errs_dev), L.h_errs, 0));
        SMP_HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void**>(&L.h_flag_dev), L.h_flag, 0));
      }
      SMP_HIP_TRY(hipDeviceSynchronize());  // the d_done clears
      const int lds = (int)std::min<size_t>((size_t)opts.max_n * 8, 160 * 1024);
      SMP_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&pactae::k_serve_chains<256>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, lds));
      SMP_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&pactae::k_serve_chains<512>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, lds));
      SMP_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&pactae::k_serve_chains<1024>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, lds));
      for (size_t i = 0; i < engines.size(); ++i)
        large_workers.emplace_back([this, i] { large_loop(*engines[i], (int)i); });
    }
    worker = std::thread([this] { loop(); });
    return SMP_OK;
  }

  // Solves one request with the single-instance pipeline; `lane` is recorded.
  void solve_plain(std::unique_lock<std::mutex>& lk, Request* r, int lane, smp_engine& plain) {
    r->state = kLaunched;
    r->t_launch = Clock::now();
    r->batch_size = 1;
    r->lane = lane;
    ++stats.launches;
    lk.unlock();
    int st = plain.run_view(r->d_pm, r->d_pw, r->n, nullptr);
    std::vector<int32_t> res;
    int64_t props = 0;
    std::string err;
    if (st == SMP_OK) {
      res.assign(plain.h_match, plain.h_match + r->n);
      props = plain.times.proposals;
    } else {
      err = pactae::last_error();
    }
    lk.lock();
    r->result = std::move(res);
    r->proposals = props;
    finish(r, st, err);
  }

  // Batched mode: the large-request worker `idx`.  With large_spill_us > 0,
  // worker 0 takes any waiting request and worker i > 0 only when more than i
  // wait or the oldest has waited large_spill_us (at low load a request then runs
  // alone on the GPU; under load all workers run).  By default any idle worker
  // takes it at once, which measured best for p50 and p99 at 1,500/s.
  bool may_take(int idx, Clock::time_point now, Clock::duration* wait_left) const {  // with mu held
    if (ready_large.empty()) return false;
    if (idx == 0 || (int)ready_large.size() > idx || opts.large_spill_us <= 0) return true;
    const auto spill = std::chrono::microseconds(opts.large_spill_us);
    const auto waited = now - ready_large.front()->t_submit;
    if (waited >= spill) return true;
    *wait_left = spill - waited;
    return false;
  }

  void large_loop(smp_engine& eng, int idx) {
    (void)hipSetDevice(opts.device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      if (!stop_large && ready_large.empty()) {
        lk.unlock();
        spin_for([this] { return large_queued.load(std::memory_order_acquire) > 0; });
        lk.lock();
      }
      for (;;) {
        if (stop_large && ready_large.empty()) break;
        Clock::duration left{};
        if (may_take(idx, Clock::now(), &left)) break;
        if (ready_large.empty())
          cv_large.wait(lk);
        else
          cv_large.wait_for(lk, left);
      }
      if (ready_large.empty()) break;
      Request* r = ready_large.front();
      ready_large.pop_front();
      large_queued.fetch_sub(1, std::memory_order_relaxed);
      ++busy_workers;
      solve_plain(lk, r, -1, eng);
      --busy_workers;
      ++stats.large_solved;
      cv_work.notify_one();  // its arena block is free for admission
    }
  }

  ~smp_server() {
    {
      std::lock_guard<std::mutex> g(mu);
      stop = true;
    }
    cv_work.notify_all();
    if (worker.joinable()) worker.join();
    {
      std::lock_guard<std::mutex> g(mu);
      stop_large = true;
    }
    cv_large.notify_all();
    for (auto& t : large_workers)
      if (t.joinable()) t.join();
    for (auto& L : lanes) {
      if (L.stream) (void)hipStreamSynchronize(L.stream);
      (void)hipHostFree(L.h_reqs);
      (void)hipHostFree(L.h_match);
      (void)hipHostFree(L.h_props);
      (void)hipHostFree(L.h_errs);
      (void)hipHostFree(L.h_flag);
      (void)hipFree(L.d_reqs);
      (void)hipFree(L.d_errs);
      (void)hipFree(L.d_done);
      if (L.done) (void)hipEventDestroy(L.done);
      if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    engines.clear();
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    (void)hipFree(d_arena);
  }

  void fail_all(const std::string& msg) {  // with mu held
    fatal = true;
    fatal_msg = msg;
    for (auto& kv : table) {
      Request* r = kv.second.get();
      if (r->state != kDone && r->state != kFailed) {
        r->state = kFailed;
        r->status = SMP_ERR_HIP;
        r->error = msg;
      }
    }
    pending.clear();
    ready.clear();
    ready_large.clear();
    large_queued.store(0);
    cv_done.notify_all();
  }

  void finish(Request* r, int status, const std::string& err) {  // with mu held
    r->t_done = Clock::now();
    r->status = status;
    r->error = err;
    r->state = status == SMP_OK ? kDone : kFailed;
    if (r->has_block) arena.release(r->off, r->bytes), r->has_block = false;
    --inflight;
    ++stats.completed;
    cv_done.notify_all();
  }

  // Admit pending requests in arrival order while the arena has room.  Copies of
  // host lists are made here (outside the lock).  Returns a HIP status.
  hipError_t admit(std::unique_lock<std::mutex>& lk, bool* progressed) {
    while (!pending.empty()) {
      Request* r = pending.front();
      r->to_worker = route_to_worker(*r);
      r->bytes = footprint(*r);
      if (!arena.alloc(r->bytes, &r->off)) break;
      r->has_block = true;
      pending.pop_front();
      ++inflight;
      stats.peak_inflight = std::max<int32_t>(stats.peak_inflight, inflight);
      *progressed = true;
      if (!r->on_device) {
        const size_t nn = (size_t)r->n * r->n;
        const size_t rank_b = rank_bytes(*r);
        const size_t list_b = (nn * 4 + 255) & ~size_t(255);
        int32_t* dm = reinterpret_cast<int32_t*>(d_arena + r->off + rank_b);
        int32_t* dw = reinterpret_cast<int32_t*>(d_arena + r->off + rank_b + list_b);
        lk.unlock();
        hipError_t e = hipMemcpyAsync(dm, r->host_pm.data(), nn * 4, hipMemcpyHostToDevice, copy_stream);
        if (e == hipSuccess)
          e = hipMemcpyAsync(dw, r->host_pw.data(), nn * 4, hipMemcpyHostToDevice, copy_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(copy_stream);
        lk.lock();
        if (e != hipSuccess) return e;
        r->d_pm = dm;
        r->d_pw = dw;
        std::vector<int32_t>().swap(r->host_pm);
        std::vector<int32_t>().swap(r->host_pw);
      }
      r->state = kReady;
      if (r->to_worker) {
        push_large(r);
      } else {
        ready.push_back(r);
      }
    }
    return hipSuccess;
  }

  void push_large(Request* r) {  // with mu held
    ready_large.push_back(r);
    large_queued.fetch_add(1, std::memory_order_release);
    cv_large.notify_all();  // the workers decide which of them takes it
  }

  // A request for the workers with its lists in HBM needs no arena room of its
  // own (the worker's pipeline has its buffers), so when nothing waits ahead of
  // it in arrival order it goes straight to them from the submitting thread.
  bool admit_large_now(Request* r) {  // with mu held
    if (!r->on_device || !pending.empty() || !route_to_worker(*r)) return false;
    r->to_worker = true;
    r->bytes = footprint(*r);
    if (!arena.alloc(r->bytes, &r->off)) return false;
    r->has_block = true;
    ++inflight;
    stats.peak_inflight = std::max<int32_t>(stats.peak_inflight, inflight);
    r->state = kReady;
    push_large(r);
    return true;
  }

  hipError_t launch(Lane& L, int lane_id) {  // with mu held
    const int count = std::min<int>((int)ready.size(), opts.max_batch);
    L.reqs.assign(ready.begin(), ready.begin() + count);
    ready.erase(ready.begin(), ready.begin() + count);
    int rows = 0, bmax = 0;
    const auto now = Clock::now();
    for (int i = 0; i < count; ++i) {
      Request* r = L.reqs[i];
      r->state = kLaunched;
      r->t_launch = now;
      r->batch_size = count;
      r->lane = lane_id;
      int32_t* dm = L.h_match_dev + (size_t)i * opts.max_n;
      L.h_reqs[i] = pactae::ServeReq{r->d_pm, r->d_pw, reinterpret_cast<int16_t*>(d_arena + r->off), dm,
                                     r->n, rows};
      rows += r->n;
      bmax = std::max(bmax, r->n);
    }
    ++L.seq;
    hipError_t e = hipSuccess;
    if (bmax <= pactae::kFusedLaneN && opts.fused_lanes) {
      // One fused kernel, one workgroup per request.
      const size_t lds = pactae::fused_lds(1024, bmax);
      hipLaunchKernelGGL(pactae::k_serve_fused<1024>, dim3(count), dim3(1024), lds, L.stream, L.h_reqs_dev, count,
                         L.h_props_dev, L.h_errs_dev, L.d_done, L.h_flag_dev, L.seq, bmax);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      if ((e = hipEventRecord(L.done, L.stream)) != hipSuccess) return e;
      L.busy = true;
      ++stats.launches;
      return hipSuccess;
    }
    hipLaunchKernelGGL(pactae::k_serve_prep, dim3(1), dim3(256), 0, L.stream, L.h_reqs_dev, L.d_reqs, count,
                       L.d_errs);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = pactae::launch_rank_rows16(L.stream, L.d_reqs, count, static_cast<const int*>(nullptr), rows,
                                        bmax, L.d_errs, 1, dev.cus, dev.max_lds)) != hipSuccess)
      return e;
    const size_t lds = (size_t)bmax * 8;
    if (opts.block_threads == 256)
      hipLaunchKernelGGL(pactae::k_serve_chains<256>, dim3(count), dim3(256), lds, L.stream, L.d_reqs,
                         count, L.d_errs, L.h_props_dev, L.h_errs_dev, L.d_done, L.h_flag_dev, L.seq);
    else if (opts.block_threads == 512)
      hipLaunchKernelGGL(pactae::k_serve_chains<512>, dim3(count), dim3(512), lds, L.stream, L.d_reqs,
                         count, L.d_errs, L.h_props_dev, L.h_errs_dev, L.d_done, L.h_flag_dev, L.seq);
    else
      hipLaunchKernelGGL(pactae::k_serve_chains<1024>, dim3(count), dim3(1024), lds, L.stream, L.d_reqs,
                         count, L.d_errs, L.h_props_dev, L.h_errs_dev, L.d_done, L.h_flag_dev, L.seq);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // The event only backs up the flag: it tells a failed launch or a device
    // error (or a stream that drained without the flag) from one still running.
    if ((e = hipEventRecord(L.done, L.stream)) != hipSuccess) return e;
    L.busy = true;
    ++stats.launches;
    return hipSuccess;
  }

  bool lane_flag_seen(const Lane& L) const { return __atomic_load_n(L.h_flag, __ATOMIC_ACQUIRE) == L.seq; }

  void complete(Lane& L) {  // with mu held; the lane's flag holds its seq
    for (size_t i = 0; i < L.reqs.size(); ++i) {
      Request* r = L.reqs[i];
      const int32_t* hm = L.h_match + i * opts.max_n;
      r->result.assign(hm, hm + r->n);
      r->proposals = (int64_t)L.h_props[i];
      if (L.h_errs[i])
        finish(r, SMP_ERR_INVALID, "invalid preference lists (device flags " + std::to_string(L.h_errs[i]) + ")");
      else
        finish(r, SMP_OK, "");
    }
    L.reqs.clear();
    L.busy = false;
  }

  void loop() {
    (void)hipSetDevice(opts.device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      bool progressed = false;
      if (!fatal && opts.mode == SMP_SERVE_PLAIN) {
        if (hipError_t e = admit(lk, &progressed); e != hipSuccess) {
          fail_all(std::string("server: ") + hipGetErrorString(e));
          continue;
        }
        if (!ready.empty()) {
          Request* r = ready.front();
          ready.pop_front();
          solve_plain(lk, r, 0, *engines[0]);
          progressed = true;
        }
      } else if (!fatal) {
        int busy = 0;
        for (auto& L : lanes) {
          if (L.busy) {
            if (lane_flag_seen(L)) {
              complete(L);
              progressed = true;
              continue;
            }
            // Not finished by the flag: ask the runtime now and then, so that a
            // failed launch or a device error ends the wait.
            if ((++lane_polls & 255) != 0) continue;
            hipError_t q = hipEventQuery(L.done);
            if (q == hipErrorNotReady) continue;
            if (q == hipSuccess && lane_flag_seen(L)) {
              complete(L);
              progressed = true;
              continue;
            }
            fail_all(std::string("server: lane launch ended without its results (") +
                     (q == hipSuccess ? "no completion flag" : hipGetErrorString(q)) + ")");
            break;
          }
        }
        if (hipError_t e = admit(lk, &progressed); e != hipSuccess) {
          fail_all(std::string("server: ") + hipGetErrorString(e));
          continue;
        }
        for (size_t li = 0; li < lanes.size() && !ready.empty(); ++li) {
          if (lanes[li].busy) continue;
          if (hipError_t e = launch(lanes[li], (int)li); e != hipSuccess) {
            fail_all(std::string("server launch: ") + hipGetErrorString(e));
            break;
          }
          progressed = true;
        }
        for (auto& L : lanes) busy += L.busy;
        stats.peak_lanes_busy = std::max<int32_t>(stats.peak_lanes_busy, busy);
        if (busy && !progressed) {
          // A lane is in flight: poll again soon (a timed sleep would last tens
          // of microseconds past its request, the timer slack).
          lk.unlock();
          std::this_thread::yield();
          lk.lock();
          continue;
        }
      }
      stats.arena_bytes = (int64_t)arena.used();
      stats.peak_arena_bytes = (int64_t)arena.peak();
      const bool idle = pending.empty() && ready.empty() && inflight == 0;
      if (stop && (idle || fatal)) break;
      if (!progressed && pending.empty() && !stop) {
        const uint64_t seen = submits.load(std::memory_order_acquire);
        lk.unlock();
        spin_for([&] { return submits.load(std::memory_order_acquire) != seen; });
        lk.lock();
        if (submits.load(std::memory_order_acquire) != seen) continue;
      }
      if (!progressed) cv_work.wait_for(lk, std::chrono::milliseconds(2));
    }
  }
};

extern "C" {

void smp_server_default_opts(smp_server_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->device = 0;
  o->mode = SMP_SERVE_BATCHED;
  o->lanes = 3;
  o->max_batch = 64;
  o->max_n = 8192;
  o->block_threads = 0;
  o->hbm_budget_bytes = 1ll << 30;
  o->large_n = 512;
  o->large_workers = 2;
  o->large_spill_us = 0;
  o->small_to_idle_worker = -1;
  o->fused_lanes = -1;
}

int smp_server_create(const smp_server_opts* opts, smp_server** out) {
  if (!out) return SMP_ERR_INVALID;
  *out = nullptr;
  auto s = std::make_unique<smp_server>();
  if (opts)
    s->opts = *opts;
  else
    smp_server_default_opts(&s->opts);
  auto& o = s->opts;
  if (o.lanes <= 0) o.lanes = 3;
  if (o.max_batch <= 0) o.max_batch = 64;
  if (o.max_n <= 0) o.max_n = 8192;
  if (o.hbm_budget_bytes <= 0) o.hbm_budget_bytes = 1ll << 30;
  if (o.block_threads != 256 && o.block_threads != 512 && o.block_threads != 1024) o.block_threads = 1024;
  if (o.large_n <= 0) o.large_n = 512;
  if (o.small_to_idle_worker < 0) o.small_to_idle_worker = 1;
  if (o.fused_lanes < 0) o.fused_lanes = 1;
  if (o.large_workers <= 0) o.large_workers = 2;
  if (o.large_workers > 8) {
    pactae::set_error("server: large_workers must be <= 8");
    return SMP_ERR_INVALID;
  }
  if (o.max_n > 20000 || o.lanes > 16 || o.max_batch > 4096 ||
      (o.mode != SMP_SERVE_BATCHED && o.mode != SMP_SERVE_PLAIN)) {
    pactae::set_error("server: max_n must be <= 20000 (LDS partner words), lanes <= 16, max_batch <= 4096");
    return SMP_ERR_INVALID;
  }
  if (int st = s->setup()) return st;
  *out = s.release();
  return SMP_OK;
}

int smp_server_submit(smp_server* s, const int32_t* pm, const int32_t* pw, int32_t n, int32_t on_device,
                      uint64_t* ticket) {
  if (!s || !pm || !pw || !ticket || n <= 0 || n > s->opts.max_n) {
    pactae::set_error("server submit: bad arguments (n must be in [1, max_n])");
    return SMP_ERR_INVALID;
  }
  auto r = std::make_unique<smp_server::Request>();
  r->n = n;
  r->on_device = on_device != 0;
  r->t_submit = Clock::now();
  if (r->on_device) {
    r->d_pm = pm;
    r->d_pw = pw;
  } else {
    r->host_pm.assign(pm, pm + (size_t)n * n);
    r->host_pw.assign(pw, pw + (size_t)n * n);
  }
  // (Admission decides where the request goes; a large one that the workers
  // take needs no rank block.)
  r->to_worker = s->is_large(*r) && !s->large_workers.empty();
  if (s->footprint(*r) > s->arena.size()) {
    pactae::set_error("server submit: request footprint exceeds the HBM budget");
    return SMP_ERR_NOMEM;
  }
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->fatal) {
      pactae::set_error(s->fatal_msg);
      return SMP_ERR_HIP;
    }
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    smp_server::Request* rp = r.get();
    s->table.emplace(r->ticket, std::move(r));
    ++s->stats.submitted;
    if (s->admit_large_now(rp)) return SMP_OK;
    s->pending.push_back(rp);
    s->submits.fetch_add(1, std::memory_order_release);
  }
  s->cv_work.notify_one();
  return SMP_OK;
}

static int collect(smp_server* s, std::map<uint64_t, std::unique_ptr<smp_server::Request>>::iterator it,
                   int32_t* match, smp_request_info* info) {
  smp_server::Request* r = it->second.get();
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->queue_ms = std::chrono::duration<double, std::milli>(r->t_launch - r->t_submit).count();
    info->service_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_launch).count();
    info->latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    info->proposals = r->proposals;
    info->n = r->n;
    info->batch_size = r->batch_size;
    info->lane = r->lane;
  }
  const int st = r->status;
  if (st == SMP_OK && match) std::memcpy(match, r->result.data(), sizeof(int32_t) * r->n);
  if (st != SMP_OK) pactae::set_error(r->error);
  s->table.erase(it);
  return st;
}

int smp_server_poll(smp_server* s, uint64_t ticket, int32_t* match, smp_request_info* info) {
  if (!s) return SMP_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->table.find(ticket);
  if (it == s->table.end()) {
    pactae::set_error("server poll: unknown ticket");
    return SMP_ERR_INVALID;
  }
  const auto st = it->second->state;
  if (st != smp_server::kDone && st != smp_server::kFailed) return SMP_PENDING;
  return collect(s, it, match, info);
}

int smp_server_wait(smp_server* s, uint64_t ticket, int32_t* match, smp_request_info* info,
                    int32_t timeout_ms) {
  if (!s) return SMP_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->table.find(ticket);
  if (it == s->table.end()) {
    pactae::set_error("server wait: unknown ticket");
    return SMP_ERR_INVALID;
  }
  smp_server::Request* r = it->second.get();
  auto done = [&] { return r->state == smp_server::kDone || r->state == smp_server::kFailed; };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, done);
  } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) {
    pactae::set_error("server wait: timed out");
    return SMP_ERR_TIMEOUT;
  }
  return collect(s, it, match, info);
}

int smp_server_stats_get(smp_server* s, smp_server_stats* st) {
  if (!s || !st) return SMP_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  *st = s->stats;
  st->arena_bytes = (int64_t)s->arena.used();
  st->peak_arena_bytes = (int64_t)s->arena.peak();
  return SMP_OK;
}

void smp_server_destroy(smp_server* s) { delete s; }

}  // extern "C"
