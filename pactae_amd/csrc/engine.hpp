// engine.hpp — the single-instance engine (AbsSmpEngine-shaped) shared by the C ABI
// and by the looped baseline of the batched path.
//
// Host structure follows the reference engine framework:
//   AbsSmpEngine::FindStableMatching   include/smp/smp_engine_abs.h:10-36
//     precheck (IsPerfect) -> InitProc -> CoreProc, each timed with a steady clock
//   SmpEngineBamboo                    src/smp/smp_engine_bamboo.cu:1-307
// with these MI355X-specific changes:
//   * every phase runs on one HIP stream of the engine; exec returns only after all
//     device work and the host tail have finished (the reference detaches its
//     kernel thread, smp_engine_bamboo.cu:108-112);
//   * the hand-off to the host is decided by a device counter of free men that the
//     proposal kernel maintains, not by 10 ms polls of a racing snapshot
//     (smp_engine_bamboo.cu:232-287); the kernel parks its chains and exits, so the
//     host works on a consistent state;
//   * the host tail reads a compact men-major PR prefix copied while the kernel
//     runs (the reference copies all 8n^2 bytes of its PR matrix, untimed,
//     smp_engine_abs.h:23); rows beyond the prefix are gathered on demand.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.hpp"

namespace pactae {

using Clock = std::chrono::steady_clock;
inline double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

struct DeviceInfo {
  int cus = 256;
  int max_lds = 65536;  // bytes a workgroup may allocate
};

inline int query_device(int dev, DeviceInfo* info) {
  int count = 0;
  SMP_HIP_TRY(hipGetDeviceCount(&count));
  if (dev < 0 || dev >= count) {
    set_error("device ordinal " + std::to_string(dev) + " not present (" + std::to_string(count) +
              " devices)");
    return SMP_ERR_HIP;
  }
  SMP_HIP_TRY(hipSetDevice(dev));
  hipDeviceProp_t p;
  SMP_HIP_TRY(hipGetDeviceProperties(&p, dev));
  info->cus = p.multiProcessorCount;
  // gfx950 offers 160 KiB of LDS per CU; sharedMemPerBlock reports what a single
  // workgroup may take.
  info->max_lds = (int)std::max<size_t>(p.sharedMemPerBlock, 65536);
  return SMP_OK;
}

// ---------------------------------------------------------------------------
// AbsSmpEngine template method (smp_engine_abs.h:10-36), with status codes.
class AbsSmpEngine {
 public:
  virtual ~AbsSmpEngine() = default;

  int FindStableMatching() {
    auto t0 = Clock::now();
    bool perfect = false;
    int st = IsPerfect(&perfect);
    precheck_time_ = ms_since(t0);
    if (st) return st;
    init_time_ = core_time_ = 0;
    if (!perfect) {
      t0 = Clock::now();
      st = InitProc();
      init_time_ = ms_since(t0);
      if (st) return st;
      t0 = Clock::now();
      st = CoreProc();
      core_time_ = ms_since(t0);
      if (st) return st;
    }
    t0 = Clock::now();
    st = PostProc(perfect);
    post_time_ = ms_since(t0);
    return st;
  }

 protected:
  virtual int IsPerfect(bool* perfect) = 0;
  virtual int InitProc() = 0;
  virtual int CoreProc() = 0;
  virtual int PostProc(bool perfect) = 0;

  double precheck_time_ = 0, init_time_ = 0, core_time_ = 0, post_time_ = 0;
};

}  // namespace pactae

using pactae::Clock;
using pactae::ms_since;

// ---------------------------------------------------------------------------
struct smp_engine : public pactae::AbsSmpEngine {
  using Scalars = pactae::Scalars;

  int n = 0;        // size of the instance being solved
  int n_alloc = 0;  // size the buffers were allocated for (>= n)
  smp_opts opts{};
  pactae::DeviceInfo dev;
  int K = 0;  // PR prefix columns
  int T = 0;  // hand-off threshold (0 = GPU finishes)
  bool owns_lists = false;
  bool have_result = false;

  const int32_t* d_pref_m = nullptr;
  const int32_t* d_pref_w = nullptr;
  int32_t* d_rank_w = nullptr;
  uint32_t* d_pr = nullptr;
  unsigned long long* d_partner = nullptr;
  int32_t* d_match = nullptr;
  int* d_seen = nullptr;
  Scalars* d_sc = nullptr;
  int2* d_parked = nullptr;
  uint32_t* d_rowbuf = nullptr;

  uint32_t* h_pr = nullptr;  // pinned copy of the PR prefix for the host tail
  unsigned long long* h_partner = nullptr;
  Scalars* h_sc = nullptr;
  int2* h_parked = nullptr;
  uint32_t* h_rowbuf = nullptr;
  std::vector<int32_t> h_match;

  hipStream_t stream = nullptr, copy_stream = nullptr;
  hipEvent_t ev[8] = {};

  smp_times times{};
  static constexpr int kRowFetch = 2048;

  ~smp_engine() override { release(); }

  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    if (owns_lists) {
      (void)hipFree(const_cast<int32_t*>(d_pref_m));
      (void)hipFree(const_cast<int32_t*>(d_pref_w));
    }
    d_pref_m = d_pref_w = nullptr;
    (void)hipFree(d_rank_w);
    (void)hipFree(d_pr);
    (void)hipFree(d_partner);
    (void)hipFree(d_match);
    (void)hipFree(d_seen);
    (void)hipFree(d_sc);
    (void)hipFree(d_parked);
    (void)hipFree(d_rowbuf);
    (void)hipHostFree(h_pr);
    (void)hipHostFree(h_partner);
    (void)hipHostFree(h_sc);
    (void)hipHostFree(h_parked);
    (void)hipHostFree(h_rowbuf);
    d_rank_w = nullptr;
    d_pr = nullptr;
    d_partner = nullptr;
    d_match = nullptr;
    d_seen = nullptr;
    d_sc = nullptr;
    d_parked = nullptr;
    d_rowbuf = nullptr;
    h_pr = nullptr;
    h_partner = nullptr;
    h_sc = nullptr;
    h_parked = nullptr;
    h_rowbuf = nullptr;
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e), e = nullptr;
    if (stream) (void)hipStreamDestroy(stream), stream = nullptr;
    if (copy_stream) (void)hipStreamDestroy(copy_stream), copy_stream = nullptr;
  }

  int setup() {
    if (int st = pactae::query_device(opts.device, &dev)) return st;
    const size_t nn = (size_t)n * n;
    n_alloc = n;
    configure(n);
    SMP_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    SMP_HIP_TRY(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
    for (auto& e : ev) SMP_HIP_TRY(hipEventCreate(&e));
    if (hipMalloc(&d_rank_w, nn * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&d_pr, (size_t)n * K * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&d_partner, (size_t)n * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&d_match, (size_t)n * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&d_seen, (size_t)n * sizeof(int)) != hipSuccess ||
        hipMalloc(&d_sc, sizeof(Scalars)) != hipSuccess ||
        hipMalloc(&d_parked, (size_t)n * sizeof(int2)) != hipSuccess ||
        hipMalloc(&d_rowbuf, kRowFetch * sizeof(uint32_t)) != hipSuccess) {
      (void)hipGetLastError();
      pactae::set_error("device allocation failed for n=" + std::to_string(n));
      return SMP_ERR_NOMEM;
    }
    if (hipHostMalloc(&h_pr, (size_t)n * K * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&h_partner, (size_t)n * sizeof(unsigned long long), hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc(&h_sc, sizeof(Scalars), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&h_parked, (size_t)n * sizeof(int2), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&h_rowbuf, kRowFetch * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      pactae::set_error("pinned host allocation failed for n=" + std::to_string(n));
      return SMP_ERR_NOMEM;
    }
    const int chunk_cols = std::max(1024, (dev.max_lds / 4) & ~3);
    if (hipError_t e = pactae::set_invert_lds_limit(std::min(n, chunk_cols) * 4); e != hipSuccess) {
      pactae::set_error(std::string("hipFuncSetAttribute(LDS) failed: ") + hipGetErrorString(e));
      return SMP_ERR_HIP;
    }
    h_match.assign(n, -1);
    return SMP_OK;
  }

  // Prefix width and hand-off threshold for an instance of size ni <= n_alloc.
  int prefix_for(int ni) const {
    return opts.prefix_cols <= 0 ? std::min(ni, 128) : std::min(ni, opts.prefix_cols);
  }
  void configure(int ni) {
    n = ni;
    K = prefix_for(ni);
    if (opts.tail_mode == SMP_TAIL_GPU)
      T = 0;
    else
      T = opts.tail_threshold > 0 ? std::min(opts.tail_threshold, ni) : std::min(ni, 32);
  }

  // Solve another instance (lists already in HBM, size <= n_alloc) with this
  // engine's buffers: the plain single-instance pipeline, used by the looped
  // baseline of the batched path.
  int run_view(const int32_t* pm, const int32_t* pw, int ni, int32_t* d_out) {
    if (ni <= 0 || ni > n_alloc) return SMP_ERR_INVALID;
    d_pref_m = pm;
    d_pref_w = pw;
    configure(ni);
    if (int st = exec()) return st;
    if (d_out) {
      SMP_HIP_TRY(hipMemcpyAsync(d_out, d_match, (size_t)ni * sizeof(int32_t),
                                 hipMemcpyDeviceToDevice, stream));
      SMP_HIP_TRY(hipStreamSynchronize(stream));
    }
    return SMP_OK;
  }

  int read_scalars() {
    SMP_HIP_TRY(hipMemcpyAsync(h_sc, d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, stream));
    SMP_HIP_TRY(hipStreamSynchronize(stream));
    return SMP_OK;
  }

  int check_device_error(const char* phase) {
    if (h_sc->error == 0) return SMP_OK;
    std::string what;
    if (h_sc->error & pactae::kErrEntryRange) what += " entry outside [0,n);";
    if (h_sc->error & pactae::kErrNotPerm) what += " a woman's row is not a permutation;";
    if (h_sc->error & pactae::kErrListEnd) what += " a man exhausted his list;";
    if (h_sc->error & pactae::kErrUnmatched) what += " a woman was left unmatched;";
    pactae::set_error(std::string("invalid preference lists (") + phase + "):" + what);
    return SMP_ERR_INVALID;
  }

  // -- precheck (smp_engine_bamboo.h:25-33) ---------------------------------
  int IsPerfect(bool* perfect) override {
    times = smp_times{};
    have_result = false;
    SMP_HIP_TRY(pactae::launch_reset_scalars(d_sc, n, stream));
    if (!opts.precheck) {
      *perfect = false;
      return SMP_OK;
    }
    SMP_HIP_TRY(pactae::launch_fill_i32(d_seen, n, 0, stream));
    SMP_HIP_TRY(pactae::launch_first_choices(d_pref_m, n, d_seen, d_sc, stream));
    if (int st = read_scalars()) return st;
    if (int st = check_device_error("precheck")) return st;
    *perfect = h_sc->dup_first == 0;
    times.is_perfect = *perfect ? 1 : 0;
    return SMP_OK;
  }

  // -- init: rank matrix + PR prefix gather (smp_engine_bamboo.cu:33-106) -----
  int InitProc() override {
    const int chunk = std::min(n, std::max(1024, (dev.max_lds / 4) & ~3));
    SMP_HIP_TRY(pactae::launch_fill_u64(d_partner, n, pactae::kEmpty, stream));
    SMP_HIP_TRY(hipEventRecord(ev[0], stream));
    SMP_HIP_TRY(pactae::launch_invert_rows(d_pref_w, d_rank_w, n, chunk, opts.validate, 1024, d_sc,
                                           stream));
    SMP_HIP_TRY(hipEventRecord(ev[1], stream));
    SMP_HIP_TRY(pactae::launch_gather_pr(d_pref_m, d_rank_w, d_pr, n, K, d_sc, stream));
    SMP_HIP_TRY(hipEventRecord(ev[2], stream));
    if (T > 0) {
      // PR prefix to the host, overlapped with the proposal kernel.
      SMP_HIP_TRY(hipStreamWaitEvent(copy_stream, ev[2], 0));
      SMP_HIP_TRY(hipMemcpyAsync(h_pr, d_pr, (size_t)n * K * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 copy_stream));
    }
    if (int st = read_scalars()) return st;
    return check_device_error("init");
  }

  // -- core: persistent proposal kernel + host tail --------------------------
  int CoreProc() override {
    pactae::ProposeArgs a;
    a.pr = d_pr;
    a.pref_m = d_pref_m;
    a.rank_w = d_rank_w;
    a.partner = d_partner;
    a.parked = d_parked;
    a.sc = d_sc;
    a.n = n;
    a.K = K;
    a.T = T;
    const int wpc = opts.waves_per_cu > 0 ? opts.waves_per_cu : 16;
    long long waves = std::min<long long>((long long)dev.cus * wpc, (long long)n);
    const int grid = (int)std::max<long long>(1, (waves + 3) / 4);
    SMP_HIP_TRY(hipEventRecord(ev[3], stream));
    SMP_HIP_TRY(pactae::launch_propose(a, grid, stream));
    SMP_HIP_TRY(hipEventRecord(ev[4], stream));
    if (int st = read_scalars()) return st;
    if (int st = check_device_error("proposal kernel")) return st;
    times.atomics = (int64_t)h_sc->atomics;
    // Men [0, waves) were taken by the waves' first iterations, the rest from the cursor.
    const int started = (int)std::min<long long>(n, (long long)grid * 4 + h_sc->cursor);
    const int handoff = h_sc->parked_count + (n - started);
    times.handoff_free = handoff;
    if (handoff > 0) {
      if (int st = host_tail(started)) return st;
    }
    return SMP_OK;
  }

  // Host tail: McVitie-Wilson chains (smp_engine_mw_1_f9.cu:15-79, and the
  // reference's host LAProcedure smp_engine_bamboo.cu:197-230) over the packed
  // partner words, starting from the state the kernel parked.
  int host_tail(int started) {
    auto t0 = Clock::now();
    SMP_HIP_TRY(hipMemcpyAsync(h_partner, d_partner, (size_t)n * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, stream));
    const int parked = h_sc->parked_count;
    if (parked > 0)
      SMP_HIP_TRY(hipMemcpyAsync(h_parked, d_parked, (size_t)parked * sizeof(int2),
                                 hipMemcpyDeviceToHost, stream));
    SMP_HIP_TRY(hipStreamSynchronize(stream));
    SMP_HIP_TRY(hipStreamSynchronize(copy_stream));
    times.handoff_ms = ms_since(t0);

    t0 = Clock::now();
    std::vector<int2> todo(h_parked, h_parked + parked);
    for (int m = started; m < n; ++m) todo.push_back(make_int2(m, 0));
    struct Ext {
      int k0 = 0;
      std::vector<uint32_t> e;
    };
    std::unordered_map<int, Ext> ext;
    long long props = 0;
    int fetches = 0;
    for (const int2& start : todo) {
      int m = start.x, k = start.y;
      for (;;) {
        if (k >= n) {
          pactae::set_error("host tail: a man exhausted his list (rows are not permutations)");
          return SMP_ERR_INVALID;
        }
        uint32_t e;
        if (k < K) {
          e = h_pr[(size_t)m * K + k];
        } else {
          Ext& x = ext[m];
          if (x.e.empty() || k < x.k0 || k >= x.k0 + (int)x.e.size()) {
            const int cnt = std::min(kRowFetch, n - k);
            SMP_HIP_TRY(pactae::launch_gather_row(d_pref_m, d_rank_w, n, m, k, cnt, d_rowbuf, stream));
            SMP_HIP_TRY(hipMemcpyAsync(h_rowbuf, d_rowbuf, cnt * sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, stream));
            SMP_HIP_TRY(hipStreamSynchronize(stream));
            x.k0 = k;
            x.e.assign(h_rowbuf, h_rowbuf + cnt);
            ++fetches;
          }
          e = x.e[k - x.k0];
        }
        ++props;
        const uint32_t w = pactae::pr_woman(e);
        if (w >= (uint32_t)n) {
          pactae::set_error("host tail: preference entry outside [0,n)");
          return SMP_ERR_INVALID;
        }
        const unsigned long long mine = pactae::pw_pack(pactae::pr_rank(e), (uint32_t)m, (uint32_t)k);
        const unsigned long long cur = h_partner[w];
        if (mine < cur) {
          h_partner[w] = mine;
          if (cur == pactae::kEmpty) break;
          m = (int)pactae::pw_man(cur);
          k = (int)pactae::pw_pos(cur) + 1;
        } else {
          ++k;
        }
      }
    }
    times.host_proposals = props;
    times.host_row_fetches = fetches;
    SMP_HIP_TRY(hipMemcpyAsync(d_partner, h_partner, (size_t)n * sizeof(unsigned long long),
                               hipMemcpyHostToDevice, stream));
    times.host_tail_ms = ms_since(t0);
    return SMP_OK;
  }

  // -- post (smp_engine_bamboo.cu:289-306) ----------------------------------
  int PostProc(bool perfect) override {
    if (perfect) {
      SMP_HIP_TRY(pactae::launch_first_match(d_pref_m, n, d_match, stream));
    } else {
      SMP_HIP_TRY(pactae::launch_post(d_partner, n, d_match, d_sc, stream));
    }
    SMP_HIP_TRY(hipMemcpyAsync(h_match.data(), d_match, (size_t)n * sizeof(int32_t),
                               hipMemcpyDeviceToHost, stream));
    SMP_HIP_TRY(hipMemcpyAsync(h_sc, d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, stream));
    SMP_HIP_TRY(hipStreamSynchronize(stream));
    if (int st = check_device_error("post")) return st;
    if (perfect) {
      long long p = n;  // every man proposes once
      times.proposals = p;
    } else {
      times.proposals = (int64_t)h_sc->proposals;
    }
    return SMP_OK;
  }

  int exec() {
    int st = FindStableMatching();
    if (st) return st;
    times.precheck_ms = precheck_time_;
    times.init_ms = init_time_;
    times.core_ms = core_time_;
    times.post_ms = post_time_;
    times.total_ms = precheck_time_ + init_time_ + core_time_ + post_time_;
    times.prefix_cols = K;
    times.tail_threshold = T;
    if (!times.is_perfect) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) times.rank_kernel_ms = ms;
      if (hipEventElapsedTime(&ms, ev[1], ev[2]) == hipSuccess) times.gather_kernel_ms = ms;
      if (hipEventElapsedTime(&ms, ev[3], ev[4]) == hipSuccess) times.propose_kernel_ms = ms;
    }
    have_result = true;
    return SMP_OK;
  }
};
