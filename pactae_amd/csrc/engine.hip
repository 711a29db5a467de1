// engine.hip — C ABI of the single-instance engine (declared in include/smp_hip.h);
// the engine itself is in engine.hpp.

#include "engine.hpp"

namespace pactae {

namespace {
thread_local std::string g_last_error;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

}  // namespace pactae

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

void smp_default_opts(smp_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->prefix_cols = 0;
  o->tail_mode = SMP_TAIL_AUTO;
  o->tail_threshold = 0;
  o->device = 0;
  o->precheck = 1;
  o->waves_per_cu = 0;
  o->validate = 1;
  o->coherent_hint = 0;
  o->poll_interval = 16;
  o->host_spin = 0;
  o->prefix_copy = 0;
  o->schedule = SMP_SCHEDULE_BAMBOO;
  o->init_mode = SMP_INIT_AUTO;
}

static int make_engine(int32_t n, const smp_opts* opts, smp_engine** out) {
  if (!out) {
    pactae::set_error("out is null");
    return SMP_ERR_INVALID;
  }
  *out = nullptr;
  if (n <= 0 || n > SMP_MAX_N) {
    pactae::set_error("n must be in [1, " + std::to_string(SMP_MAX_N) + "], got " +
                      std::to_string(n));
    return SMP_ERR_INVALID;
  }
  auto e = std::make_unique<smp_engine>();
  e->n = n;
  if (opts)
    e->opts = *opts;
  else
    smp_default_opts(&e->opts);
  if (e->opts.init_mode < SMP_INIT_AUTO || e->opts.init_mode > SMP_INIT_SPARSE ||
      e->opts.schedule < SMP_SCHEDULE_BAMBOO || e->opts.schedule >= SMP_SCHEDULE_COUNT ||
      e->opts.tail_window_entries < 0 || e->opts.start_order < 0 || e->opts.start_order > 1) {
    pactae::set_error("opts: unknown init_mode, schedule or start_order, or a negative tail_window_entries");
    return SMP_ERR_INVALID;
  }
  if (int st = e->setup()) return st;
  *out = e.release();
  return SMP_OK;
}

int smp_create(const int32_t* pref_m, const int32_t* pref_w, int32_t n, const smp_opts* opts,
               smp_engine** out) {
  if (!pref_m || !pref_w) {
    pactae::set_error("null preference list");
    return SMP_ERR_INVALID;
  }
  smp_engine* e = nullptr;
  if (int st = make_engine(n, opts, &e)) return st;
  const size_t bytes = (size_t)n * n * sizeof(int32_t);
  int32_t *dm = nullptr, *dw = nullptr;
  if (hipMalloc(&dm, bytes) != hipSuccess || hipMalloc(&dw, bytes) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(dm);
    pactae::set_error("device allocation of the preference lists failed");
    delete e;
    return SMP_ERR_NOMEM;
  }
  e->d_pref_m = dm;
  e->d_pref_w = dw;
  e->owns_lists = true;
  if (hipMemcpy(dm, pref_m, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dw, pref_w, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    pactae::set_error("H2D copy of the preference lists failed");
    delete e;
    return SMP_ERR_HIP;
  }
  *out = e;
  return SMP_OK;
}

int smp_create_device(const int32_t* d_pref_m, const int32_t* d_pref_w, int32_t n,
                      const smp_opts* opts, smp_engine** out) {
  if (!d_pref_m || !d_pref_w) {
    pactae::set_error("null device preference list");
    return SMP_ERR_INVALID;
  }
  smp_engine* e = nullptr;
  if (int st = make_engine(n, opts, &e)) return st;
  e->d_pref_m = d_pref_m;
  e->d_pref_w = d_pref_w;
  e->owns_lists = false;
  *out = e;
  return SMP_OK;
}

// Ranks given (the women's inverse rank matrix derived by the caller): no
// women's lists, so only this build's pipeline: the reference's walks map ranks
// back to men through the women's lists, and they and the frontier schedule
// build the rank matrix from them.
static int check_ranked(const smp_opts* opts) {
  if (opts && opts->schedule != SMP_SCHEDULE_BAMBOO) {
    pactae::set_error("opts: the MW/LA and frontier schedules need the women's lists, not their ranks");
    return SMP_ERR_INVALID;
  }
  return SMP_OK;
}

int smp_create_ranked(const int32_t* pref_m, const int32_t* rank_w, int32_t n, const smp_opts* opts,
                      smp_engine** out) {
  if (!pref_m || !rank_w) {
    pactae::set_error("null preference list or rank matrix");
    return SMP_ERR_INVALID;
  }
  if (int st = check_ranked(opts)) return st;
  smp_engine* e = nullptr;
  if (int st = make_engine(n, opts, &e)) return st;
  const size_t bytes = (size_t)n * n * sizeof(int32_t);
  int32_t *dm = nullptr, *dr = nullptr;
  if (hipMalloc(&dm, bytes) != hipSuccess || hipMalloc(&dr, bytes) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(dm);
    pactae::set_error("device allocation of the preference lists failed");
    delete e;
    return SMP_ERR_NOMEM;
  }
  e->d_pref_m = dm;
  e->d_rank_in = dr;
  e->owns_lists = true;
  if (hipMemcpy(dm, pref_m, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dr, rank_w, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    pactae::set_error("H2D copy of the preference lists failed");
    delete e;
    return SMP_ERR_HIP;
  }
  *out = e;
  return SMP_OK;
}

int smp_create_device_ranked(const int32_t* d_pref_m, const int32_t* d_rank_w, int32_t n, const smp_opts* opts,
                             smp_engine** out) {
  if (!d_pref_m || !d_rank_w) {
    pactae::set_error("null device preference list or rank matrix");
    return SMP_ERR_INVALID;
  }
  if (int st = check_ranked(opts)) return st;
  smp_engine* e = nullptr;
  if (int st = make_engine(n, opts, &e)) return st;
  e->d_pref_m = d_pref_m;
  e->d_rank_in = d_rank_w;
  e->owns_lists = false;
  *out = e;
  return SMP_OK;
}

int smp_exec(smp_engine* e) {
  if (!e) {
    pactae::set_error("null engine");
    return SMP_ERR_INVALID;
  }
  if (hipSetDevice(e->opts.device) != hipSuccess) {
    pactae::set_error("hipSetDevice failed");
    return SMP_ERR_HIP;
  }
  return e->exec();
}

int smp_result(smp_engine* e, int32_t* match) {
  if (!e || !match) {
    pactae::set_error("null argument");
    return SMP_ERR_INVALID;
  }
  if (!e->have_result) {
    pactae::set_error("smp_result before a successful smp_exec");
    return SMP_ERR_STATE;
  }
  std::memcpy(match, e->h_match, sizeof(int32_t) * e->n);
  return SMP_OK;
}

int smp_result_device(smp_engine* e, const int32_t** d_match) {
  if (!e || !d_match) {
    pactae::set_error("null argument");
    return SMP_ERR_INVALID;
  }
  if (!e->have_result) {
    pactae::set_error("smp_result_device before a successful smp_exec");
    return SMP_ERR_STATE;
  }
  return e->device_match(d_match);
}

int smp_profile(const smp_engine* e, smp_times* t) {
  if (!e || !t) {
    pactae::set_error("null argument");
    return SMP_ERR_INVALID;
  }
  *t = e->times;
  return SMP_OK;
}

int smp_init_exec(smp_engine* e, int32_t init_pr, smp_init_times* t) {
  if (!e || !t) {
    pactae::set_error("null argument");
    return SMP_ERR_INVALID;
  }
  SMP_HIP_TRY(hipSetDevice(e->opts.device));
  return e->init_only(init_pr != 0, t);
}

int smp_tune(smp_engine* e, int32_t trials, smp_tune_result* out) {
  if (!e) {
    pactae::set_error("null engine");
    return SMP_ERR_INVALID;
  }
  SMP_HIP_TRY(hipSetDevice(e->opts.device));
  return e->tune(trials, out);
}

int smp_rank_matrix(smp_engine* e, int32_t* rank_out) {
  if (!e || !rank_out) {
    pactae::set_error("null argument");
    return SMP_ERR_INVALID;
  }
  if (!e->matrices_valid) {
    pactae::set_error("smp_rank_matrix: the last run built no rank matrix (first-choice or master-list path)");
    return SMP_ERR_STATE;
  }
  const size_t nn = (size_t)e->n * e->n;
  if (e->d_rank_in) {  // given as input
    SMP_HIP_TRY(hipStreamSynchronize(e->stream));
    SMP_HIP_TRY(hipMemcpy(rank_out, e->d_rank_in, nn * sizeof(int32_t), hipMemcpyDeviceToHost));
    return SMP_OK;
  }
  if (int st = e->ensure_rank_matrix()) return st;  // a sparse prefix pass skips it
  SMP_HIP_TRY(hipStreamSynchronize(e->stream));
  std::vector<uint16_t> r16(nn);
  SMP_HIP_TRY(hipMemcpy(r16.data(), e->d_rank_w, nn * sizeof(uint16_t), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < nn; ++i) rank_out[i] = (int32_t)r16[i];
  return SMP_OK;
}

int smp_pr_matrix(smp_engine* e, int32_t* pr_out, int32_t cols_cap, int32_t* cols) {
  if (!e || !cols) {
    pactae::set_error("null argument");
    return SMP_ERR_INVALID;
  }
  if (!e->matrices_valid) {
    pactae::set_error("smp_pr_matrix: the last run built no PR prefix (first-choice or master-list path)");
    return SMP_ERR_STATE;
  }
  const int n = e->n, K = e->K;
  *cols = K;
  if (!pr_out) return SMP_OK;  // size query
  if (cols_cap < K) {
    pactae::set_error("smp_pr_matrix: the buffer holds " + std::to_string(cols_cap) + " columns per man, the PR prefix has " +
                      std::to_string(K));
    return SMP_ERR_INVALID;
  }
  SMP_HIP_TRY(hipStreamSynchronize(e->stream));
  std::vector<uint32_t> packed(pactae::pr_entries(n, K));
  SMP_HIP_TRY(hipMemcpy(packed.data(), e->d_pr, packed.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  for (int m = 0; m < n; ++m)
    for (int k = 0; k < K; ++k) {
      const uint32_t v = packed[pactae::pr_index(m, k, n)];
      const size_t i = (size_t)m * K + k;
      pr_out[2 * i] = (int32_t)pactae::pr_woman(v);
      pr_out[2 * i + 1] = (int32_t)pactae::pr_rank(v);
    }
  return SMP_OK;
}

void smp_destroy(smp_engine* e) { delete e; }

int smp_generate(int32_t kind, int32_t n, uint64_t seed, int32_t param, int32_t* pref_m,
                 int32_t* pref_w, int32_t threads) {
  int st = pactae::generate(kind, n, seed, param, pref_m, pref_w, threads);
  if (st) pactae::set_error("smp_generate: bad kind or size");
  return st;
}

int smp_batch_plan(int32_t count, int32_t n_min, int32_t n_max, uint64_t seed, const double* mix3,
                   int32_t* sizes, int32_t* kinds) {
  int st = pactae::batch_plan(count, n_min, n_max, seed, mix3, sizes, kinds);
  if (st) pactae::set_error("smp_batch_plan: bad arguments");
  return st;
}

int smp_device_count(int32_t* count) {
  int c = 0;
  if (!count) return SMP_ERR_INVALID;
  hipError_t st = hipGetDeviceCount(&c);
  if (st != hipSuccess) {
    *count = 0;
    pactae::set_error(std::string("hipGetDeviceCount: ") + hipGetErrorString(st));
    return SMP_ERR_HIP;
  }
  *count = c;
  return SMP_OK;
}

int smp_device_alloc(size_t bytes, void** p) {
  if (!p) return SMP_ERR_INVALID;
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    pactae::set_error("hipMalloc of " + std::to_string(bytes) + " bytes failed");
    return SMP_ERR_NOMEM;
  }
  return SMP_OK;
}

int smp_device_free(void* p) {
  SMP_HIP_TRY(hipFree(p));
  return SMP_OK;
}

int smp_copy_h2d(void* d, const void* h, size_t bytes) {
  SMP_HIP_TRY(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
  return SMP_OK;
}

int smp_copy_d2h(void* h, const void* d, size_t bytes) {
  SMP_HIP_TRY(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
  return SMP_OK;
}

int smp_device_copy(void* d, const void* src, size_t bytes) {
  if (!d || !src || bytes % 16 != 0) {
    pactae::set_error("smp_device_copy: null pointer or size not a multiple of 16");
    return SMP_ERR_INVALID;
  }
  int dev = 0;
  SMP_HIP_TRY(hipGetDevice(&dev));
  int cus = 0;
  SMP_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  SMP_HIP_TRY(pactae::launch_copy16(d, src, bytes, cus, nullptr));
  SMP_HIP_TRY(hipDeviceSynchronize());
  return SMP_OK;
}

int smp_device_sync(void) {
  SMP_HIP_TRY(hipDeviceSynchronize());
  return SMP_OK;
}

const char* smp_last_error(void) { return pactae::last_error(); }
const char* smp_version(void) { return "pactae_amd 0.1.0 (gfx950)"; }

}  // extern "C"
