// smp_oracle.cpp — CPU oracle and CPU baseline for the stable-marriage proposal path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: tests/, the smoke entry and
// the cpu_baseline leg of bench.py load it (as oracle/_build/liboracle*.so).  The
// product library never links, loads or calls it, and it is never the thing
// measured as the GPU result.
//
// What it restates (algorithms re-derived in this file's own code; citations are
// relative to the reference repository root):
//   oracle_gs_seq   SmpEngineGsF9       src/smp/smp_engine_gs_1_f9.cu:13-79
//   oracle_mw_seq   SmpEngineMwF9       src/smp/smp_engine_mw_1_f9.cu:15-79
//   oracle_gs_par   SmpEngineGs2F9      src/smp/smp_engine_gs_2_f9.cu:19-121
//   oracle_mw_par   SmpEngineMw2F9      src/smp/smp_engine_mw_2_f9.cu:17-96
//   oracle_*_ranked the same four engines with the women's rank matrix given as
//                   input (the single-instance 2,000 configuration derives it on the
//                   host), so init skips the inversion; pinned by equality with
//                   the engines above and by brute force in the tests
//   oracle_is_perfect          SmpEngineBamboo::IsPerfect include/smp/smp_engine_bamboo.h:25-33
//   oracle_invert_rows         F9RankMatrixInitKernel src/smp/kernel.cu:672-686 (host form
//                              as in smp_engine_gs_1_f9.cu:15-20)
//   oracle_times               AbsSmpEngine::FindStableMatching include/smp/smp_engine_abs.h:10-36
// Independent checks (no reference counterpart; the reference has no tests):
//   oracle_count_blocking_pairs  stability check by definition
//   oracle_brute_force           n <= 9: enumerate all n! matchings, keep the stable
//                                ones, return the one every man weakly prefers
//   oracle_serial_dictatorship   closed form when all men share one list L:
//                                woman L[k] takes her favourite man still free
//
// Pinning (each function is checked by something that does not restate it):
//   gs_seq / mw_seq / gs_par / mw_par : brute force on all n <= 8 cases in the tests,
//       serial dictatorship closed form, first-choice closed form, mutual agreement,
//       blocking-pair count == 0, and a planted wrong matching that the checks reject.
//   count_blocking_pairs : hand-built unstable matchings with known blocking-pair
//       counts, and brute-force enumeration.
//   brute_force : the n = 1..3 cases worked by hand in the tests; lattice property.
//   serial_dictatorship : equals brute force for n <= 8 congested inputs.
//
// Time accounting mirrors the reference: precheck + init + core, steady clock, ms.
// Matchings are int32[n] indexed by man, holding the woman.

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <queue>
#include <set>
#include <thread>
#include <vector>
#include <algorithm>

namespace {

using Clock = std::chrono::steady_clock;

inline double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

inline size_t at(size_t row, size_t n, size_t col) { return row * n + col; }

// Row-wise inverse permutation: rank[w][pref[w][r]] = r.
// Host form of the rank-matrix init used by every F9 CPU engine
// (smp_engine_gs_1_f9.cu:15-20).
void invert_rows(int n, const int* pref, int* rank) {
  for (size_t w = 0; w < (size_t)n; ++w)
    for (size_t r = 0; r < (size_t)n; ++r)
      rank[at(w, n, (size_t)pref[at(w, n, r)])] = (int)r;
}

// match[m] = w from partner_rank[w] (the reference's PostProc, e.g.
// smp_engine_gs_1_f9.cu:70-79).
void post_from_partner_rank(int n, const int* pref_w, const int* partner_rank, int* match) {
  for (int w = 0; w < n; ++w) {
    int r = partner_rank[w];
    match[pref_w[at(w, n, r)]] = w;
  }
}

}  // namespace

extern "C" {

typedef struct oracle_times {
  double precheck_ms;
  double init_ms;
  double core_ms;
  double total_ms;
  long long proposals;  // every (man, woman) evaluation made by the core loop
} oracle_times;

static void fill_total(oracle_times* t) {
  if (t) t->total_ms = t->precheck_ms + t->init_ms + t->core_ms;
}

// ---------------------------------------------------------------------------
// GS-Seq (SmpEngineGsF9, smp_engine_gs_1_f9.cu:13-79).
// One FIFO queue of free men; the current man walks his list until accepted,
// then the displaced man (if any) is queued and the next free man is popped.
static int gs_seq_impl(int n, const int* pref_m, const int* pref_w, const int* rank_given, int* match,
                       oracle_times* t) {
  if (n <= 0 || !pref_m || !pref_w || !match) return 1;
  oracle_times tt{};
  auto t0 = Clock::now();
  tt.precheck_ms = ms_since(t0);  // IsPerfect() is `return false` in this engine

  t0 = Clock::now();
  std::vector<int> rank_own;
  const int* rank_w = rank_given;
  if (!rank_w) {
    rank_own.resize((size_t)n * n);
    invert_rows(n, pref_w, rank_own.data());
    rank_w = rank_own.data();
  }
  std::queue<int> free_men;
  std::vector<int> partner_rank(n), next(n);
  for (int i = 0; i < n; ++i) {
    free_men.push(i);
    partner_rank[i] = n;
    next[i] = 0;
  }
  tt.init_ms = ms_since(t0);

  t0 = Clock::now();
  long long props = 0;
  int m = free_men.front();
  free_men.pop();
  int k = 0;
  for (;;) {
    int w = pref_m[at(m, n, k)];
    int r = rank_w[at(w, n, m)];
    ++k;
    ++props;
    int p = partner_rank[w];
    if (r < p) {
      partner_rank[w] = r;
      if (p != n) free_men.push(pref_w[at(w, n, p)]);
      next[m] = k;
      if (free_men.empty()) break;
      m = free_men.front();
      free_men.pop();
      k = next[m];
    }
  }
  tt.core_ms = ms_since(t0);
  post_from_partner_rank(n, pref_w, partner_rank.data(), match);
  tt.proposals = props;
  fill_total(&tt);
  if (t) *t = tt;
  return 0;
}

// ---------------------------------------------------------------------------
// MW-Seq (SmpEngineMwF9, smp_engine_mw_1_f9.cu:15-79).
// McVitie-Wilson: men enter one at a time; each entering man's chain of
// displacements runs to completion before the next man enters.
static int mw_seq_impl(int n, const int* pref_m, const int* pref_w, const int* rank_given, int* match,
                       oracle_times* t) {
  if (n <= 0 || !pref_m || !pref_w || !match) return 1;
  oracle_times tt{};
  auto t0 = Clock::now();
  tt.precheck_ms = ms_since(t0);

  t0 = Clock::now();
  std::vector<int> rank_own;
  const int* rank_w = rank_given;
  if (!rank_w) {
    rank_own.resize((size_t)n * n);
    invert_rows(n, pref_w, rank_own.data());
    rank_w = rank_own.data();
  }
  std::vector<int> partner_rank(n, n), next(n, 0);
  tt.init_ms = ms_since(t0);

  t0 = Clock::now();
  long long props = 0;
  for (int start = 0; start < n; ++start) {
    int m = start;
    int k = next[m];
    for (;;) {
      int w = pref_m[at(m, n, k)];
      int r = rank_w[at(w, n, m)];
      ++k;
      ++props;
      int p = partner_rank[w];
      if (p == n) {
        next[m] = k;
        partner_rank[w] = r;
        break;
      }
      if (p > r) {
        next[m] = k;
        partner_rank[w] = r;
        m = pref_w[at(w, n, p)];
        k = next[m];
      }
    }
  }
  tt.core_ms = ms_since(t0);
  post_from_partner_rank(n, pref_w, partner_rank.data(), match);
  tt.proposals = props;
  fill_total(&tt);
  if (t) *t = tt;
  return 0;
}

// ---------------------------------------------------------------------------
// GS-Par (SmpEngineGs2F9, smp_engine_gs_2_f9.cu:19-121).
// Men are split into `threads` contiguous blocks, one private FIFO queue per
// thread; women's partner ranks are std::atomic<int> updated by CAS.  Displaced
// men join the queue of the thread that displaced them.  Thread count is
// clamped to n (the reference indexes an empty queue otherwise).
static int gs_par_impl(int n, const int* pref_m, const int* pref_w, const int* rank_given, int threads, int* match,
                       oracle_times* t) {
  if (n <= 0 || !pref_m || !pref_w || !match) return 1;
  if (threads <= 0) threads = (int)std::min(std::thread::hardware_concurrency(), 96u);
  threads = std::max(1, std::min(threads, n));
  oracle_times tt{};
  auto t0 = Clock::now();
  tt.precheck_ms = ms_since(t0);

  t0 = Clock::now();
  std::vector<int> rank_own;
  const int* rank_w = rank_given;
  if (!rank_w) {
    rank_own.resize((size_t)n * n);
    invert_rows(n, pref_w, rank_own.data());
    rank_w = rank_own.data();
  }
  std::vector<std::atomic<int>> partner(n);
  std::vector<std::atomic<int>> next(n);
  for (int i = 0; i < n; ++i) {
    partner[i].store(n);
    next[i].store(0, std::memory_order_relaxed);
  }
  std::vector<std::queue<int>> queues(threads);
  int per = n / threads;
  for (int i = 0; i < threads; ++i) {
    int s = i * per, e = (i != threads - 1) ? (i + 1) * per : n;
    for (int j = s; j < e; ++j) queues[i].push(j);
  }
  tt.init_ms = ms_since(t0);

  t0 = Clock::now();
  std::atomic<long long> props_total{0};
  auto task = [&](int tid) {
    std::queue<int> q = queues[tid];
    long long props = 0;
    int m = q.front();
    q.pop();
    int k = 0;
    bool done = false;
    while (!done) {
      int w = pref_m[at(m, n, k)];
      int r = rank_w[at(w, n, m)];
      ++k;
      ++props;
      int p = partner[w].load();
      if (p < r) continue;
      while (r < p) {
        int expected = p;
        if (partner[w].compare_exchange_strong(expected, r)) {
          next[m].store(k, std::memory_order_relaxed);
          if (p != n) q.push(pref_w[at(w, n, p)]);
          if (!q.empty()) {
            m = q.front();
            q.pop();
            k = next[m].load(std::memory_order_relaxed);
          } else {
            done = true;
          }
          p = r;
        } else {
          p = expected;
        }
      }
    }
    props_total += props;
  };
  std::vector<std::thread> pool;
  pool.reserve(threads);
  for (int i = 0; i < threads; ++i) pool.emplace_back(task, i);
  for (auto& th : pool) th.join();
  tt.core_ms = ms_since(t0);

  std::vector<int> pr(n);
  for (int w = 0; w < n; ++w) pr[w] = partner[w].load();
  post_from_partner_rank(n, pref_w, pr.data(), match);
  tt.proposals = props_total.load();
  fill_total(&tt);
  if (t) *t = tt;
  return 0;
}

// ---------------------------------------------------------------------------
// MW-Par (SmpEngineMw2F9, smp_engine_mw_2_f9.cu:17-96).
// Men split into ceil(n/threads)-sized blocks; each thread enters its men one
// at a time and follows each displacement chain; CAS on atomic partner ranks.
static int mw_par_impl(int n, const int* pref_m, const int* pref_w, const int* rank_given, int threads, int* match,
                       oracle_times* t) {
  if (n <= 0 || !pref_m || !pref_w || !match) return 1;
  if (threads <= 0) threads = (int)std::min(std::thread::hardware_concurrency(), 96u);
  threads = std::max(1, std::min(threads, n));
  oracle_times tt{};
  auto t0 = Clock::now();
  tt.precheck_ms = ms_since(t0);

  t0 = Clock::now();
  std::vector<int> rank_own;
  const int* rank_w = rank_given;
  if (!rank_w) {
    rank_own.resize((size_t)n * n);
    invert_rows(n, pref_w, rank_own.data());
    rank_w = rank_own.data();
  }
  std::vector<std::atomic<int>> partner(n);
  std::vector<std::atomic<int>> next(n);
  for (int i = 0; i < n; ++i) {
    partner[i].store(n);
    next[i].store(0, std::memory_order_relaxed);
  }
  tt.init_ms = ms_since(t0);

  t0 = Clock::now();
  std::atomic<long long> props_total{0};
  const size_t per = ((size_t)n + threads - 1) / threads;
  auto task = [&](int tid) {
    long long props = 0;
    size_t s = tid * per, e = std::min((size_t)n, s + per);
    for (size_t start = s; start < e; ++start) {
      int m = (int)start;
      int k = next[m].load(std::memory_order_relaxed);
      bool married = false;
      while (!married) {
        int w = pref_m[at(m, n, k)];
        int r = rank_w[at(w, n, m)];
        ++k;
        ++props;
        int p = partner[w].load();
        if (p < r) continue;
        while (r < p) {
          int expected = p;
          if (partner[w].compare_exchange_strong(expected, r)) {
            next[m].store(k, std::memory_order_relaxed);
            if (expected == n) {
              married = true;
            } else {
              m = pref_w[at(w, n, expected)];
              k = next[m].load(std::memory_order_relaxed);
            }
            break;
          }
          p = expected;
        }
      }
    }
    props_total += props;
  };
  std::vector<std::thread> pool;
  pool.reserve(threads);
  for (int i = 0; i < threads; ++i) pool.emplace_back(task, i);
  for (auto& th : pool) th.join();
  tt.core_ms = ms_since(t0);

  std::vector<int> pr(n);
  for (int w = 0; w < n; ++w) pr[w] = partner[w].load();
  post_from_partner_rank(n, pref_w, pr.data(), match);
  tt.proposals = props_total.load();
  fill_total(&tt);
  if (t) *t = tt;
  return 0;
}

// The four engines as the reference runs them (rank matrix built in init), and
// with the women's inverse rank matrix given (rank_w[w*n + m] = position of man m
// in woman w's list; init then only sets up the partner and queue state).
int oracle_gs_seq(int n, const int* pref_m, const int* pref_w, int* match, oracle_times* t) {
  return gs_seq_impl(n, pref_m, pref_w, nullptr, match, t);
}
int oracle_mw_seq(int n, const int* pref_m, const int* pref_w, int* match, oracle_times* t) {
  return mw_seq_impl(n, pref_m, pref_w, nullptr, match, t);
}
int oracle_gs_par(int n, const int* pref_m, const int* pref_w, int threads, int* match, oracle_times* t) {
  return gs_par_impl(n, pref_m, pref_w, nullptr, threads, match, t);
}
int oracle_mw_par(int n, const int* pref_m, const int* pref_w, int threads, int* match, oracle_times* t) {
  return mw_par_impl(n, pref_m, pref_w, nullptr, threads, match, t);
}
int oracle_gs_seq_ranked(int n, const int* pref_m, const int* pref_w, const int* rank_w, int* match,
                         oracle_times* t) {
  return rank_w ? gs_seq_impl(n, pref_m, pref_w, rank_w, match, t) : 1;
}
int oracle_mw_seq_ranked(int n, const int* pref_m, const int* pref_w, const int* rank_w, int* match,
                         oracle_times* t) {
  return rank_w ? mw_seq_impl(n, pref_m, pref_w, rank_w, match, t) : 1;
}
int oracle_gs_par_ranked(int n, const int* pref_m, const int* pref_w, const int* rank_w, int threads, int* match,
                         oracle_times* t) {
  return rank_w ? gs_par_impl(n, pref_m, pref_w, rank_w, threads, match, t) : 1;
}
int oracle_mw_par_ranked(int n, const int* pref_m, const int* pref_w, const int* rank_w, int threads, int* match,
                         oracle_times* t) {
  return rank_w ? mw_par_impl(n, pref_m, pref_w, rank_w, threads, match, t) : 1;
}

// ---------------------------------------------------------------------------
// IsPerfect (smp_engine_bamboo.h:25-33): true iff all first choices differ;
// then each man's first choice is the answer (smp_engine_bamboo.cu:291-297).
int oracle_is_perfect(int n, const int* pref_m) {
  std::set<int> firsts;
  for (int m = 0; m < n; ++m) firsts.insert(pref_m[at(m, n, 0)]);
  return (int)firsts.size() == n;
}

void oracle_invert_rows(int n, const int* pref, int* rank) { invert_rows(n, pref, rank); }

// ---------------------------------------------------------------------------
// Stability by definition.  Returns -1 if `match` is not a perfect matching
// (not a permutation of 0..n-1), else the number of blocking pairs (m, w):
// m prefers w to match[m] and w prefers m to her husband.
long long oracle_count_blocking_pairs(int n, const int* pref_m, const int* pref_w,
                                      const int* match) {
  std::vector<int> husband(n, -1);
  for (int m = 0; m < n; ++m) {
    int w = match[m];
    if (w < 0 || w >= n || husband[w] != -1) return -1;
    husband[w] = m;
  }
  std::vector<int> rank_w((size_t)n * n);
  invert_rows(n, pref_w, rank_w.data());
  long long blocking = 0;
  for (int m = 0; m < n; ++m) {
    const int* row = pref_m + at(m, n, 0);
    for (int k = 0; k < n && row[k] != match[m]; ++k) {
      int w = row[k];
      if (rank_w[at(w, n, m)] < rank_w[at(w, n, husband[w])]) ++blocking;
    }
  }
  return blocking;
}

// ---------------------------------------------------------------------------
// Brute force for n <= 9: enumerate every perfect matching, keep the stable
// ones, and return the one that every man weakly prefers to every other stable
// matching.  Returns the number of stable matchings found, or -1 on bad n, or
// -2 if no man-optimal one exists (which the theory rules out).
int oracle_brute_force(int n, const int* pref_m, const int* pref_w, int* match) {
  if (n <= 0 || n > 9) return -1;
  std::vector<int> rank_m((size_t)n * n), perm(n);
  invert_rows(n, pref_m, rank_m.data());
  for (int i = 0; i < n; ++i) perm[i] = i;
  std::vector<std::vector<int>> stable;
  do {
    if (oracle_count_blocking_pairs(n, pref_m, pref_w, perm.data()) == 0) stable.push_back(perm);
  } while (std::next_permutation(perm.begin(), perm.end()));
  for (const auto& cand : stable) {
    bool best = true;
    for (const auto& other : stable) {
      for (int m = 0; m < n && best; ++m)
        if (rank_m[at(m, n, cand[m])] > rank_m[at(m, n, other[m])]) best = false;
      if (!best) break;
    }
    if (best) {
      std::memcpy(match, cand.data(), sizeof(int) * n);
      return (int)stable.size();
    }
  }
  return -2;
}

// ---------------------------------------------------------------------------
// Closed form when every man's list is the same permutation L: going down L,
// woman L[k] takes her most-preferred man not yet taken (serial dictatorship
// by the women in L's order).
int oracle_serial_dictatorship(int n, const int* L, const int* pref_w, int* match) {
  std::vector<char> taken(n, 0);
  for (int k = 0; k < n; ++k) {
    int w = L[k];
    const int* row = pref_w + at(w, n, 0);
    int r = 0;
    while (taken[row[r]]) ++r;
    taken[row[r]] = 1;
    match[row[r]] = w;
  }
  return 0;
}

}  // extern "C"
