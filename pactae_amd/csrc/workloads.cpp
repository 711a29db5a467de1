// workloads.cpp — seeded preference-list generators and the batch plan.
//
// The reference's generators (include/utils/generate_workloads_flat.h:14-280) draw
// from std::random_device and libstdc++'s std::shuffle, so their outputs cannot be
// reproduced.  These generators keep the reference's workload definitions but are
// deterministic: row r of side s draws from its own xoshiro256** stream, seeded by
// SplitMix64 from (seed, side, row).  The same bytes therefore come out on any
// host and any thread count, which is what lets a GPU result be compared bit for
// bit with a CPU run on identical input.
//
// Sides: 0 = men's rows, 1 = women's rows, 2 and 3 = shared lists (L, P, ids).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "smp_internal.hpp"

namespace pactae {

namespace {

inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// xoshiro256** (Blackman & Vigna, 2018).
struct Xoshiro {
  uint64_t s[4];
  Xoshiro(uint64_t seed, uint64_t side, uint64_t row) {
    uint64_t x = seed ^ (side * 0xD1B54A32D192ED03ull) ^ (row * 0x8CB92BA72F3D8DD7ull);
    for (auto& v : s) v = splitmix64(x);
  }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // Unbiased draw in [0, range) (Lemire, "Fast random integer generation in an
  // interval", 2019): multiply-shift with rejection of the short first interval.
  uint32_t bounded(uint32_t range) {
    uint64_t x = next() >> 32;
    uint64_t m = x * (uint64_t)range;
    uint32_t l = (uint32_t)m;
    if (l < range) {
      uint32_t t = (uint32_t)(-range) % range;
      while (l < t) {
        x = next() >> 32;
        m = x * (uint64_t)range;
        l = (uint32_t)m;
      }
    }
    return (uint32_t)(m >> 32);
  }
  double uniform01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Fisher-Yates over a[0..len) (descending i, j uniform in [0, i]).
void shuffle(int32_t* a, int len, Xoshiro& g) {
  for (int i = len - 1; i > 0; --i) {
    int j = (int)g.bounded((uint32_t)i + 1);
    std::swap(a[i], a[j]);
  }
}

void identity(int32_t* a, int n) {
  for (int i = 0; i < n; ++i) a[i] = i;
}

void uniform_row(int32_t* row, int n, uint64_t seed, int side, int r) {
  Xoshiro g(seed, side, (uint64_t)r);
  identity(row, n);
  shuffle(row, n, g);
}

std::vector<int32_t> shared_perm(int n, uint64_t seed, int side) {
  std::vector<int32_t> v(n);
  uniform_row(v.data(), n, seed, side, 0);
  return v;
}

template <class F>
void parallel_rows(int n, int threads, F&& f) {
  if (threads <= 0) threads = (int)std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
  threads = std::max(1, std::min(threads, n));
  if (threads == 1) {
    for (int r = 0; r < n; ++r) f(r);
    return;
  }
  std::vector<std::thread> pool;
  const int per = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    int r0 = t * per, r1 = std::min(n, r0 + per);
    if (r0 >= r1) break;
    pool.emplace_back([=, &f] {
      for (int r = r0; r < r1; ++r) f(r);
    });
  }
  for (auto& th : pool) th.join();
}

// Reference RANDOM row (generate_workloads_flat.h:85-105): positions are cut into
// groups of g consecutive slots of a shared shuffled id order; each group is
// shuffled independently per row.  g >= n gives a uniform row.
void ref_random_side(int32_t* flat, int n, int g, uint64_t seed, int side, int ids_side,
                     int threads) {
  std::vector<int32_t> ids = shared_perm(n, seed, ids_side);
  parallel_rows(n, threads, [&](int r) {
    int32_t* row = flat + (size_t)r * n;
    std::copy(ids.begin(), ids.end(), row);
    Xoshiro gen(seed, side, (uint64_t)r);
    for (int s = 0; s < n; s += g) shuffle(row + s, std::min(g, n - s), gen);
  });
}

// Reference PERFECT row (generate_workloads_flat.h:50-82): a rotation starting at
// the row's distinct first choice, with the tail [1, n) shuffled.
void ref_perfect_side(int32_t* flat, int n, uint64_t seed, int side, int first_side,
                      int threads) {
  std::vector<int32_t> first = shared_perm(n, seed, first_side);
  parallel_rows(n, threads, [&](int r) {
    int32_t* row = flat + (size_t)r * n;
    for (int i = 0; i < n; ++i) row[i] = (first[r] + i) % n;
    Xoshiro gen(seed, side, (uint64_t)r);
    shuffle(row + 1, n - 1, gen);
  });
}

// Reference SOLO men (generate_workloads_flat.h:140-165): man m (m = n-2 for the
// last man) ranks w1 = m first, w2 = (m-1) mod (n-1) second to last and n-1 last;
// the other women fill ranks 1..n-3 in shuffled order.
void ref_solo_men(int32_t* flat, int n, uint64_t seed, int threads) {
  parallel_rows(n, threads, [&](int r) {
    int32_t* row = flat + (size_t)r * n;
    int m = (r == n - 1) ? n - 2 : r;
    int w1 = m, w2 = (m - 1 + n - 1) % (n - 1), wl = n - 1;
    row[0] = w1;
    row[n - 2] = w2;
    row[n - 1] = wl;
    int w = 0;
    for (int k = 1; k < n - 2; ++k) {
      while (w == w1 || w == w2) ++w;
      row[k] = w++;
    }
    Xoshiro gen(seed, 0, (uint64_t)r);
    shuffle(row + 1, n - 3, gen);
  });
}

// Reference SOLO women (generate_workloads_flat.h:193-222): woman w < n-1 ranks
// m1 = (w+1) mod (n-1) first and m2 = (m1 + n - 1) mod n second, the rest shuffled;
// the last woman's list is a full shuffle.
void ref_solo_women(int32_t* flat, int n, uint64_t seed, int threads) {
  parallel_rows(n, threads, [&](int r) {
    int32_t* row = flat + (size_t)r * n;
    Xoshiro gen(seed, 1, (uint64_t)r);
    if (r < n - 1) {
      int m1 = (r + 1) % (n - 1), m2 = (m1 + n - 1) % n;
      row[0] = m1;
      row[1] = m2;
      int m = 0;
      for (int k = 2; k < n; ++k) {
        while (m == m1 || m == m2) ++m;
        row[k] = m++;
      }
      shuffle(row + 2, n - 2, gen);
    } else {
      for (int k = 0; k < n; ++k) row[k] = n - 1 - k;
      shuffle(row, n, gen);
    }
  });
}

}  // namespace

int generate(int kind, int n, uint64_t seed, int param, int32_t* pm, int32_t* pw, int threads) {
  if (n <= 0 || n > SMP_MAX_N || !pm || !pw) return SMP_ERR_INVALID;
  if ((kind == SMP_WL_REF_SOLO) && n < 4) return SMP_ERR_INVALID;
  auto uniform_side = [&](int32_t* flat, int side) {
    parallel_rows(n, threads, [&](int r) { uniform_row(flat + (size_t)r * n, n, seed, side, r); });
  };
  switch (kind) {
    case SMP_WL_UNIFORM:
      uniform_side(pm, 0);
      uniform_side(pw, 1);
      return SMP_OK;
    case SMP_WL_CONGESTED: {
      std::vector<int32_t> L = shared_perm(n, seed, 2);
      parallel_rows(n, threads, [&](int r) { std::copy(L.begin(), L.end(), pm + (size_t)r * n); });
      uniform_side(pw, 1);
      return SMP_OK;
    }
    case SMP_WL_SHIFTED: {
      std::vector<int32_t> L = shared_perm(n, seed, 2);
      parallel_rows(n, threads, [&](int r) {
        int32_t* row = pm + (size_t)r * n;
        for (int k = 0; k < n; ++k) row[k] = L[(size_t)(r + k) % n];
      });
      uniform_side(pw, 1);
      return SMP_OK;
    }
    case SMP_WL_CORRELATED: {
      std::vector<int32_t> base = shared_perm(n, seed, 2);
      const int swaps = param > 0 ? param : n / 10;
      parallel_rows(n, threads, [&](int r) {
        int32_t* row = pm + (size_t)r * n;
        std::copy(base.begin(), base.end(), row);
        Xoshiro g(seed, 0, (uint64_t)r);
        for (int t = 0; t < swaps; ++t) {
          int a = (int)g.bounded((uint32_t)n), b = (int)g.bounded((uint32_t)n);
          std::swap(row[a], row[b]);
        }
      });
      uniform_side(pw, 1);
      return SMP_OK;
    }
    case SMP_WL_REF_RANDOM: {
      int g = param > 0 ? param : 5;
      ref_random_side(pm, n, g, seed, 0, 2, threads);
      ref_random_side(pw, n, g, seed, 1, 3, threads);
      return SMP_OK;
    }
    case SMP_WL_REF_CONGESTED: {
      std::vector<int32_t> L = shared_perm(n, seed, 2), P = shared_perm(n, seed, 3);
      parallel_rows(n, threads, [&](int r) {
        std::copy(L.begin(), L.end(), pm + (size_t)r * n);
        std::copy(P.begin(), P.end(), pw + (size_t)r * n);
      });
      return SMP_OK;
    }
    case SMP_WL_REF_PERFECT:
      ref_perfect_side(pm, n, seed, 0, 2, threads);
      ref_perfect_side(pw, n, seed, 1, 3, threads);
      return SMP_OK;
    case SMP_WL_REF_SOLO:
      ref_solo_men(pm, n, seed, threads);
      ref_solo_women(pw, n, seed, threads);
      return SMP_OK;
    default:
      return SMP_ERR_INVALID;
  }
}

int batch_plan(int count, int n_min, int n_max, uint64_t seed, const double* mix3, int32_t* sizes,
               int32_t* kinds) {
  if (count <= 0 || n_min <= 0 || n_max < n_min || n_max > SMP_MAX_N || !sizes || !kinds)
    return SMP_ERR_INVALID;
  double w[3] = {0.6, 0.2, 0.2};
  if (mix3)
    for (int i = 0; i < 3; ++i) w[i] = mix3[i] < 0 ? 0 : mix3[i];
  double tot = w[0] + w[1] + w[2];
  if (!(tot > 0)) return SMP_ERR_INVALID;
  static const int32_t kind_of[3] = {SMP_WL_UNIFORM, SMP_WL_CONGESTED, SMP_WL_CORRELATED};
  const double l0 = std::log((double)n_min), l1 = std::log((double)n_max + 1.0);
  for (int i = 0; i < count; ++i) {
    Xoshiro g(seed, 4, (uint64_t)i);
    double u = g.uniform01();
    int n = (int)std::floor(std::exp(l0 + u * (l1 - l0)));
    sizes[i] = std::max(n_min, std::min(n_max, n));
    double v = g.uniform01() * tot;
    kinds[i] = v < w[0] ? kind_of[0] : (v < w[0] + w[1] ? kind_of[1] : kind_of[2]);
  }
  return SMP_OK;
}

}  // namespace pactae
