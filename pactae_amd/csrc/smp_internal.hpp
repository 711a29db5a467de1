// smp_internal.hpp — definitions shared by the kernels, the engines and the C ABI.
#pragma once

#include <cstdint>
#include <stdint.h>
#include <cstdio>
#include <string>

#include "../../include/smp_hip.h"

#if defined(__HIPCC__)
#define SMP_HD __host__ __device__ inline
#else
#define SMP_HD inline
#endif

namespace pactae {

// ---------------------------------------------------------------------------
// Men-major PR entry (the reference's PRNode {idx_, rank_}, include/utils/types.h:58-61,
// packed into 32 bits): woman in the high half, her rank of this man in the low half.
SMP_HD uint32_t pr_pack(uint32_t w, uint32_t r) { return (w << 16) | (r & 0xFFFFu); }
SMP_HD uint32_t pr_woman(uint32_t e) { return e >> 16; }
SMP_HD uint32_t pr_rank(uint32_t e) { return e & 0xFFFFu; }

// The PR prefix is stored column-blocked: blocks of kPrBlock list positions,
// each block holding every man's kPrBlock entries contiguously
// (index ((k / kPrBlock) * n + m) * kPrBlock + k % kPrBlock).  A 64-position
// segment of one man is still two or three whole 128-byte lines for the GPU, and
// the first block (positions 0..31 of every man, where nearly all late proposals
// fall) is one contiguous piece that the host can copy and cache first.
constexpr int kPrBlock = 32;
// A list entry the extension windows of the host tail could not fill (past the
// list's end, or malformed lists).
constexpr uint32_t kWinInvalid = 0xFFFFFFFFu;
SMP_HD size_t pr_index(int m, int k, int n) {
  return ((size_t)(k >> 5) * (size_t)n + (size_t)m) * kPrBlock + (size_t)(k & (kPrBlock - 1));
}
SMP_HD size_t pr_entries(int n, int K) {  // allocation for positions [0, K)
  return (size_t)((K + kPrBlock - 1) / kPrBlock) * (size_t)n * kPrBlock;
}

// Woman's partner word, claimed with a 64-bit atomicMin:
//   bits 42..62 rank of the man in her list (decides the min: ranks are distinct)
//   bits 21..41 the man
//   bits  0..20 the man's list position of this woman (so a displaced man
//               resumes at position + 1 without any per-man side array)
// An unclaimed woman holds all ones, which is larger than every real word.
constexpr uint64_t kEmpty = ~0ull;
constexpr int kFieldBits = 21;
constexpr uint64_t kFieldMask = (1ull << kFieldBits) - 1;
SMP_HD uint64_t pw_pack(uint32_t rank, uint32_t man, uint32_t pos) {
  return ((uint64_t)rank << 42) | ((uint64_t)man << 21) | (uint64_t)pos;
}
SMP_HD uint32_t pw_rank(uint64_t v) { return (uint32_t)(v >> 42); }
SMP_HD uint32_t pw_man(uint64_t v) { return (uint32_t)((v >> 21) & kFieldMask); }
SMP_HD uint32_t pw_pos(uint64_t v) { return (uint32_t)(v & kFieldMask); }

// Device-side status flags (bit set in Scalars::error).
enum : int {
  kErrEntryRange = 1,   // a preference entry outside [0, n)
  kErrNotPerm = 2,      // a woman's row is not a permutation
  kErrListEnd = 4,      // a man ran past the end of his list (rows not permutations)
  kErrUnmatched = 8,    // a woman left unclaimed at the end
};

// Small device-resident control block of one solve.
struct Scalars {
  int active;        // proposal waves that still hold work (a chain or unstarted men)
  int prefix_overflow;  // sparse prefix pass: a woman had more suitors than her bucket holds
  int parked_count;  // chains parked for the host tail
  int error;         // kErr* bits
  int dup_first;     // precheck: 1 if two men share a first choice
  int first_differs; // precheck: 1 if two men have different first choices
  int not_master;    // precheck: 1 if some man's list differs from man 0's
  int export_blocks; // k_export_state: blocks finished (the last one resets it)
  int restart;       // proposal kernel: deep chains seen, every chain parked for a
                     // restart over the whole PR matrix
  int women_differ;  // master-list path: some woman's list differs from woman 0's
  int scans_done;    // master-list path: scan workgroups finished
  int both_master;   // master-list path: every scan done, every man's list equal to
                     // man 0's and every woman's to woman 0's
  int pad0;
  unsigned long long deep_steps; // proposal kernel: steps past the PR prefix so far
                                 // (flushed by the waves every 256 steps)
  unsigned long long atomics;    // atomicMin attempts
  unsigned long long steps;      // segment scans
  unsigned long long proposals;  // sum over women of (partner's list position + 1)
};

// Thread-local last error text for the C ABI.
void set_error(const std::string& msg);
const char* last_error();

int generate(int kind, int n, uint64_t seed, int param, int32_t* pm, int32_t* pw, int threads);
int batch_plan(int count, int n_min, int n_max, uint64_t seed, const double* mix3, int32_t* sizes,
               int32_t* kinds);

}  // namespace pactae

#define SMP_HIP_TRY(expr)                                                                   \
  do {                                                                                      \
    hipError_t _st = (expr);                                                                \
    if (_st != hipSuccess) {                                                                \
      ::pactae::set_error(std::string(#expr) + " failed: " + hipGetErrorString(_st) + " (" + \
                          __FILE__ + ":" + std::to_string(__LINE__) + ")");                 \
      return SMP_ERR_HIP;                                                                   \
    }                                                                                       \
  } while (0)
