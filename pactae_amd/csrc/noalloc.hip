// noalloc.hip — smp_exec_device_noalloc: one solve of preference lists already
// resident in HBM that allocates nothing, for callers that cannot call hipMalloc
// (code under graph capture, tight serving loops, harnesses that forbid it).
//
// It runs the engine's GPU-only full-init pipeline (the kernels of smp_exec with
// init_mode FULL, tail_mode GPU, precheck 0): reset -> women's rank inversion ->
// PR prefix gather -> persistent proposal kernel -> counter fold -> matching.
// Everything beyond the caller's lists and result lives in one file-scope device
// array, `scratch`, of fixed size, which every GPU the library runs on holds
// until the process ends.  Its device address is looked up once per device with
// hipGetSymbolAddress (the array's name is a host-side stand-in, not a device
// address) and only that address is handed to kernels.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU
// (two host threads, or a call racing an earlier one's kernels on another
// stream) race on it.  Each call clears what must start cleared (control block,
// partner words) with a fill kernel and overwrites the rest before reading it.
#include <algorithm>
#include <mutex>
#include <string>

#include "kernels.hpp"

namespace pactae {
namespace {

constexpr int kNaMaxN = 4096;        // largest instance accepted
constexpr int kNaK = 128;            // PR prefix columns (positions past it read the rank matrix)
constexpr int kNaMaxWaves = 8192;    // proposal-kernel waves (the grid is clamped to it)
constexpr int kNaMaxDevices = 64;

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaPrEntries = (size_t)((kNaK + kPrBlock - 1) / kPrBlock) * kNaMaxN * kPrBlock;
// Layout of the scratch array (byte offsets).
constexpr size_t kOffSc = 0;
constexpr size_t kOffPartner = kOffSc + align256(sizeof(Scalars));
constexpr size_t kOffParked = kOffPartner + align256((size_t)kNaMaxN * sizeof(unsigned long long));
constexpr size_t kOffWstats = kOffParked + align256((size_t)kNaMaxN * sizeof(int2));
constexpr size_t kOffPr = kOffWstats + align256((size_t)kNaMaxWaves * 2 * sizeof(unsigned long long));
constexpr size_t kOffRank = kOffPr + align256(kNaPrEntries * sizeof(uint32_t));
constexpr size_t kScratchBytes = kOffRank + align256((size_t)kNaMaxN * kNaMaxN * sizeof(uint16_t));
static_assert(kScratchBytes < (36u << 20), "no-allocation scratch stays under 36 MiB");

}  // namespace

// Device memory of the no-allocation entry point: 35,848,448 bytes (34.2 MiB).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct DeviceScratch {
  std::once_flag once;
  hipError_t status = hipSuccess;
  unsigned char* ptr = nullptr;  // device address of `scratch` on this device
  int cus = 0;
};
DeviceScratch g_scratch[kNaMaxDevices];

// Once per device, with that device current: the device address of `scratch`
// and the CU count (no allocation in either).
int scratch_for(int dev, DeviceScratch** out) {
  DeviceScratch& d = g_scratch[dev];
  std::call_once(d.once, [&] {
    void* p = nullptr;
    d.status = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (d.status == hipSuccess) d.ptr = static_cast<unsigned char*>(p);
    if (d.status == hipSuccess)
      d.status = hipDeviceGetAttribute(&d.cus, hipDeviceAttributeMultiprocessorCount, dev);
  });
  if (d.status != hipSuccess || !d.ptr || d.cus <= 0) {
    set_error(std::string("smp_exec_device_noalloc: scratch address lookup failed: ") +
              hipGetErrorString(d.status));
    return SMP_ERR_HIP;
  }
  *out = &d;
  return SMP_OK;
}

}  // namespace
}  // namespace pactae

extern "C" int smp_exec_device_noalloc(const int32_t* d_pref_m, const int32_t* d_pref_w, int32_t n,
                                       int32_t device, int32_t* d_match, int64_t* proposals) {
  using namespace pactae;
  if (!d_pref_m || !d_pref_w || !d_match) {
    set_error("smp_exec_device_noalloc: null device pointer");
    return SMP_ERR_INVALID;
  }
  if (n <= 0 || n > kNaMaxN) {
    set_error("smp_exec_device_noalloc: n must be in [1, " + std::to_string(kNaMaxN) +
              "] (its fixed scratch holds at most that size), got " + std::to_string(n));
    return SMP_ERR_INVALID;
  }
  int count = 0;
  SMP_HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count || device >= kNaMaxDevices) {
    set_error("smp_exec_device_noalloc: device ordinal " + std::to_string(device) + " not present");
    return SMP_ERR_INVALID;
  }
  SMP_HIP_TRY(hipSetDevice(device));
  DeviceScratch* ds = nullptr;
  if (int st = scratch_for(device, &ds)) return st;
  unsigned char* base = ds->ptr;
  Scalars* sc = reinterpret_cast<Scalars*>(base + kOffSc);
  unsigned long long* partner = reinterpret_cast<unsigned long long*>(base + kOffPartner);
  int2* parked = reinterpret_cast<int2*>(base + kOffParked);
  unsigned long long* wstats = reinterpret_cast<unsigned long long*>(base + kOffWstats);
  uint32_t* pr = reinterpret_cast<uint32_t*>(base + kOffPr);
  uint16_t* rank = reinterpret_cast<uint16_t*>(base + kOffRank);

  const hipStream_t s = nullptr;
  const int K = std::min(n, kNaK);
  // Proposal grid as the engine sizes it (16 waves per CU, at most one per man).
  const long long waves = std::min<long long>(std::min<long long>((long long)ds->cus * 16, n), kNaMaxWaves);
  const int grid = (int)std::max<long long>(1, (waves + 3) / 4);
  SMP_HIP_TRY(launch_reset_scalars(sc, std::min(n, grid * 4), s, nullptr, 0, nullptr, 0, partner, (size_t)n));
  // One LDS pass per row (2n bytes, at most 8 KiB here).
  SMP_HIP_TRY(launch_invert_rows(d_pref_w, rank, n, n, /*validate=*/1, sc, s, /*nontemporal=*/1));
  SMP_HIP_TRY(launch_gather_pr(d_pref_m, rank, pr, n, K, sc, s));
  ProposeArgs a{};
  a.pr = pr;
  a.pref_m = d_pref_m;
  a.rank_w = rank;
  a.rank32 = nullptr;
  a.partner = partner;
  a.parked = parked;
  a.sc = sc;
  a.wstats = wstats;
  a.n = n;
  a.K = K;
  a.T = 0;  // the GPU finishes every chain
  a.coherent_hint = 0;
  a.poll = 16;
  a.starts = nullptr;
  a.nstarts = 0;
  a.rank_sparse = 0;  // the full rank matrix is built
  SMP_HIP_TRY(launch_propose(a, grid, s));
  SMP_HIP_TRY(launch_fold_wstats(wstats, grid * 4, sc, s));
  SMP_HIP_TRY(launch_post(partner, n, d_match, sc, s));
  Scalars h{};
  SMP_HIP_TRY(hipMemcpyAsync(&h, sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
  SMP_HIP_TRY(hipStreamSynchronize(s));
  if (h.error != 0) {
    std::string what;
    if (h.error & kErrEntryRange) what += " entry outside [0,n);";
    if (h.error & kErrNotPerm) what += " a woman's row is not a permutation;";
    if (h.error & kErrListEnd) what += " a man exhausted his list;";
    if (h.error & kErrUnmatched) what += " a woman was left unmatched;";
    set_error("invalid preference lists (smp_exec_device_noalloc):" + what);
    return SMP_ERR_INVALID;
  }
  if (proposals) *proposals = (int64_t)h.proposals;
  return SMP_OK;
}
