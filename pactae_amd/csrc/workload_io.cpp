// workload_io.cpp — the reference's on-disk workload cache format.
//
// Format (include/utils/generate_workloads_flat.h:302-352): a text file whose first
// line is n, followed by n lines holding the men's rows and n lines holding the
// women's rows, each value followed by one space, each row ended by a newline.
// File names follow GetWorkloadFilename (:285-300):
//   smp_workload_random_<n>_<group>.txt | solo_<n> | congested_<n> | perfect_<n>
// Writing uses std::to_chars into a large buffer and reading uses std::from_chars
// over a memory image of the file, so a 20,000 x 20,000 pair (8e8 numbers) is
// bounded by disk bandwidth rather than iostream parsing.

#include <charconv>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "smp_internal.hpp"

namespace pactae {

namespace {

int write_rows(std::FILE* f, const int32_t* a, int n, std::vector<char>& buf) {
  for (int r = 0; r < n; ++r) {
    size_t pos = 0;
    const int32_t* row = a + (size_t)r * n;
    for (int c = 0; c < n; ++c) {
      auto res = std::to_chars(buf.data() + pos, buf.data() + buf.size(), row[c]);
      pos = (size_t)(res.ptr - buf.data());
      buf[pos++] = ' ';
    }
    buf[pos++] = '\n';
    if (std::fwrite(buf.data(), 1, pos, f) != pos) return SMP_ERR_INVALID;
  }
  return SMP_OK;
}

const char* skip_ws(const char* p, const char* end) {
  while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
  return p;
}

}  // namespace

int workload_save(const char* path, int n, const int32_t* pm, const int32_t* pw) {
  if (!path || n <= 0 || n > SMP_MAX_N || !pm || !pw) return SMP_ERR_INVALID;
  std::FILE* f = std::fopen(path, "wb");
  if (!f) {
    set_error(std::string("cannot open ") + path + " for writing");
    return SMP_ERR_INVALID;
  }
  std::vector<char> buf((size_t)n * 12 + 16);
  int st = std::fprintf(f, "%d\n", n) > 0 ? SMP_OK : SMP_ERR_INVALID;
  if (st == SMP_OK) st = write_rows(f, pm, n, buf);
  if (st == SMP_OK) st = write_rows(f, pw, n, buf);
  if (std::fclose(f) != 0) st = SMP_ERR_INVALID;
  if (st) set_error(std::string("write error on ") + path);
  return st;
}

int workload_load(const char* path, int* n_out, int32_t* pm, int32_t* pw, size_t capacity) {
  if (!path || !n_out) return SMP_ERR_INVALID;
  std::FILE* f = std::fopen(path, "rb");
  if (!f) {
    set_error(std::string("cannot open ") + path);
    return SMP_ERR_INVALID;
  }
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (size <= 0) {
    std::fclose(f);
    set_error(std::string("empty workload file ") + path);
    return SMP_ERR_INVALID;
  }
  // The header alone when the caller only asks for n.
  if (!pm || !pw) {
    char head[32] = {0};
    const size_t got = std::fread(head, 1, sizeof(head) - 1, f);
    std::fclose(f);
    int n = 0;
    const char* p = skip_ws(head, head + got);
    auto r = std::from_chars(p, head + got, n);
    if (r.ec != std::errc() || n <= 0 || n > SMP_MAX_N) {
      set_error(std::string("bad header in ") + path);
      return SMP_ERR_INVALID;
    }
    *n_out = n;
    return SMP_OK;
  }
  std::vector<char> text((size_t)size);
  const size_t got = std::fread(text.data(), 1, text.size(), f);
  std::fclose(f);
  const char* p = text.data();
  const char* end = p + got;
  int n = 0;
  p = skip_ws(p, end);
  auto r = std::from_chars(p, end, n);
  if (r.ec != std::errc() || n <= 0 || n > SMP_MAX_N) {
    set_error(std::string("bad header in ") + path);
    return SMP_ERR_INVALID;
  }
  p = r.ptr;
  const size_t nn = (size_t)n * n;
  if (capacity < nn) {
    set_error("workload_load: output capacity " + std::to_string(capacity) + " < n*n = " +
              std::to_string(nn));
    *n_out = n;
    return SMP_ERR_INVALID;
  }
  for (int32_t* dst : {pm, pw}) {
    for (size_t i = 0; i < nn; ++i) {
      p = skip_ws(p, end);
      auto q = std::from_chars(p, end, dst[i]);
      if (q.ec != std::errc()) {
        set_error(std::string("truncated or malformed workload file ") + path);
        return SMP_ERR_INVALID;
      }
      p = q.ptr;
    }
  }
  *n_out = n;
  return SMP_OK;
}

// GetWorkloadFilename with the reference's WorkloadType numbering
// (PERFECT = 0, RANDOM = 1, CONGESTED = 2, SOLO = 3; include/utils/types.h:85-90).
int workload_filename(int type, int n, int group, char* buf, size_t len) {
  std::string s = "smp_workload_";
  switch (type) {
    case 1: s += "random_" + std::to_string(n) + "_" + std::to_string(group); break;
    case 3: s += "solo_" + std::to_string(n); break;
    case 2: s += "congested_" + std::to_string(n); break;
    case 0: s += "perfect_" + std::to_string(n); break;
    default: return SMP_ERR_INVALID;
  }
  s += ".txt";
  if (!buf || len <= s.size()) return SMP_ERR_INVALID;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return SMP_OK;
}

}  // namespace pactae

extern "C" {

int smp_workload_save(const char* path, int32_t n, const int32_t* pm, const int32_t* pw) {
  return pactae::workload_save(path, n, pm, pw);
}

int smp_workload_load(const char* path, int32_t* n_out, int32_t* pm, int32_t* pw, size_t capacity) {
  return pactae::workload_load(path, n_out, pm, pw, capacity);
}

int smp_workload_filename(int32_t ref_type, int32_t n, int32_t group, char* buf, size_t len) {
  int st = pactae::workload_filename(ref_type, n, group, buf, len);
  if (st) pactae::set_error("smp_workload_filename: bad type or buffer too small");
  return st;
}

}  // extern "C"
