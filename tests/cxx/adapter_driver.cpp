// adapter_driver.cpp — runs the adapter the way the reference's executor runs an
// engine (SmpEngineExecutor, benchmark/smp_engine_exec.h:6-32): build the SmpObj
// from flat lists, construct the engine, FindStableMatching, PrintProfilingInfo.
// Usage: adapter_driver <n> <lists.bin> <match.bin>
//   lists.bin: n*n int32 men's lists, then n*n int32 women's lists
//   match.bin: written, n int32 (man -> woman)
#include <cstdio>
#include <fstream>

#include "engine_frame.h"
#include "smp_engine_bamboo_mi355x.h"

template <typename Engine>
static Matching Execute(const bamboosmp::SmpObj& smp, size_t n) {
  Engine engine(smp, n);
  std::cout << "Executing " << engine.GetEngineName() << ", size: " << n << std::endl;
  Matching m = engine.FindStableMatching();
  engine.PrintProfilingInfo();
  const smp_times& t = engine.LibraryTimes();
  std::cout << "library: is_perfect " << t.is_perfect << " proposals " << t.proposals << std::endl;
  return m;
}

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s n lists.bin match.bin\n", argv[0]);
    return 2;
  }
  const size_t n = std::stoul(argv[1]);
  std::vector<int> m(n * n), w(n * n);
  std::ifstream in(argv[2], std::ios::binary);
  in.read(reinterpret_cast<char*>(m.data()), (std::streamsize)(m.size() * sizeof(int)));
  in.read(reinterpret_cast<char*>(w.data()), (std::streamsize)(w.size() * sizeof(int)));
  if (!in) {
    std::fprintf(stderr, "short read of %s\n", argv[2]);
    return 2;
  }
  bamboosmp::SmpObj smp(std::move(m), std::move(w));
  try {
    Matching match = Execute<bamboosmp::SmpEngineBambooMI355X>(smp, n);
    std::ofstream out(argv[3], std::ios::binary);
    out.write(reinterpret_cast<const char*>(match.data()), (std::streamsize)(match.size() * sizeof(int)));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "engine error: %s\n", e.what());
    return 1;
  }
  return 0;
}
