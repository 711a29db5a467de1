// engine_frame.h — a stand-alone base class with the shape of the reference's
// engine framework (AbsSmpEngine, include/smp/smp_engine_abs.h:10-66, driven by
// SmpEngineExecutor, benchmark/smp_engine_exec.h:6-32), written for this test so
// that the adapter in smp_engine_bamboo_mi355x.h compiles and runs here exactly
// as it would inside the reference tree.  Phases are timed with a steady clock;
// a perfect instance skips init and core and says so.
#pragma once
#include <chrono>
#include <cstdint>
#include <iostream>
#include <string>
#include <vector>

using Matching = std::vector<int>;
using String = std::string;
using ms_t = double;

namespace bamboosmp {
struct SmpObj {  // the flat lists the reference's engines borrow
  std::vector<int> flatten_pref_lists_m_vec, flatten_pref_lists_w_vec;
  int* flatten_pref_lists_m = nullptr;
  int* flatten_pref_lists_w = nullptr;
  SmpObj(std::vector<int> m, std::vector<int> w)
      : flatten_pref_lists_m_vec(std::move(m)), flatten_pref_lists_w_vec(std::move(w)) {
    flatten_pref_lists_m = flatten_pref_lists_m_vec.data();
    flatten_pref_lists_w = flatten_pref_lists_w_vec.data();
  }
};
}  // namespace bamboosmp

class AbsSmpEngine {
 public:
  virtual ~AbsSmpEngine() = default;
  Matching FindStableMatching() {
    auto t0 = std::chrono::steady_clock::now();
    const bool perfect = IsPerfect();
    precheck_time_ = since(t0);
    if (perfect) {
      std::cout << "Perfect matching identified; skipping execution steps." << std::endl;
    } else {
      t0 = std::chrono::steady_clock::now();
      InitProc();
      init_time_ = since(t0);
      BambooTransfer();
      t0 = std::chrono::steady_clock::now();
      CoreProc();
      core_time_ = since(t0);
    }
    return PostProc();
  }
  void PrintProfilingInfo() {
    const String name = GetEngineName();
    std::cout << name << " Precheck Time: " << precheck_time_ << " ms\n"
              << name << " Init Time: " << init_time_ << " ms\n"
              << name << " Core Time: " << core_time_ << " ms\n"
              << name << " Total Time: " << precheck_time_ + init_time_ + core_time_ << " ms" << std::endl;
  }
  virtual String GetEngineName() const = 0;

 protected:
  virtual bool IsPerfect() const = 0;
  virtual void InitProc() = 0;
  virtual void CoreProc() = 0;
  virtual Matching PostProc() = 0;
  virtual void BambooTransfer() {}
  ms_t precheck_time_ = 0, init_time_ = 0, core_time_ = 0;

 private:
  static double since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};
