// smp_engine_bamboo_mi355x.h — the adapter a maintainer adds to the reference
// tree (beside include/smp/smp_engine_bamboo.h) to run the MI355X engine behind
// the reference's own AbsSmpEngine / SmpEngineExecutor interface.  It talks to
// the engine only through the C ABI of include/smp_hip.h.
//
// The library runs precheck, init and core in one call.  IsPerfect() (the first
// hook FindStableMatching calls) makes that call and returns the library's
// first-choice verdict, so the framework skips InitProc/CoreProc and prints its
// "perfect matching" line only for instances that really are perfect.  InitProc
// and CoreProc are then no-ops, and PostProc copies the matching and puts the
// library's own phase times into the framework's timers.
#pragma once
#include <stdexcept>

#include "smp_hip.h"

namespace bamboosmp {
class SmpEngineBambooMI355X : public AbsSmpEngine {
 public:
  SmpEngineBambooMI355X(const SmpObj& smp, size_t n) : n_(n) {
    smp_opts o;
    smp_default_opts(&o);
    if (smp_create(smp.flatten_pref_lists_m, smp.flatten_pref_lists_w, (int32_t)n, &o, &h_) != SMP_OK)
      throw std::runtime_error(smp_last_error());
  }
  ~SmpEngineBambooMI355X() override { smp_destroy(h_); }
  String GetEngineName() const override { return "BambooSMP-MI355X"; }
  const smp_times& LibraryTimes() const { return t_; }

 protected:
  bool IsPerfect() const override {
    auto* self = const_cast<SmpEngineBambooMI355X*>(this);
    if (smp_exec(h_) != SMP_OK) throw std::runtime_error(smp_last_error());
    smp_profile(h_, &self->t_);
    return t_.is_perfect != 0;
  }
  void InitProc() override {}
  void CoreProc() override {}
  Matching PostProc() override {
    Matching m(n_);
    if (smp_result(h_, m.data()) != SMP_OK) throw std::runtime_error(smp_last_error());
    precheck_time_ = t_.precheck_ms;
    init_time_ = t_.init_ms;
    core_time_ = t_.core_ms + t_.post_ms;
    return m;
  }

 private:
  smp_engine* h_ = nullptr;
  size_t n_;
  smp_times t_{};
};
}  // namespace bamboosmp
