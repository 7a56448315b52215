// Host executor of the nicamhip kernel functors, for CPU tests of the kernel
// source: every functor runs as a loop over its threads (OpenMP across
// blocks).  Shared by both precision variants of the host build.  Test
// infrastructure: the product library is the gfx950 build of csrc/.
#pragma once

#define NH_HD
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nicamhip.h"

namespace nh {

struct HostProf {
  bool on = false;
  std::string only;
  std::vector<std::pair<const char*, double>> rec;
};

class HostExec {
 public:
  HostExec(int, void*, void** prof_slot) {
    if (prof_slot) {
      if (!*prof_slot) *prof_slot = new HostProf();
      prof_ = static_cast<HostProf*>(*prof_slot);
    }
  }
  int prof_only(const char* name) {
    if (!prof_) return -2;
    prof_->only = name ? name : "";
    return 0;
  }
  int prof_enable(bool on) {
    if (!prof_) return -2;
    prof_->on = on;
    return 0;
  }
  int prof_report(int max_kernels, char* names, double* total_ms, int64_t* launches) {
    if (!prof_) return 0;
    int n = 0;
    for (auto& r : prof_->rec) {
      int slot = -1;
      for (int q = 0; q < n; ++q)
        if (std::strncmp(names + 32 * q, r.first, 31) == 0) slot = q;
      if (slot < 0) {
        if (n >= max_kernels) continue;
        slot = n++;
        std::strncpy(names + 32 * slot, r.first, 31);
        names[32 * slot + 31] = 0;
        total_ms[slot] = 0.0;
        launches[slot] = 0;
      }
      total_ms[slot] += r.second;
      launches[slot] += 1;
    }
    prof_->rec.clear();
    return n;
  }
  void prof_release() {
    delete prof_;
    prof_ = nullptr;
  }
  void* alloc(size_t bytes) { return std::aligned_alloc(256, (bytes + 255) / 256 * 256); }
  void release(void* p) { std::free(p); }
  void zero(void* p, size_t bytes) { std::memset(p, 0, bytes); }
  void h2d(void* dst, const void* src, size_t bytes) { std::memcpy(dst, src, bytes); }
  void d2d(void* dst, const void* src, size_t bytes) { std::memmove(dst, src, bytes); }
  void d2h(void* dst, const void* src, size_t bytes) { std::memcpy(dst, src, bytes); }
  template <class F>
  void run(const F& f, long n) {
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(static)
    for (long t = 0; t < n; ++t) f(t);
    if (prof_ && prof_->on && (prof_->only.empty() || prof_->only == F::kName))
      prof_->rec.emplace_back(F::kName,
                              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  }
  template <class F, class D>
  void run_pts(const F& f, const D& d) {
    run_pts_n(f, d, d.lall * d.kall + 2 * d.kall);
  }
  template <class F, class D>
  void run_pts_n(const F& f, const D& d, int ny) {
    auto t0 = std::chrono::steady_clock::now();
    const int nx = d.g * d.g;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) f(y, x);
    if (prof_ && prof_->on && (prof_->only.empty() || prof_->only == F::kName))
      prof_->rec.emplace_back(F::kName,
                              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  }
  template <class F>
  void run_tiles(const F& f, int nblocks, int lds_elems) {
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel
    {
      std::vector<typename F::LdsT> lds((size_t)lds_elems + 1);
#pragma omp for schedule(static)
      for (int b = 0; b < nblocks; ++b) {
        if constexpr (F::kStage)
          for (int t = 0; t < 256; ++t) f.phase0(b, t, lds.data());
        for (int t = 0; t < 256; ++t) f.phase1(b, t, lds.data());
        for (int t = 0; t < 256; ++t) f.phase2(b, t, lds.data());
      }
    }
    if (prof_ && prof_->on && (prof_->only.empty() || prof_->only == F::kName))
      prof_->rec.emplace_back(F::kName,
                              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  }
  template <class F>
  void run_cols(const F& f, int nblocks, int lds_elems) {
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel
    {
      std::vector<typename F::LdsT> lds((size_t)lds_elems + 1);
      constexpr int nt = F::kThreads;
      std::vector<typename F::State> st(nt);
#pragma omp for schedule(static)
      for (int b = 0; b < nblocks; ++b) {
        int nl = 0;
        for (int t = 0; t < nt; ++t) nl = f.init(b, t, st[t]);
        for (int s = 0; s < nl; ++s) {
          for (int t = 0; t < nt; ++t) f.level_a(s, st[t], lds.data());
          for (int t = 0; t < nt; ++t) f.level_b(s, st[t], lds.data());
        }
      }
    }
    if (prof_ && prof_->on && (prof_->only.empty() || prof_->only == F::kName))
      prof_->rec.emplace_back(F::kName,
                              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  }
  void sync() {}
  int check(const char*) { return 0; }
  std::string message() const { return "host"; }

 private:
  HostProf* prof_ = nullptr;
};

}  // namespace nh

