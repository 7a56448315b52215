// HIP executor of the kernel functors (gfx950): launch shapes, kernel timers,
// device memory.  Shared by both precision variants; every function is inline.
//
// Functors run as 256-thread workgroups (4 wavefronts): flat (k_run), on the
// point grid with an XCD-aware block order (k_pts), or as two-phase LDS tiles
// (k_tiles).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#ifndef NH_HD
#define NH_HD __host__ __device__
#endif

#include "../../include/nicamhip.h"

namespace nh {

template <class F>
__global__ __launch_bounds__(256) void k_run(F f, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) f(t);
}

// point grid: nx blocks per plane, ny planes (region levels, then pole levels),
// launched as one linear grid.  The hardware deals workgroups to the 8 XCDs
// round-robin; the remap gives each XCD a contiguous range of blocks, so the
// neighbouring rows (j +- 1) and levels (k +- 1) a stencil reads sit in the
// same XCD's L2.
constexpr int kXcds = 8;

template <class F>
__global__ __launch_bounds__(256) void k_pts(F f, int nx, int nblk) {
  const int L = blockIdx.x;
  const int per = (nblk + kXcds - 1) / kXcds;
  const int xcd = L % kXcds, idx = L / kXcds;
  // the first (nblk % kXcds) XCDs get 'per' blocks, the others one fewer
  const int rem = nblk % kXcds;
  int logical;
  if (rem == 0 || xcd < rem) logical = xcd * per + idx;
  else logical = rem * per + (xcd - rem) * (per - 1) + idx;
  const int plane = logical / nx;
  const int xb = logical - plane * nx;
  f(plane, xb * (int)blockDim.x + (int)threadIdx.x);
}

// column tiles: per block init(), then per level level_a, barrier, level_b;
// the number of levels is uniform over the block
template <class F>
__global__ __launch_bounds__(F::kThreads) void k_cols(F f, int nblk) {
  extern __shared__ __align__(16) unsigned char nh_lds_bytes[];
  typename F::LdsT* lds = reinterpret_cast<typename F::LdsT*>(nh_lds_bytes);
  const int L = blockIdx.x;
  const int per = (nblk + kXcds - 1) / kXcds;
  const int xcd = L % kXcds, idx = L / kXcds;
  const int rem = nblk % kXcds;
  const int blk = (rem == 0 || xcd < rem) ? xcd * per + idx : rem * per + (xcd - rem) * (per - 1) + idx;
  typename F::State st;
  const int nl = f.init(blk, (int)threadIdx.x, st);
  for (int s = 0; s < nl; ++s) {
    f.level_a(s, st, lds);
    __syncthreads();
    f.level_b(s, st, lds);
  }
}

// Kernel timers: one event pair per launch while enabled, recorded on the
// launch stream around the kernel; aggregated per kernel name on report.
struct HipProf {
  bool on = false;
  std::string only;                    // time only the launches of this kernel ("" = all)
  std::vector<hipEvent_t> ev;          // pairs (start, stop)
  size_t used = 0;                     // pairs in use
  std::vector<const char*> name;       // per used pair
};

// tiled kernels: LDS shared by the block (elements of F::LdsT), an optional
// staging phase (F::kStage), then two phases, with a barrier between phases
template <class F>
__global__ __launch_bounds__(256) void k_tiles(F f, int nblk) {
  extern __shared__ __align__(16) unsigned char nh_lds_bytes[];
  typename F::LdsT* lds = reinterpret_cast<typename F::LdsT*>(nh_lds_bytes);
  const int L = blockIdx.x;
  const int per = (nblk + kXcds - 1) / kXcds;
  const int xcd = L % kXcds, idx = L / kXcds;
  const int rem = nblk % kXcds;
  const int blk = (rem == 0 || xcd < rem) ? xcd * per + idx : rem * per + (xcd - rem) * (per - 1) + idx;
  if constexpr (F::kStage) {
    f.phase0(blk, (int)threadIdx.x, lds);
    __syncthreads();
  }
  f.phase1(blk, (int)threadIdx.x, lds);
  __syncthreads();
  f.phase2(blk, (int)threadIdx.x, lds);
}

class HipExec {
 public:
  HipExec(int device, void* stream, void** prof_slot);
  int prof_enable(bool on);
  int prof_only(const char* name) {
    if (!prof_) return NH_ERR_STATE;
    prof_->only = name ? name : "";
    return 0;
  }
  int prof_report(int max_kernels, char* names, double* total_ms, int64_t* launches);
  void prof_release();
  void* alloc(size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      note(e, "hipMalloc");
      return nullptr;
    }
    return p;
  }
  void release(void* p) {
    if (p) (void)hipFree(p);
  }
  void zero(void* p, size_t bytes) { note(hipMemset(p, 0, bytes), "hipMemset"); }
  void h2d(void* dst, const void* src, size_t bytes) {
    note(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
  void d2d(void* dst, const void* src, size_t bytes) {
    note(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s_), "hipMemcpyAsync D2D");
  }
  // ordered after the stream's work; returns when the bytes are on the host
  void d2h(void* dst, const void* src, size_t bytes) {
    note(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s_), "hipMemcpyAsync D2H");
    note(hipStreamSynchronize(s_), "hipStreamSynchronize");
  }
  template <class F>
  void run(const F& f, long n) {
    if (n <= 0) return;
    const int bs = 256;
    long nb = (n + bs - 1) / bs;
    hipEvent_t* pair = timed(F::kName) ? next_pair(F::kName) : nullptr;
    if (pair) note(hipEventRecord(pair[0], s_), "hipEventRecord");
    hipLaunchKernelGGL(k_run<F>, dim3((unsigned)nb), dim3(bs), 0, s_, f, n);
    note(hipGetLastError(), "kernel launch");
    if (pair) note(hipEventRecord(pair[1], s_), "hipEventRecord");
    if (sync_each_) note(hipStreamSynchronize(s_), "kernel execution");
  }
  template <class F, class D>
  void run_pts(const F& f, const D& d) {
    run_pts_n(f, d, d.lall * d.kall + 2 * d.kall);
  }
  // ny planes of the point grid (the functor maps its plane index)
  template <class F, class D>
  void run_pts_n(const F& f, const D& d, int ny) {
    const int bs = 256;
    const int nx = (d.g * d.g + bs - 1) / bs;
    const int nblk = nx * ny;
    hipEvent_t* pair = timed(F::kName) ? next_pair(F::kName) : nullptr;
    if (pair) note(hipEventRecord(pair[0], s_), "hipEventRecord");
    hipLaunchKernelGGL(k_pts<F>, dim3(nblk), dim3(bs), 0, s_, f, nx, nblk);
    note(hipGetLastError(), "kernel launch");
    if (pair) note(hipEventRecord(pair[1], s_), "hipEventRecord");
    if (sync_each_) note(hipStreamSynchronize(s_), "kernel execution");
  }
  template <class F>
  void run_tiles(const F& f, int nblocks, int lds_elems) {
    hipEvent_t* pair = timed(F::kName) ? next_pair(F::kName) : nullptr;
    if (pair) note(hipEventRecord(pair[0], s_), "hipEventRecord");
    hipLaunchKernelGGL(k_tiles<F>, dim3(nblocks), dim3(256), (size_t)lds_elems * sizeof(typename F::LdsT), s_, f,
                       nblocks);
    note(hipGetLastError(), "kernel launch");
    if (pair) note(hipEventRecord(pair[1], s_), "hipEventRecord");
    if (sync_each_) note(hipStreamSynchronize(s_), "kernel execution");
  }
  template <class F>
  void run_cols(const F& f, int nblocks, int lds_elems) {
    hipEvent_t* pair = timed(F::kName) ? next_pair(F::kName) : nullptr;
    if (pair) note(hipEventRecord(pair[0], s_), "hipEventRecord");
    hipLaunchKernelGGL(k_cols<F>, dim3(nblocks), dim3(F::kThreads), (size_t)lds_elems * sizeof(typename F::LdsT), s_, f,
                       nblocks);
    note(hipGetLastError(), "kernel launch");
    if (pair) note(hipEventRecord(pair[1], s_), "hipEventRecord");
    if (sync_each_) note(hipStreamSynchronize(s_), "kernel execution");
  }
  void sync() { note(hipStreamSynchronize(s_), "hipStreamSynchronize"); }
  int check(const char* what) {
    if (err_ == hipSuccess) note(hipGetLastError(), what);
    return err_ == hipSuccess ? 0 : NH_ERR_DEVICE;
  }
  std::string message() const {
    return std::string(where_) + ": " + hipGetErrorString(err_);
  }

 private:
  void note(hipError_t e, const char* where) {
    if (e != hipSuccess && err_ == hipSuccess) {
      err_ = e;
      where_ = where;
    }
  }
  bool timed(const char* name) const { return prof_ && prof_->on && (prof_->only.empty() || prof_->only == name); }
  hipEvent_t* next_pair(const char* name) {
    HipProf* p = prof_;
    if (p->used * 2 == p->ev.size()) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess) return nullptr;
      if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return nullptr; }
      p->ev.push_back(a);
      p->ev.push_back(b);
    }
    p->name.push_back(name);
    return &p->ev[2 * p->used++];
  }
  HipProf* prof_ = nullptr;
  hipStream_t s_;
  hipError_t err_ = hipSuccess;
  const char* where_ = "";
  bool sync_each_ = false;
};

inline HipExec::HipExec(int device, void* stream, void** prof_slot) : s_((hipStream_t)stream) {
  int cur = -1;
  err_ = hipGetDevice(&cur);
  if (err_ == hipSuccess && cur != device) err_ = hipSetDevice(device);
  const char* e = std::getenv("NH_SYNC");
  sync_each_ = (e && e[0] == '1');
  if (prof_slot) {
    if (!*prof_slot) *prof_slot = new HipProf();
    prof_ = static_cast<HipProf*>(*prof_slot);
  }
}

inline int HipExec::prof_enable(bool on) {
  if (!prof_) return NH_ERR_STATE;
  prof_->on = on;
  return 0;
}

inline int HipExec::prof_report(int max_kernels, char* names, double* total_ms, int64_t* launches) {
  if (!prof_) return 0;
  HipProf* p = prof_;
  int n = 0;
  for (size_t u = 0; u < p->used; ++u) {
    hipEvent_t a = p->ev[2 * u], b = p->ev[2 * u + 1];
    hipError_t e = hipEventSynchronize(b);
    if (e != hipSuccess) { note(e, "hipEventSynchronize"); return -1; }
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, a, b);
    if (e != hipSuccess) { note(e, "hipEventElapsedTime"); return -1; }
    int slot = -1;
    for (int q = 0; q < n; ++q)
      if (std::strncmp(names + 32 * q, p->name[u], 31) == 0) { slot = q; break; }
    if (slot < 0) {
      if (n >= max_kernels) continue;
      slot = n++;
      std::strncpy(names + 32 * slot, p->name[u], 31);
      names[32 * slot + 31] = 0;
      total_ms[slot] = 0.0;
      launches[slot] = 0;
    }
    total_ms[slot] += ms;
    launches[slot] += 1;
  }
  p->used = 0;
  p->name.clear();
  return n;
}

inline void HipExec::prof_release() {
  if (!prof_) return;
  for (hipEvent_t ev : prof_->ev) (void)hipEventDestroy(ev);
  delete prof_;
  prof_ = nullptr;
}

}  // namespace nh
