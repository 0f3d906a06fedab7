// Device helpers shared by the HIP kernels of the hot path (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "mdft_tables.hpp"

#define MDFT_HIP_CHECK(expr)                                                                       \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      throw ::mdft::Error(std::string(#expr " failed: ") + hipGetErrorString(e_) + " at " __FILE__ \
                          ":" + std::to_string(__LINE__));                                         \
  } while (0)

namespace mdft {

template <typename T>
struct cvec;
template <>
struct cvec<double> {
  using type = double2;
};
template <>
struct cvec<float> {
  using type = float2;
};
template <typename T>
using cplx = typename cvec<T>::type;

template <typename T>
__host__ __device__ inline cplx<T> mkc(T re, T im) {
  cplx<T> z;
  z.x = re;
  z.y = im;
  return z;
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return mkc<T>(a.x + b.x, a.y + b.y);
}
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
  return mkc<T>(a.x - b.x, a.y - b.y);
}
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mkc<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// acc + a * b
template <typename T>
__device__ __forceinline__ cplx<T> cfma(cplx<T> acc, cplx<T> a, cplx<T> b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}
// acc + a * conj(b)
template <typename T>
__device__ __forceinline__ cplx<T> cfma_conjb(cplx<T> acc, cplx<T> a, cplx<T> b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.y, b.x, acc.y);
  acc.y = fma(-a.x, b.y, acc.y);
  return acc;
}
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) {
  return mkc<T>(a.x, -a.y);
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) {
  return mkc<T>(a.x * s, a.y * s);
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Stage boundaries of one functional evaluation (the reference's Solver
// timers, src/mini-mdft.hpp:140-192): start | angl2proj | FFT fwd | OZ |
// FFT bwd | proj2angl + delta_f.
constexpr int kStageStart = 0, kStageConvStart = 1, kStageFftFwd = 2, kStageOz = 3, kStageFftBwd = 4,
              kStageEnd = 5, kNumMarks = 6, kNumStages = 5;

// hipEvent marks recorded on the work stream, one set per evaluation, read
// back only when collected (no synchronisation between evaluations).  With
// oz_only, mark() records nothing (each recorded event leaves the GPU idle
// for a few microseconds between kernels): the OZ launch itself records the
// set's FftFwd and Oz events as its start and end timestamps
// (hipExtLaunchKernel), and only that stage accumulates.
class StageTimer {
 public:
  explicit StageTimer(bool oz_only = false) : oz_only_(oz_only) {}
  ~StageTimer() {
    for (auto& set : sets_)
      for (hipEvent_t e : set) (void)hipEventDestroy(e);
  }
  StageTimer(const StageTimer&) = delete;
  StageTimer& operator=(const StageTimer&) = delete;

  void begin_eval() {
    if (used_ == kMaxSets) collect();
    if (used_ == static_cast<int>(sets_.size())) {
      std::array<hipEvent_t, kNumMarks> set{};
      for (auto& e : set) MDFT_HIP_CHECK(hipEventCreate(&e));
      sets_.push_back(set);
    }
    cur_ = used_++;
  }
  void mark(hipStream_t s, int i) {
    if (oz_only_) return;
    MDFT_HIP_CHECK(hipEventRecord(sets_[cur_][i], s));
  }
  bool oz_only() const { return oz_only_; }
  // mark i of the current evaluation's set
  hipEvent_t event(int i) const { return sets_[cur_][i]; }
  // waits for the recorded evaluations and adds their stage times
  void collect() {
    for (int k = 0; k < used_; ++k) {
      MDFT_HIP_CHECK(hipEventSynchronize(sets_[k][oz_only_ ? kStageOz : kStageEnd]));
      for (int i = 0; i < kNumStages; ++i) {
        if (oz_only_ && i != kStageFftFwd) continue;  // stage FftFwd -> Oz is the OZ pass
        float ms = 0.f;
        MDFT_HIP_CHECK(hipEventElapsedTime(&ms, sets_[k][i], sets_[k][i + 1]));
        acc_[i] += ms;
      }
      ++count_;
    }
    used_ = 0;
  }
  void reset() {
    collect();
    acc_.fill(0.0);
    count_ = 0;
  }
  const std::array<double, kNumStages>& acc() const { return acc_; }
  long count() const { return count_; }

 private:
  static constexpr int kMaxSets = 512;
  bool oz_only_ = false;
  std::vector<std::array<hipEvent_t, kNumMarks>> sets_;
  int used_ = 0, cur_ = 0;
  std::array<double, kNumStages> acc_{};
  long count_ = 0;
};

inline unsigned grid_for(size_t n, unsigned block, unsigned cap = 1u << 20) {
  size_t g = (n + block - 1) / block;
  if (g > cap) g = cap;
  if (g == 0) g = 1;
  return static_cast<unsigned>(g);
}

}  // namespace mdft
