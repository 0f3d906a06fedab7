// Batched 3D c2c FFT plans (see mdft_fft.hip).
#pragma once

#include <type_traits>

#include "mdft_device.hpp"

namespace mdft {

constexpr int kFftMaxStages = 16;

struct FftStages {
  int n;
  int nstages;
  int radix[kFftMaxStages];
};

// One axis: radix schedule (4s first, then 2, 3, 5, 7, 11, 13) and the FP64 /
// FP32 twiddle tables e^{-2 pi i t/n}, t < n, in device memory.
struct FftPlan1D {
  FftStages st{};
  explicit FftPlan1D(int n);
  ~FftPlan1D();
  FftPlan1D(const FftPlan1D&) = delete;
  FftPlan1D& operator=(const FftPlan1D&) = delete;
  template <typename T>
  const cplx<T>* twiddles() const;

 private:
  double2* tw64_ = nullptr;
  float2* tw32_ = nullptr;
};

// How a 3D transform runs: either the fused z/y pass over each (y, z) plane
// (plane_threads per block) then the x pass, or one pass per axis; w* are
// the lines per tile of each per-axis pass (0: the default).
struct FftConfig {
  bool plane = false;
  int plane_threads = 512;
  int wz = 0, wy = 0, wx = 0;
};

// In-place unnormalised transform over the last three axes of a
// (batch, nx, ny, nz) LayoutRight complex array: z, then y, then x.
struct Fft3dPlan {
  int nx, ny, nz;
  FftPlan1D px, py, pz;
  Fft3dPlan(int nx, int ny, int nz);
  // with the tuned configuration once autotune ran for T, else the default
  template <typename T>
  void execute(cplx<T>* data, long batch, int sign, hipStream_t stream) const;
  template <typename T>
  void execute_config(const FftConfig& c, cplx<T>* data, long batch, int sign, hipStream_t stream) const;
  // times the candidates on a scratch array (synchronises; not for graph capture)
  template <typename T>
  FftConfig autotune(long batch, hipStream_t s) const;
  template <typename T>
  FftConfig default_config() const;
  template <typename T>
  bool tuned() const { return tuned_[std::is_same<T, double>::value ? 1 : 0]; }
  bool plane_ok(size_t elem, int threads) const;

 private:
  mutable FftConfig cfg_[2];
  mutable bool tuned_[2] = {false, false};
};

}  // namespace mdft
