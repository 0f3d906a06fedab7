// Batched 3D c2c FFT plans (see mdft_fft.hip).
#pragma once

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

// In-place unnormalised transform over the last three axes of a
// (batch, nx, ny, nz) LayoutRight complex array: z, then y, then x.
struct Fft3dPlan {
  int nx, ny, nz;
  FftPlan1D px, py, pz;
  Fft3dPlan(int nx, int ny, int nz);
  template <typename T>
  void execute(cplx<T>* data, long batch, int sign, hipStream_t stream) const;
};

}  // namespace mdft
