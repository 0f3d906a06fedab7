// Device helpers of the element passes over (N, no) fields: the shared
// constants, the local terms of the functional, fixed-order block reductions
// and the element loop (used by mdft_functional.hip and mdft_engine.hip).
#pragma once

#include <type_traits>

#include "mdft_device.hpp"
#include "mdft_functional.hpp"

namespace mdft {
namespace elem {

constexpr int kFThreads = 256;
constexpr int kFWaves = kFThreads / 64;
constexpr int kMaxBlocks = 2048;

// Constants the element passes share.
struct Elem {
  const double* w;   // no
  const double* h0;  // no
  double rho0, kT, dv, gscale;  // gscale = 2 prefactor / N
  long n;                       // N * no
  int no, step;                 // step = (gridDim * threads) mod no
};

// the element passes read and write each field once per pass, in streams far
// larger than the caches: non-temporal accesses (measured 2.7% faster solves)
template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void st_nt(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}

// the local terms at one element, from one ln x2: accumulates
// w (x2 ln x2 - x2 + 1) into fid and w x2 V into fext, and returns the
// gradient dv w (2 rho0 x (kT ln x2 + V) + gscale df)
__device__ __forceinline__ double local_terms(double x, double v, double df, double wio, const Elem& e, double& fid,
                                              double& fext) {
  const double x2 = fmax(x * x, 1e-300);
  const double l = log(x2);
  fid += wio * (x2 * l - x2 + 1.0);
  fext += wio * x2 * v;
  return (e.dv * wio) * (2.0 * e.rho0 * x * (e.kT * l + v) + e.gscale * df);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// One partial per block and accumulator, accumulator k at part[k * nb + block];
// with LAST_MAX the last accumulator is a maximum.
template <int NA, bool LAST_MAX>
__device__ __forceinline__ void store_partials(const double (&acc)[NA], double* __restrict__ part, int nb) {
  __shared__ double red[kFWaves][NA];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    const double v = (LAST_MAX && k == NA - 1) ? wave_max(acc[k]) : wave_sum(acc[k]);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NA; k += kFThreads) {
    double v = red[0][k];
    for (int w = 1; w < kFWaves; ++w) v = (LAST_MAX && k == NA - 1) ? fmax(v, red[w][k]) : v + red[w][k];
    part[static_cast<long>(k) * nb + blockIdx.x] = v;
  }
}

// the element loop of every pass: i over [0, n) with a grid stride, io = i mod no
#define MDFT_ELEMENT_LOOP(e)                                                        \
  const long stride_ = static_cast<long>(gridDim.x) * kFThreads;                  \
  long i = static_cast<long>(blockIdx.x) * kFThreads + threadIdx.x;               \
  int io = static_cast<int>(i % (e).no);                                          \
  for (; i < (e).n; i += stride_, io = (io + (e).step >= (e).no) ? io + (e).step - (e).no : io + (e).step)


// calls f(std::integral_constant<int, k>) for k = 0 .. kLbfgsMaxPairs
template <typename F>
void dispatch_k(int k, F&& f) {
  switch (k) {
    case 0: f(std::integral_constant<int, 0>{}); break;
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    default: throw Error("L-BFGS: too many stored pairs");
  }
}
static_assert(kLbfgsMaxPairs == 8, "dispatch_k covers 0 .. 8 pairs");

}  // namespace elem
}  // namespace mdft
