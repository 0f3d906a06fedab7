// C++ facade over include/mdft_hip.h: execution space, device views,
// precision forms and error translation.
//
// The reference's classes (common/MDFT_*.hpp) are templates over a Kokkos
// execution space and take Kokkos::View arguments.  Here the execution space
// is MDFT::HIP (a HIP stream) and the views are MDFT::View: a LayoutRight
// device array with the Kokkos accessors the reference code uses (data(),
// extent(i), size(), label()).  Errors of the C ABI are rethrown as
// std::runtime_error, as the reference's Impl::Throw_If does
// (common/MDFT_Asserts.hpp:15-28).
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <complex>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../mdft_hip.h"

namespace MDFT {

namespace Impl {
inline void Throw_If(bool condition, const std::string& message) {
  if (condition) throw std::runtime_error(message);
}
// status of a C-ABI call -> exception carrying mdft_last_error()
inline void check(int status) {
  if (status != 0) throw std::runtime_error(mdft_last_error());
}
inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace Impl

// The execution space: an in-order HIP stream (the null stream by default).
// Work is stream ordered; fence() waits for it.
struct HIP {
  hipStream_t stream = nullptr;
  HIP() = default;
  explicit HIP(hipStream_t s) : stream(s) {}
  void fence() const { Impl::hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }
  void* handle() const { return static_cast<void*>(stream); }
};

// Precision forms of the transforms (mdft_hip.h): Default follows
// ScalarType (double -> FP64, float -> FP32).  The mixed forms take
// complex<float> projections and double orientation fields, except MixedX6F
// (float orientation fields).
enum class Precision {
  Default = -1,
  FP64 = MDFT_FP64,
  FP32 = MDFT_FP32,
  Mixed = MDFT_MIXED,
  MixedX6 = MDFT_MIXED_X6,
  MixedX6F = MDFT_MIXED_X6F
};

template <typename ScalarType>
constexpr int precision_code(Precision p) {
  if (p != Precision::Default) return static_cast<int>(p);
  return std::is_same<ScalarType, float>::value ? MDFT_FP32 : MDFT_FP64;
}

// LayoutRight device array of rank R (owning, or a view of external memory).
template <typename T, int R>
class View {
 public:
  using value_type = T;
  using non_const_value_type = typename std::remove_const<T>::type;
  static constexpr int rank = R;

  View() = default;
  // owning: View("label", n0, n1, ...)
  template <typename... I>
  explicit View(const std::string& label, I... n) : label_(label), extents_{static_cast<size_t>(n)...} {
    static_assert(sizeof...(I) == R, "one extent per rank");
    T* p = nullptr;
    if (size() > 0) Impl::hip_check(hipMalloc(reinterpret_cast<void**>(&p), size() * sizeof(T)), "hipMalloc");
    data_ = std::shared_ptr<T>(p, [](T* q) {
      if (q) (void)hipFree(q);
    });
  }
  // unmanaged: View(ptr, n0, n1, ...) over caller-owned device memory
  template <typename... I>
  explicit View(T* ptr, I... n) : label_("unmanaged"), extents_{static_cast<size_t>(n)...} {
    static_assert(sizeof...(I) == R, "one extent per rank");
    data_ = std::shared_ptr<T>(ptr, [](T*) {});
  }

  T* data() const { return data_.get(); }
  size_t extent(int i) const { return extents_[i]; }
  int extent_int(int i) const { return static_cast<int>(extents_[i]); }
  size_t size() const {
    size_t s = 1;
    for (size_t e : extents_) s *= e;
    return s;
  }
  const std::string& label() const { return label_; }

 private:
  std::string label_;
  std::array<size_t, R> extents_{};
  std::shared_ptr<T> data_;
};

// host <-> device copies (the reference's create_mirror_view_and_copy / deep_copy)
template <typename T, int R>
std::vector<typename std::remove_const<T>::type> create_mirror_view_and_copy(const View<T, R>& v) {
  std::vector<typename std::remove_const<T>::type> h(v.size());
  if (!h.empty()) Impl::hip_check(hipMemcpy(h.data(), v.data(), v.size() * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
  return h;
}
template <typename T, int R>
void deep_copy(const View<T, R>& dst, const std::vector<typename std::remove_const<T>::type>& src) {
  Impl::Throw_If(src.size() != dst.size(), "deep_copy: size mismatch");
  if (!src.empty()) Impl::hip_check(hipMemcpy(dst.data(), src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

}  // namespace MDFT
