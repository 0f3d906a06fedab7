// The Solvent free functions on the hot path (common/MDFT_Solvent.hpp:
// 575-596 get_delta_rho, :613-660 get_delta_f) and the c-proj evaluation of
// Solver::energy_cproj_mrso (src/mini-mdft.hpp:137-193) over the C ABI.
#pragma once

#include "Convolution.hpp"

namespace MDFT {

namespace Impl {
template <typename T>
constexpr int dtype_code() {
  static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value, "double or float views");
  return std::is_same<T, double>::value ? MDFT_DOUBLE : MDFT_FLOAT;
}
}  // namespace Impl

// delta_rho = rho0 (xi^2 - 1), elementwise; throws if the sizes differ
template <typename ExecutionSpace, typename ViewType, typename ScalarType>
void get_delta_rho(const ExecutionSpace& exec_space, const ViewType& xi, const ViewType& delta_rho,
                   const ScalarType rho0) {
  using T = typename ViewType::non_const_value_type;
  Impl::Throw_If(delta_rho.size() != xi.size(), "size of delta_rho must be the same as the size of xi");
  Impl::check(mdft_delta_rho(xi.data(), delta_rho.data(), xi.size(), static_cast<double>(rho0),
                             Impl::dtype_code<T>(), exec_space.handle()));
}

// delta_f = 2 rho0 xi vexc and ff = sum rho0 prefactor vexc w (xi^2 - 1);
// xi, vexc, delta_f are (nx, ny, nz, no), w is (no) double.  ff is returned
// after the stream has completed (FP64, deterministic order).
template <typename ExecutionSpace, typename View1DType, typename View4DType, typename ScalarType>
void get_delta_f(const ExecutionSpace& exec_space, const View4DType& xi, const View4DType& vexc,
                 const View1DType& w, const View4DType& delta_f, ScalarType& ff, const ScalarType rho0,
                 const ScalarType prefactor) {
  using T = typename View4DType::non_const_value_type;
  static_assert(std::is_same<typename View1DType::non_const_value_type, double>::value, "w is a double view");
  for (int i = 0; i < 4; i++)
    Impl::Throw_If(xi.extent(i) != vexc.extent(i), "angular grid size of xi and vexc must be the same");
  const size_t nxyz = xi.extent(0) * xi.extent(1) * xi.extent(2);
  const int no = static_cast<int>(xi.extent(3));
  Impl::Throw_If(w.size() != static_cast<size_t>(no), "w must hold one weight per orientation");
  Impl::Throw_If(delta_f.size() != xi.size(), "size of delta_f must be the same as the size of xi");
  double f = 0.0;
  Impl::check(mdft_delta_f(xi.data(), vexc.data(), w.data(), delta_f.data(), nxyz, no, static_cast<double>(rho0),
                           static_cast<double>(prefactor), &f, Impl::dtype_code<T>(), exec_space.handle()));
  ff = static_cast<ScalarType>(f);
}

// One c-proj excess functional + gradient evaluation with delta_rho fused
// into the projection and get_delta_f into proj2angl (the stage sequence of
// Solver::energy_cproj_mrso).  xi and df are (nx, ny, nz, no): double for the
// FP64 and mixed forms, float for FP32.  The precision form is the
// Convolution's.
template <typename ExecutionSpace, typename ScalarType>
class ExcessFunctional {
 public:
  ExcessFunctional(const Convolution<ExecutionSpace, ScalarType>& conv, double rho0, double prefactor,
                   ExecutionSpace exec = ExecutionSpace())
      : m_exec(exec), m_precision(conv.precision()) {
    Impl::check(mdft_solver_create(conv.handle(), rho0, prefactor, m_precision, &m_h));
  }
  ~ExcessFunctional() {
    if (m_h) (void)mdft_solver_destroy(m_h);
  }
  ExcessFunctional(const ExcessFunctional&) = delete;
  ExcessFunctional& operator=(const ExcessFunctional&) = delete;

  // ff (kJ/mol) and df; waits for the stream
  template <typename View4DType>
  void energy_cproj_mrso(double& ff, const View4DType& xi, const View4DType& df) {
    check_views(xi, df);
    Impl::check(mdft_energy_cproj(m_h, xi.data(), df.data(), &ff, m_exec.handle()));
  }
  // stream-ordered form: ff lands in one device double
  template <typename View4DType>
  void energy_cproj_mrso_async(double* ff_device, const View4DType& xi, const View4DType& df) {
    check_views(xi, df);
    Impl::check(mdft_energy_cproj_async(m_h, xi.data(), df.data(), ff_device, m_exec.handle()));
  }
  size_t workspace_bytes() const {
    size_t b = 0;
    Impl::check(mdft_solver_workspace_bytes(m_h, &b));
    return b;
  }
  mdft_solver_t handle() const { return m_h; }

 private:
  template <typename View4DType>
  void check_views(const View4DType& xi, const View4DType& df) const {
    using T = typename View4DType::non_const_value_type;
    Impl::Throw_If(sizeof(T) != (m_precision == MDFT_FP32 ? sizeof(float) : sizeof(double)),
                   "xi / df element type does not match the precision form");
    Impl::Throw_If(xi.size() != df.size(), "xi and df must have the same size");
  }

  ExecutionSpace m_exec;
  int m_precision;
  mdft_solver_t m_h = nullptr;
};

}  // namespace MDFT
