// The full functional F = F_id + F_ext + F_exc, the solute's external
// potential and the FP64 L-BFGS minimiser over the C ABI.  The reference
// keeps an unused m_vext (common/MDFT_Solvent.hpp:160), builds the rotation
// tables AngularGrid::m_rot* (common/MDFT_Grid.hpp:265-333) without using them,
// and makes one excess evaluation (src/mini-mdft.hpp:129-133); these classes
// are what a caller of that Solver adds to minimise.
#pragma once

#include <array>
#include <vector>

#include "Solvent.hpp"

namespace MDFT {

// AngularGrid's rotation matrices, rot[io] row major (common/MDFT_Grid.hpp:295-318)
template <typename ExecutionSpace, typename ScalarType>
std::vector<std::array<double, 9>> rotation_matrices(const AngularGrid<ExecutionSpace, ScalarType>& ag) {
  std::vector<std::array<double, 9>> rot(ag.m_no);
  Impl::check(mdft_rotation_matrices(ag.m_mmax, ag.m_molrotsymorder, rot.data()->data()));
  return rot;
}

// V_ext (kJ/mol) of the solute on the grid into vext, (nx, ny, nz, no) doubles:
// Lennard-Jones (Lorentz-Berthelot) and qfact q q / d over site pairs at the
// minimum-image distance d, capped at vmax.  Waits for the stream.
template <typename ExecutionSpace, typename ScalarType, typename View4DType>
void compute_vext(const ExecutionSpace& exec, const SpatialGrid<ExecutionSpace, ScalarType>& sg,
                  const AngularGrid<ExecutionSpace, ScalarType>& ag, const std::vector<Site>& solute,
                  const std::vector<Site>& solvent, double vmax, const View4DType& vext) {
  static_assert(std::is_same<typename View4DType::non_const_value_type, double>::value, "vext is a double view");
  Impl::Throw_If(vext.size() != static_cast<size_t>(sg.m_nx) * sg.m_ny * sg.m_nz * ag.m_no,
                 "vext must hold one value per grid node and orientation");
  static_assert(sizeof(Site) == 6 * sizeof(double), "Site is six doubles");
  Impl::check(mdft_vext(sg.m_nx, sg.m_ny, sg.m_nz, sg.m_lx, sg.m_ly, sg.m_lz, ag.m_mmax, ag.m_molrotsymorder,
                        static_cast<int>(solute.size()), reinterpret_cast<const double*>(solute.data()),
                        static_cast<int>(solvent.size()), reinterpret_cast<const double*>(solvent.data()), vmax,
                        vext.data(), exec.handle()));
}

// xi = exp(-V / (2 kT)) from vext (both (nx, ny, nz, no) doubles): the
// minimum of F_id + F_ext alone, a start point for Functional::minimize (the
// reference starts from xi = 1, common/MDFT_Solvent.hpp:287-296, having no
// V_ext).  Waits for the stream.
template <typename ExecutionSpace, typename View4DType>
void xi_from_vext(const ExecutionSpace& exec, const View4DType& vext, double kT, const View4DType& xi) {
  static_assert(std::is_same<typename View4DType::non_const_value_type, double>::value, "xi is a double view");
  Impl::Throw_If(vext.size() != xi.size(), "xi_from_vext: vext and xi differ in size");
  Impl::check(mdft_xi_from_vext(vext.data(), static_cast<long long>(vext.size()), kT, xi.data(), exec.handle()));
}

struct MinimizeResult {
  std::array<double, 4> F;  // F, F_id, F_ext, F_exc (kJ/mol)
  int iterations, evaluations;
  int status;  // 0 converged, 1 max_iter reached, 2 no acceptable step
  double residual;  // max |h0 g|
  std::vector<double> history;  // F at the start and after every iteration
};

// F and dF/dxi over an ExcessFunctional of the FP64 or a mixed form (xi
// double); vext is borrowed and must outlive the object.
template <typename ExecutionSpace, typename ScalarType>
class Functional {
 public:
  template <typename View4DType>
  Functional(ExcessFunctional<ExecutionSpace, ScalarType>& excess, const View4DType& vext, double kT,
             ExecutionSpace exec = ExecutionSpace())
      : m_exec(exec), m_n(vext.size()) {
    static_assert(std::is_same<typename View4DType::non_const_value_type, double>::value, "vext is a double view");
    Impl::check(mdft_functional_create(excess.handle(), vext.data(), kT, &m_h));
  }
  ~Functional() {
    if (m_h) (void)mdft_functional_destroy(m_h);
  }
  Functional(const Functional&) = delete;
  Functional& operator=(const Functional&) = delete;

  // (F, F_id, F_ext, F_exc); g = dF/dxi; waits for the stream
  template <typename View4DType>
  std::array<double, 4> eval(const View4DType& xi, const View4DType& g) {
    check_view(xi);
    check_view(g);
    std::array<double, 4> F{};
    Impl::check(mdft_functional_eval(m_h, xi.data(), g.data(), F.data(), m_exec.handle()));
    return F;
  }

  // L-BFGS from xi, in place, until max |h0 g| < gtol
  template <typename View4DType>
  MinimizeResult minimize(const View4DType& xi, int m = 5, int max_iter = 200, double gtol = 1e-5) {
    check_view(xi);
    MinimizeResult r{};
    r.history.resize(static_cast<size_t>(max_iter) + 1);
    Impl::check(mdft_lbfgs_minimize(m_h, xi.data(), m, max_iter, gtol, r.F.data(), &r.iterations, &r.evaluations,
                                    &r.status, &r.residual, r.history.data(), m_exec.handle()));
    r.history.resize(static_cast<size_t>(r.iterations) + 1);
    return r;
  }

 private:
  template <typename View4DType>
  void check_view(const View4DType& v) const {
    static_assert(std::is_same<typename View4DType::non_const_value_type, double>::value,
                  "the functional's fields are double views");
    Impl::Throw_If(v.size() != m_n, "field size does not match the grid");
  }
  ExecutionSpace m_exec;
  size_t m_n;
  mdft_functional_t m_h = nullptr;
};

}  // namespace MDFT
