// Convolution (common/MDFT_Convolution.hpp:20-549) over the C ABI: same
// class name, constructor (Luc file path, grids, map, np_luc), in-place
// execute(deltarho_p) on (np, nx, ny, nz) complex and the accessors
// gamma_p_map(), mnmunukhi_q(), dq() (as host copies).  The optional
// Precision selects the form; the mixed forms run on complex<float>.
#pragma once

#include "OrientationProjectionTransform.hpp"

namespace MDFT {

template <typename ExecutionSpace, typename ScalarType>
class Convolution {
  using SpatialGridType = SpatialGrid<ExecutionSpace, ScalarType>;
  using AngularGridType = AngularGrid<ExecutionSpace, ScalarType>;
  using OrientationProjectionMapType = OrientationProjectionMap<ExecutionSpace, ScalarType>;

 public:
  Convolution() = delete;
  // (:106-313): reads np_luc tuples and the c(q) rows the grid needs from the
  // Luc file (common/IO/MDFT_ReadCLuc.hpp)
  Convolution(const std::string filename, const SpatialGridType& spatial_grid, const AngularGridType& angular_grid,
              const OrientationProjectionMapType& map, const int np_luc, Precision precision = Precision::Default,
              ExecutionSpace exec = ExecutionSpace())
      : m_exec(exec), m_precision(precision_code<ScalarType>(precision)) {
    Impl::check_map(map, angular_grid);
    Impl::check(mdft_conv_create(filename.c_str(), spatial_grid.m_nx, spatial_grid.m_ny, spatial_grid.m_nz,
                                 spatial_grid.m_lx, spatial_grid.m_ly, spatial_grid.m_lz, angular_grid.m_mmax,
                                 angular_grid.m_molrotsymorder, np_luc, m_precision, &m_h));
    init(spatial_grid, angular_grid);
  }
  // the same from tuples (np_luc x 5: m, n, mu, nu, khi) and values
  // (np_luc x nq_table complex) already in memory
  Convolution(const std::vector<int>& tuples, const std::vector<std::complex<double>>& c, int nq_table, double dq,
              const SpatialGridType& spatial_grid, const AngularGridType& angular_grid,
              const OrientationProjectionMapType& map, Precision precision = Precision::Default,
              ExecutionSpace exec = ExecutionSpace())
      : m_exec(exec), m_precision(precision_code<ScalarType>(precision)) {
    Impl::check_map(map, angular_grid);
    Impl::Throw_If(tuples.size() % 5 != 0 || c.size() != tuples.size() / 5 * static_cast<size_t>(nq_table),
                   "c(q) table shape does not match the tuples");
    Impl::check(mdft_conv_create_from_table(
        spatial_grid.m_nx, spatial_grid.m_ny, spatial_grid.m_nz, spatial_grid.m_lx, spatial_grid.m_ly,
        spatial_grid.m_lz, angular_grid.m_mmax, angular_grid.m_molrotsymorder, static_cast<int>(tuples.size() / 5),
        tuples.data(), reinterpret_cast<const double*>(c.data()), nq_table, dq, m_precision, &m_h));
    init(spatial_grid, angular_grid);
  }
  ~Convolution() {
    if (m_h) (void)mdft_conv_destroy(m_h);
  }
  Convolution(const Convolution&) = delete;
  Convolution& operator=(const Convolution&) = delete;

  // (L, 3) (ix, iy, iz) of the (q, -q) list in the reference's order (:126-176)
  std::vector<int> gamma_p_map() const {
    std::vector<int> out(static_cast<size_t>(m_npairs) * 3);
    Impl::check(mdft_qpairs(m_nx, m_ny, m_nz, out.data()));
    return out;
  }
  // (np_new, nq) c(q) table with the signs folded in (:210-312)
  std::vector<std::complex<double>> mnmunukhi_q() const {
    std::vector<std::complex<double>> out(static_cast<size_t>(m_np_new) * m_nq);
    Impl::check(mdft_conv_table(m_h, reinterpret_cast<double*>(out.data())));
    return out;
  }
  double dq() const { return m_dq; }
  int nq() const { return m_nq; }
  int np_new() const { return m_np_new; }

  // (:324-548), in place
  template <typename View>
  void execute(const View& deltarho_p) {
    const bool dbl = m_precision == MDFT_FP64;
    Impl::Throw_If(sizeof(typename View::value_type) != (dbl ? 2 * sizeof(double) : 2 * sizeof(float)),
                   "projection view element type does not match the precision form");
    Impl::Throw_If(deltarho_p.size() != m_np * static_cast<size_t>(m_nx) * m_ny * m_nz,
                   "projection view size must be np * nx * ny * nz");
    Impl::check(mdft_conv_execute(m_h, deltarho_p.data(), m_exec.handle()));
  }

  mdft_conv_t handle() const { return m_h; }
  int precision() const { return m_precision; }

 private:
  void init(const SpatialGridType& g, const AngularGridType& a) {
    m_nx = g.m_nx, m_ny = g.m_ny, m_nz = g.m_nz;
    m_np = static_cast<size_t>(a.m_np);
    Impl::check(mdft_conv_info(m_h, &m_nq, &m_np_new, &m_npairs, &m_dq));
  }

  ExecutionSpace m_exec;
  int m_precision;
  mdft_conv_t m_h = nullptr;
  int m_nx = 0, m_ny = 0, m_nz = 0, m_nq = 0, m_np_new = 0, m_npairs = 0;
  size_t m_np = 0;
  double m_dq = 0.0;
};

}  // namespace MDFT
