// OrientationProjectionMap and OrientationProjectionTransform
// (common/MDFT_OrientationProjectionTransform.hpp:20-155 and :161-485) over
// the C ABI: same class names, constructor arguments, methods and argument
// meaning.  angl2proj(o, p): o (N, ntheta, nphi, npsi) real LayoutRight ->
// p (np, nx, ny, nz) complex; proj2angl(p, o) the reverse.  The optional
// Precision selects the form (mdft_hip.h); with the mixed forms o is double
// and p complex<float>.
#pragma once

#include "Grid.hpp"

namespace MDFT {

template <typename ExecutionSpace, typename ScalarType>
struct OrientationProjectionMap {
  using AngularGridType = AngularGrid<ExecutionSpace, ScalarType>;

  explicit OrientationProjectionMap(const AngularGridType& angular_grid)
      : m_mmax(angular_grid.m_mmax),
        m_mrso(angular_grid.m_molrotsymorder),
        m_ntheta(angular_grid.m_ntheta),
        m_np(angular_grid.m_np) {
    const int nmu2 = m_mmax / m_mrso + 1;
    m_p.resize(static_cast<size_t>(m_mmax + 1) * (2 * m_mmax + 1) * nmu2);
    m_m.resize(m_np);
    m_mup.resize(m_np);
    m_mu.resize(m_np);
    m_mu2.resize(m_np);
    std::vector<double> d(static_cast<size_t>(m_ntheta) * m_np);
    Impl::check(mdft_angular_tables(m_mmax, m_mrso, nullptr, nullptr, nullptr, m_p.data(), m_m.data(),
                                    m_mup.data(), m_mu.data(), m_mu2.data(), d.data()));
    m_wigner_small_d.assign(d.begin(), d.end());
  }

  // (ntheta, np): d^m_{mup,mu}(theta) of every projection (:118-147)
  const std::vector<ScalarType>& wigner_small_d() const { return m_wigner_small_d; }
  // (mmax+1, 2mmax+1, mmax/mrso+1): ip of (m, mup + mmax, mu2) (:76-91)
  const std::vector<int>& p() const { return m_p; }
  const std::vector<int>& m() const { return m_m; }
  const std::vector<int>& mup() const { return m_mup; }
  const std::vector<int>& mu() const { return m_mu; }
  const std::vector<int>& mu2() const { return m_mu2; }

  int m_mmax, m_mrso, m_ntheta, m_np;

 private:
  std::vector<int> m_p, m_m, m_mup, m_mu, m_mu2;
  std::vector<ScalarType> m_wigner_small_d;
};

namespace Impl {
// the map passed to a transform or convolution must describe its angular grid
template <typename MapType, typename AngularGridType>
void check_map(const MapType& map, const AngularGridType& angular_grid) {
  Throw_If(map.m_mmax != angular_grid.m_mmax || map.m_mrso != angular_grid.m_molrotsymorder ||
               map.m_np != angular_grid.m_np,
           "the OrientationProjectionMap was built for another angular grid");
}
}  // namespace Impl

template <typename ExecutionSpace, typename ScalarType>
class OrientationProjectionTransform {
  using SpatialGridType = SpatialGrid<ExecutionSpace, ScalarType>;
  using AngularGridType = AngularGrid<ExecutionSpace, ScalarType>;
  using OrientationProjectionMapType = OrientationProjectionMap<ExecutionSpace, ScalarType>;

 public:
  OrientationProjectionTransform() = delete;
  OrientationProjectionTransform(const SpatialGridType& spatial_grid, const AngularGridType& angular_grid,
                                 const OrientationProjectionMapType& map, Precision precision = Precision::Default,
                                 ExecutionSpace exec = ExecutionSpace())
      : m_exec(exec), m_precision(precision_code<ScalarType>(precision)) {
    Impl::check_map(map, angular_grid);
    Impl::check(mdft_opt_create(spatial_grid.m_nx, spatial_grid.m_ny, spatial_grid.m_nz, angular_grid.m_mmax,
                                angular_grid.m_molrotsymorder, m_precision, &m_h));
    m_no = static_cast<size_t>(angular_grid.m_no);
    m_N = static_cast<size_t>(spatial_grid.m_nx) * spatial_grid.m_ny * spatial_grid.m_nz;
    m_np = static_cast<size_t>(angular_grid.m_np);
  }
  ~OrientationProjectionTransform() {
    if (m_h) (void)mdft_opt_destroy(m_h);
  }
  OrientationProjectionTransform(const OrientationProjectionTransform&) = delete;
  OrientationProjectionTransform& operator=(const OrientationProjectionTransform&) = delete;

  // (:267-380)
  template <typename OView, typename PView>
  void angl2proj(const OView& o, const PView& p) {
    check_views(o, p);
    Impl::check(mdft_angl2proj(m_h, o.data(), p.data(), m_exec.handle()));
  }
  // (:388-484)
  template <typename PView, typename OView>
  void proj2angl(const PView& p, const OView& o) {
    check_views(o, p);
    Impl::check(mdft_proj2angl(m_h, p.data(), o.data(), m_exec.handle()));
  }

 private:
  template <typename OView, typename PView>
  void check_views(const OView& o, const PView& p) const {
    const bool dbl_o = m_precision != MDFT_FP32 && m_precision != MDFT_MIXED_X6F, dbl_p = m_precision == MDFT_FP64;
    Impl::Throw_If(sizeof(typename OView::value_type) != (dbl_o ? sizeof(double) : sizeof(float)),
                   "orientation view element type does not match the precision form");
    Impl::Throw_If(sizeof(typename PView::value_type) != (dbl_p ? 2 * sizeof(double) : 2 * sizeof(float)),
                   "projection view element type does not match the precision form");
    Impl::Throw_If(o.size() != m_N * m_no, "orientation view size must be N * ntheta * nphi * npsi");
    Impl::Throw_If(p.size() != m_np * m_N, "projection view size must be np * nx * ny * nz");
  }

  ExecutionSpace m_exec;
  int m_precision;
  mdft_opt_t m_h = nullptr;
  size_t m_N = 0, m_no = 0, m_np = 0;
};

}  // namespace MDFT
