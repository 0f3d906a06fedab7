// SpatialGrid and AngularGrid (common/MDFT_Grid.hpp:23-104 and :106-334),
// same names, constructor arguments, members and errors.  The tables are built
// by the library in FP64 on the host (mdft_kproj, mdft_angular_tables); the
// orientation weights m_w are a device view, as the reference's View1DType
// (:120, :162) that get_delta_f reads, the other tables host vectors.
#pragma once

#include <algorithm>
#include <cmath>

#include "Core.hpp"

namespace MDFT {

template <typename ExecutionSpace, typename ScalarType>
struct SpatialGrid {
  using IntType = int;
  using IntArrayType = std::array<IntType, 3>;
  using ScalarArrayType = std::array<ScalarType, 3>;

  IntArrayType m_n_nodes, m_n;
  IntType m_nx, m_ny, m_nz;
  ScalarArrayType m_length, m_l;
  ScalarType m_lx, m_ly, m_lz;
  ScalarArrayType m_dl;
  ScalarType m_dx, m_dy, m_dz;
  ScalarType m_dv;  // elemental volume
  ScalarType m_v;
  // k_i = i 2 pi / L for i <= n/2, then (i - n) 2 pi / L (:87-103)
  std::vector<ScalarType> m_kx, m_ky, m_kz;

  SpatialGrid(IntArrayType n_nodes, ScalarArrayType length) : m_n_nodes(n_nodes), m_length(length) {
    Impl::Throw_If(std::any_of(m_length.begin(), m_length.end(), [](ScalarType v) { return v <= 0; }),
                   "The supercell cannot have negative length.");
    Impl::Throw_If(std::any_of(m_n_nodes.begin(), m_n_nodes.end(), [](IntType v) { return v <= 0; }),
                   "The space is divided into grid nodes.");
    m_n = m_n_nodes;
    m_l = m_length;
    m_nx = m_n[0], m_ny = m_n[1], m_nz = m_n[2];
    m_lx = m_l[0], m_ly = m_l[1], m_lz = m_l[2];
    for (int i = 0; i < 3; ++i) m_dl[i] = m_l[i] / static_cast<ScalarType>(m_n[i]);
    m_dx = m_dl[0], m_dy = m_dl[1], m_dz = m_dl[2];
    m_dv = m_dx * m_dy * m_dz;
    m_v = m_lx * m_ly * m_lz;
    m_kx = kproj(m_nx, m_lx);
    m_ky = kproj(m_ny, m_ly);
    m_kz = kproj(m_nz, m_lz);
  }

 private:
  static std::vector<ScalarType> kproj(int n, ScalarType l) {
    std::vector<double> k(n);
    Impl::check(mdft_kproj(n, static_cast<double>(l), k.data()));
    return std::vector<ScalarType>(k.begin(), k.end());
  }
};

template <typename ExecutionSpace, typename ScalarType>
struct AngularGrid {
  using execSpace = ExecutionSpace;
  using IntType = int;

  IntType m_mmax, m_molrotsymorder, m_ntheta, m_nphi, m_npsi, m_no, m_np;
  ScalarType m_dphi, m_dpsi;
  ScalarType m_quadrature_norm = 8.0 * M_PI * M_PI;
  // theta nodes (acos of the Gauss-Legendre nodes, descending) and their
  // weights (sum 1)
  std::vector<ScalarType> m_thetaofntheta, m_wthetaofntheta;
  // the orientation weights w(io) (sum 8 pi^2 / mrso), on the device
  View<ScalarType, 1> m_w;

  AngularGrid(int mmax, int molrotsymorder) : m_mmax(mmax), m_molrotsymorder(molrotsymorder) {
    Impl::Throw_If(mmax <= 0, "The grid mmax must be greater than zero.");
    Impl::Throw_If(molrotsymorder <= 0, "The grid molrotsymorder must be greater than zero.");
    int s[6];
    Impl::check(mdft_angular_sizes(mmax, molrotsymorder, s));
    m_ntheta = s[0], m_nphi = s[1], m_npsi = s[2], m_no = s[3], m_np = s[4];
    m_dphi = static_cast<ScalarType>(2.0 * M_PI / m_nphi);
    m_dpsi = static_cast<ScalarType>(2.0 * M_PI / (m_npsi * m_molrotsymorder));
    std::vector<double> th(m_ntheta), wth(m_ntheta), w(m_no);
    Impl::check(mdft_angular_tables(mmax, molrotsymorder, th.data(), wth.data(), w.data(), nullptr, nullptr,
                                    nullptr, nullptr, nullptr, nullptr));
    m_thetaofntheta.assign(th.begin(), th.end());
    m_wthetaofntheta.assign(wth.begin(), wth.end());
    m_w = View<ScalarType, 1>("w", m_no);
    deep_copy(m_w, std::vector<ScalarType>(w.begin(), w.end()));
  }
};

}  // namespace MDFT
