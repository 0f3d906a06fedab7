// Host-side (FP64) tables of the excess-functional hot path: angular grid,
// orientation <-> projection map, Wigner-d, k vectors, rotation recursion,
// (q, -q) pair list, the OZ coefficient table and the Luc c(q) reader.
#pragma once

#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

namespace mdft {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Impl::Throw_If (common/MDFT_Asserts.hpp:15-28)
inline void throw_if(bool cond, const std::string& msg) {
  if (cond) throw Error(msg);
}

using cd = std::complex<double>;

double gl_factorial(int n);
void gauss_legendre(int n, double* x, double* w);
double wigner_small_d(int m, int mu, int mup, double theta);
std::vector<double> kproj(int n, double length);

struct AngularGrid {
  int mmax = 0, mrso = 0, ntheta = 0, nphi = 0, npsi = 0, no = 0, np = 0, nmu2 = 0;
  std::vector<double> theta, wtheta, w;  // ntheta, ntheta, no
  std::vector<int> pmap;                 // (mmax+1) x (2mmax+1) x nmu2
  std::vector<int> m, mup, mu, mu2;      // np
  std::vector<double> wigner_d;          // ntheta x np
  std::vector<double> fm;                // mmax+1
  int P(int im, int imup, int imu2) const { return pmap[(im * (2 * mmax + 1) + imup + mmax) * nmu2 + imu2]; }
};
AngularGrid make_angular_grid(int mmax, int mrso);

// recursion coefficients a, b (5 x 13 x 13) and c, d (5 x 13), l = 2..6
struct RotationCoeffs {
  static constexpr int kLmax = 6, kW = 2 * kLmax + 1;
  double a[kLmax - 1][kW][kW];
  double b[kLmax - 1][kW][kW];
  double c[kLmax - 1][kW];
  double d[kLmax - 1][kW];
};
const RotationCoeffs& rotation_coeffs();
void rotation_matrix(int mmax, const double q[3], cd* R);

// (ix, iy, iz) triples in the reference's enumeration order
std::vector<int> qpairs_reference(int nx, int ny, int nz);

struct LucTable {
  int np_luc = 0, nq = 0;
  double dq = 0;
  std::vector<int> tuples;  // np_luc x 5
  std::vector<cd> c;        // np_luc x nq
};
LucTable read_luc(const std::string& path, int np_luc, double qmaxwanted);

// Convolution constructor tables (common/MDFT_Convolution.hpp:178-312)
struct ConvTables {
  int nx, ny, nz, mmax, mrso;
  double lx, ly, lz;
  AngularGrid ang;
  std::vector<double> kx, ky, kz;
  int np_new = 0, nq = 0;
  double dq = 0;
  std::vector<cd> ctab;        // np_new x nq (reference layout, signs folded)
  std::vector<int> ia_start;   // np + 1: OZ terms of output ip are ia_start[ip] .. ia_start[ip+1)
};
ConvTables make_conv_tables(int nx, int ny, int nz, double lx, double ly, double lz, int mmax, int mrso,
                            const LucTable& luc);
double qmax_of_grid(const std::vector<double>& kx, const std::vector<double>& ky,
                    const std::vector<double>& kz, int nz);

}  // namespace mdft
