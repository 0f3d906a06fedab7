// Host-side FP64 tables of the excess-functional hot path.  Each builder
// names the reference code it follows (paths relative to the Mini-MDFT tree).
#include "mdft_tables.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>

namespace mdft {

// Factorials as doubles; exact through 22!, enough for every index of
// wigner_small_d with m <= 11 (common/MDFT_Wigner.hpp:20-54 tabulates them).
double gl_factorial(int n) {
  throw_if(n < 0 || n > 22, "Factorial index out of range");
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= static_cast<double>(i);
  return f;
}

// common/MDFT_Math_Utils.hpp:24-73: Newton iteration on P_n from
// cos(pi (i - 1/4) / (n + 1/2)) until the step is below DBL_EPSILON; weights
// 1 / ((1 - x^2) P'^2), i.e. normalised to sum 1.
void gauss_legendre(int n, double* x, double* w) {
  const int half = (n + 1) / 2;
  for (int i = 1; i <= half; ++i) {
    double z = std::cos(M_PI * (i - 0.25) / (n + 0.5));
    double step = 1.0, dp = 0.0;
    while (std::fabs(step) > DBL_EPSILON) {
      double p_cur = 1.0, p_prev = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p_old = p_prev;
        p_prev = p_cur;
        p_cur = ((2 * j - 1) * z * p_prev - (j - 1) * p_old) / j;
      }
      dp = n * (z * p_cur - p_prev) / (z * z - 1.0);
      step = -p_cur / dp;
      z += step;
    }
    x[i - 1] = z;
    w[i - 1] = 1.0 / ((1.0 - z * z) * dp * dp);
    x[n - i] = -z;
    w[n - i] = w[i - 1];
  }
}

// common/MDFT_Wigner.hpp:62-130.  d^m_{mu,mup}(theta): a Legendre-type
// recursion when one index vanishes, otherwise the explicit Wigner sum.
double wigner_small_d(int m, int mu, int mup, double theta) {
  if (std::abs(mu) > m || std::abs(mup) > m) return 0.0;
  if (m == 0) return 1.0;
  if (mu == 0 || mup == 0) {
    int k = (mu == 0) ? mup : mu;
    double t = (mu == 0) ? -theta : theta;
    double sgn = 1.0;
    if (k < 0) {
      sgn = std::pow(-1.0, k);
      k = -k;
    }
    const double c = std::cos(t);
    double pl_2 = 0.0, pl_1 = 1.0;  // running P_{l-2}, P_{l-1} of order k (scaled)
    for (int l = k + 1; l <= m; ++l) {
      double pl = (c * static_cast<double>(2 * l - 1) * pl_1 - static_cast<double>(l + k - 1) * pl_2) /
                  static_cast<double>(l - k);
      pl_2 = pl_1;
      pl_1 = pl;
    }
    return sgn * std::pow(-1.0, k) * std::sqrt(gl_factorial(m - k) / gl_factorial(m + k)) *
           gl_factorial(2 * k) / (std::pow(2.0, k) * gl_factorial(k)) * std::pow(std::sin(t), k) * pl_1;
  }
  const double ch = std::cos(0.5 * theta), sh = std::sin(0.5 * theta);
  double acc = 0.0;
  for (int s = std::max(0, mu - mup); s <= std::min(m + mu, m - mup); ++s) {
    acc += std::pow(-1.0, s) /
           (gl_factorial(m + mu - s) * gl_factorial(m - mup - s) * gl_factorial(s) * gl_factorial(s - mu + mup)) *
           std::pow(ch, 2 * m + mu - mup - 2 * s) * std::pow(sh, 2 * s - mu + mup);
  }
  return std::sqrt(gl_factorial(m + mu) * gl_factorial(m - mu) * gl_factorial(m + mup) * gl_factorial(m - mup)) *
         acc;
}

// common/MDFT_Grid.hpp:87-103
std::vector<double> kproj(int n, double length) {
  std::vector<double> k(n);
  const double dk = 2.0 * M_PI / length;
  const int pos = n / 2 + 1, neg = (n - 1) / 2;
  for (int i = 0; i < pos; ++i) k[i] = static_cast<double>(i) * dk;
  for (int i = 0; i < neg; ++i) k[pos + i] = static_cast<double>(i - neg) * dk;
  return k;
}

// common/MDFT_Grid.hpp:129-263 (sizes, GL theta, io-ordered weights and the
// Sum(w) = 8 pi^2 / mrso check), OrientationProjectionMap
// (common/MDFT_OrientationProjectionTransform.hpp:55-147, Wigner table with
// argument order (m, mup, mu) at :143-144) and f_m = sqrt(2m+1) (:233-236).
AngularGrid make_angular_grid(int mmax, int mrso) {
  throw_if(mmax <= 0, "The grid mmax must be greater than zero.");
  throw_if(mrso <= 0, "The grid molrotsymorder must be greater than zero.");
  throw_if(mmax > 11, "mmax above 11 is not supported (factorial range)");
  AngularGrid g;
  g.mmax = mmax;
  g.mrso = mrso;
  g.ntheta = mmax + 1;
  g.nphi = 2 * mmax + 1;
  g.npsi = 2 * (mmax / mrso) + 1;
  g.no = g.ntheta * g.nphi * g.npsi;
  g.nmu2 = mmax / mrso + 1;
  std::vector<double> x(g.ntheta);
  g.wtheta.resize(g.ntheta);
  gauss_legendre(g.ntheta, x.data(), g.wtheta.data());
  g.theta.resize(g.ntheta);
  for (int i = 0; i < g.ntheta; ++i) g.theta[i] = std::acos(x[i]);
  const double qnorm = 8.0 * M_PI * M_PI;
  const double wphi = 1.0 / static_cast<double>(g.nphi);
  const double wpsi = 1.0 / static_cast<double>(g.npsi * mrso);
  g.w.resize(g.no);
  for (int it = 0, io = 0; it < g.ntheta; ++it)
    for (int j = 0; j < g.nphi; ++j)
      for (int k = 0; k < g.npsi; ++k, ++io) g.w[io] = g.wtheta[it] * wphi * wpsi * qnorm;
  double total = 0.0;
  for (double v : g.w) total += v;
  throw_if(std::fabs(total - qnorm / mrso) > DBL_EPSILON * 100 * std::fabs(total),
           "sum(m_w) != quadrature_norm/molrotsymorder");
  // entries with |mup| > m or mrso*mu2 > m stay 0, as in the reference's
  // zero-initialised view
  g.pmap.assign(static_cast<size_t>(mmax + 1) * (2 * mmax + 1) * g.nmu2, 0);
  for (int m = 0; m <= mmax; ++m)
    for (int mup = -m; mup <= m; ++mup)
      for (int mu = 0; mu <= m; mu += mrso) {
        g.pmap[(m * (2 * mmax + 1) + mup + mmax) * g.nmu2 + mu / mrso] = static_cast<int>(g.m.size());
        g.m.push_back(m);
        g.mup.push_back(mup);
        g.mu.push_back(mu);
        g.mu2.push_back(mu / mrso);
      }
  g.np = static_cast<int>(g.m.size());
  g.wigner_d.resize(static_cast<size_t>(g.ntheta) * g.np);
  for (int ip = 0; ip < g.np; ++ip)
    for (int it = 0; it < g.ntheta; ++it)
      g.wigner_d[it * g.np + ip] = wigner_small_d(g.m[ip], g.mup[ip], g.mu[ip], g.theta[it]);
  g.fm.resize(mmax + 1);
  for (int m = 0; m <= mmax; ++m) g.fm[m] = std::sqrt(2.0 * m + 1.0);
  return g;
}

// common/MDFT_Rotation.hpp:33-84
static RotationCoeffs build_rotation_coeffs() {
  RotationCoeffs rc;
  std::memset(&rc, 0, sizeof(rc));
  const int L = RotationCoeffs::kLmax;
  double root[2 * L + 2] = {0.0};  // root[i] = sqrt(i - 1), root[0] = root[1] = 0
  for (int i = 2; i < 2 * L + 2; ++i) root[i] = std::sqrt(static_cast<double>(i - 1));
  const double r2 = std::sqrt(2.0);
  for (int l = 2; l <= L; ++l)
    for (int m = -l; m <= l; ++m) {
      for (int m1 = -l + 1; m1 < l; ++m1) {
        const double den = root[l + m1 + 1] * root[l - m1 + 1];
        rc.a[l - 2][m + L][m1 + L] = root[l + m + 1] * root[l - m + 1] / den;
        rc.b[l - 2][m + L][m1 + L] = root[l + m + 1] * root[l + m] / (r2 * den);
      }
      const double den = root[2 * l + 1] * root[2 * l];
      rc.c[l - 2][m + L] = r2 * root[l + m + 1] * root[l - m + 1] / den;
      rc.d[l - 2][m + L] = root[l + m + 1] * root[l + m] / den;
    }
  return rc;
}

const RotationCoeffs& rotation_coeffs() {
  static const RotationCoeffs rc = build_rotation_coeffs();
  return rc;
}

// common/MDFT_Rotation.hpp:103-283, host FP64 version (the device kernel in
// mdft_conv.hip runs the same recursion per q).  Entries outside |m|, |m1|
// <= l are zero.
void rotation_matrix(int mmax, const double q[3], cd* R) {
  const int W = 2 * mmax + 1;
  auto at = [&](int l, int a, int b) -> cd& { return R[(l * W + a + mmax) * W + b + mmax]; };
  std::fill(R, R + (mmax + 1) * W * W, cd(0.0, 0.0));
  if (mmax == 0) {
    R[0] = cd(1.0, 0.0);
    return;
  }
  at(0, 0, 0) = cd(1.0, 0.0);
  double ez[3], ey[3], ex[3];
  if (q[0] == 0.0 && q[1] == 0.0 && q[2] == 0.0) {
    ez[0] = 0.0; ez[1] = 0.0; ez[2] = 1.0;
  } else {
    ez[0] = q[0]; ez[1] = q[1]; ez[2] = q[2];
  }
  if (ez[0] != 0.0 || ez[1] != 0.0) {  // y' = z x z'
    ey[0] = -ez[1]; ey[1] = ez[0]; ey[2] = 0.0;
  } else {  // y' = z' x x
    ey[0] = 0.0; ey[1] = ez[2]; ey[2] = -ez[1];
  }
  auto unit = [](double* v) {
    const double s = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; ++i) v[i] /= s;
  };
  unit(ez);
  unit(ey);
  ex[0] = ey[1] * ez[2] - ey[2] * ez[1];
  ex[1] = ey[2] * ez[0] - ey[0] * ez[2];
  ex[2] = ey[0] * ez[1] - ey[1] * ez[0];
  const double h = 1.0 / std::sqrt(2.0);
  at(1, -1, -1) = cd((ey[1] + ex[0]) * 0.5, (ex[1] - ey[0]) * 0.5);
  at(1, -1, 0) = cd(ez[0] * h, ez[1] * h);
  at(1, -1, 1) = cd((ey[1] - ex[0]) * 0.5, (-ex[1] - ey[0]) * 0.5);
  at(1, 0, -1) = cd(ex[2] * h, -ey[2] * h);
  at(1, 0, 0) = cd(ez[2], 0.0);
  at(1, 0, 1) = cd(-ex[2] * h, -ey[2] * h);
  at(1, 1, -1) = cd((ey[1] - ex[0]) * 0.5, (ex[1] + ey[0]) * 0.5);
  at(1, 1, 0) = cd(-ez[0] * h, ez[1] * h);
  at(1, 1, 1) = cd((ey[1] + ex[0]) * 0.5, (ey[0] - ex[1]) * 0.5);
  if (mmax == 1) return;
  const RotationCoeffs& C = rotation_coeffs();
  const int L = RotationCoeffs::kLmax;
  auto prod = [](double s, cd u, cd v) {
    return cd(s * (u.real() * v.real() - u.imag() * v.imag()), s * (u.real() * v.imag() + u.imag() * v.real()));
  };
  auto mirror = [&](int l, int m, int m1) {
    const double s = ((m + m1) & 1) ? -1.0 : 1.0;
    const cd v = at(l, m, m1);
    at(l, -m, -m1) = cd(s * v.real(), -s * v.imag());
  };
  for (int l = 2; l <= mmax; ++l) {
    const int k = l - 1;
    for (int m = -l; m <= l; ++m) {
      for (int m1 = (m > 0 ? 1 : 0); m1 <= l - 1; ++m1) {
        if (m == -l) {
          at(l, m, m1) = prod(C.b[l - 2][L - m][L + m1], at(1, -1, 0), at(k, m + 1, m1));
        } else if (m == l) {
          at(l, m, m1) = prod(C.b[l - 2][L + m][L + m1], at(1, 1, 0), at(k, m - 1, m1));
        } else {
          const double A = C.a[l - 2][L + m][L + m1], Bp = C.b[l - 2][L + m][L + m1],
                       Bn = C.b[l - 2][L - m][L + m1];
          const cd r0 = at(1, 0, 0), rp = at(1, 1, 0), rn = at(1, -1, 0);
          const cd s0 = at(k, m, m1), sp = at(k, m + 1, m1), sn = at(k, m - 1, m1);
          const double re = A * (r0.real() * s0.real()) + Bp * (rp.real() * sn.real() - rp.imag() * sn.imag()) +
                            Bn * (rn.real() * sp.real() - rn.imag() * sp.imag());
          const double im = A * (r0.real() * s0.imag()) + Bp * (rp.real() * sn.imag() + rp.imag() * sn.real()) +
                            Bn * (rn.real() * sp.imag() + rn.imag() * sp.real());
          at(l, m, m1) = cd(re, im);
        }
        mirror(l, m, m1);
      }
      const int m1 = l;
      if (m == -l) {
        at(l, m, m1) = prod(C.d[l - 2][L - m], at(1, -1, 1), at(k, m + 1, m1 - 1));
      } else if (m == l) {
        at(l, m, m1) = prod(C.d[l - 2][L + m], at(1, 1, 1), at(k, m - 1, m1 - 1));
      } else {
        const double Cc = C.c[l - 2][L + m], Dp = C.d[l - 2][L + m], Dn = C.d[l - 2][L - m];
        const cd r0 = at(1, 0, 1), rp = at(1, 1, 1), rn = at(1, -1, 1);
        const cd s0 = at(k, m, m1 - 1), sp = at(k, m + 1, m1 - 1), sn = at(k, m - 1, m1 - 1);
        const double re = Cc * (r0.real() * s0.real() - r0.imag() * s0.imag()) +
                          Dp * (rp.real() * sn.real() - rp.imag() * sn.imag()) +
                          Dn * (rn.real() * sp.real() - rn.imag() * sp.imag());
        const double im = Cc * (r0.real() * s0.imag() + r0.imag() * s0.real()) +
                          Dp * (rp.real() * sn.imag() + rp.imag() * sn.real()) +
                          Dn * (rn.real() * sp.imag() + rn.imag() * sp.real());
        at(l, m, m1) = cd(re, im);
      }
      mirror(l, m, m1);
    }
  }
}

// common/MDFT_Convolution.hpp:126-176: iz in [0, nz/2], then iy, then ix; a
// point is skipped when it and its mirror are both already covered.
std::vector<int> qpairs_reference(int nx, int ny, int nz) {
  std::vector<unsigned char> done(static_cast<size_t>(nx) * ny * nz, 0);
  auto lin = [&](int ix, int iy, int iz) { return (static_cast<size_t>(ix) * ny + iy) * nz + iz; };
  auto neg = [](int i, int n) { return i == 0 ? 0 : n - i; };
  std::vector<int> out;
  for (int iz = 0; iz <= nz / 2; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const size_t a = lin(ix, iy, iz), b = lin(neg(ix, nx), neg(iy, ny), neg(iz, nz));
        if (done[a] && done[b]) continue;
        done[a] = done[b] = 1;
        out.insert(out.end(), {ix, iy, iz});
      }
  return out;
}

// common/IO/MDFT_ReadCLuc.hpp:189-259 (deltaAbscissa) and :123-187
// (read_luc): 17 header lines; tuples on lines 12-16; then rows "q (re,im)...".
LucTable read_luc(const std::string& path, int np_luc, double qmaxwanted) {
  std::ifstream f(path);
  throw_if(!f.good(), "File: " + path + "does not exist.");
  std::vector<std::string> lines;
  for (std::string s; std::getline(f, s);) lines.push_back(s);
  const int skip = 17;
  throw_if(static_cast<int>(lines.size()) < skip + 4, "Unexpected EOF while skipping lines in file: " + path);
  auto first_value = [&](int row) {
    std::istringstream is(lines[skip + row]);
    double v;
    throw_if(!(is >> v), "Error reading abscissa in file: " + path);
    return v;
  };
  LucTable t;
  t.np_luc = np_luc;
  t.dq = first_value(3) - first_value(2);
  throw_if(!(t.dq > 0.0), "Non-positive abscissa step in file: " + path);
  for (int r = 1; r + skip < static_cast<int>(lines.size()); ++r) {
    std::istringstream is(lines[skip + r]);
    double v;
    if (!(is >> v)) break;
    throw_if(((v - first_value(r - 1)) - t.dq) / t.dq > 1e-5, "Non-uniform abscissa detected in file: " + path);
  }
  t.nq = static_cast<int>(qmaxwanted / t.dq + 0.5) + 2;
  // tuples: label line (line 11), then labelled rows of np_luc ints
  std::istringstream hdr(lines[11] + " " + lines[12] + " " + lines[13] + " " + lines[14] + " " + lines[15]);
  t.tuples.assign(static_cast<size_t>(np_luc) * 5, 0);
  for (int j = 0; j < 5; ++j) {
    std::string label;
    throw_if(!(hdr >> label), "Missing tuple label in file: " + path);
    for (int i = 0; i < np_luc; ++i) throw_if(!(hdr >> t.tuples[i * 5 + j]), "Short tuple line in file: " + path);
  }
  t.c.assign(static_cast<size_t>(np_luc) * t.nq, cd(0.0, 0.0));
  for (int iq = 0; iq < t.nq; ++iq) {
    throw_if(skip + iq >= static_cast<int>(lines.size()), "Luc file has fewer q rows than the grid needs: " + path);
    std::istringstream is(lines[skip + iq]);
    double q;
    throw_if(!(is >> q), "Error reading q in file: " + path);
    for (int ip = 0; ip < np_luc; ++ip) {
      cd z;
      throw_if(!(is >> z), "Short c(q) row in file: " + path);
      t.c[static_cast<size_t>(ip) * t.nq + iq] = z;
    }
    throw_if(q > qmaxwanted + 2.0 * t.dq, "q > qmaxwanted in file.");
  }
  return t;
}

double qmax_of_grid(const std::vector<double>& kx, const std::vector<double>& ky, const std::vector<double>& kz,
                    int nz) {
  const double a = *std::max_element(kx.begin(), kx.end());
  const double b = *std::max_element(ky.begin(), ky.end());
  const double c = *std::max_element(kz.begin(), kz.begin() + nz / 2 + 1);
  return std::sqrt(a * a + b * b + c * c);
}

// common/MDFT_Convolution.hpp:178-312 with common/IO/MDFT_ReadCLuc.hpp:88-111:
// the Luc tuple index h_p(m, n, mu/mrso, nu/mrso, khi), the ia enumeration
// (ip, n, nu2) and the table rows in (m, khi, mu2, n, nu2) order with the OZ
// signs folded in.  Both enumerations visit (m, khi, mu2) in the order of ip
// (khi plays the role of mup), so row ia of the table is the ia-th term.
ConvTables make_conv_tables(int nx, int ny, int nz, double lx, double ly, double lz, int mmax, int mrso,
                            const LucTable& luc) {
  throw_if(nx <= 0 || ny <= 0 || nz <= 0, "The space is divided into grid nodes.");
  throw_if(lx <= 0 || ly <= 0 || lz <= 0, "The supercell cannot have negative length.");
  throw_if(nx > 1024 || ny > 1024 || nz > 1024, "grid sides above 1024 are not supported");
  throw_if(mmax > 6, "RotationCoeffs support mmax <= 6 (common/MDFT_Rotation.hpp:29)");
  ConvTables T;
  T.nx = nx; T.ny = ny; T.nz = nz; T.mmax = mmax; T.mrso = mrso;
  T.lx = lx; T.ly = ly; T.lz = lz;
  T.ang = make_angular_grid(mmax, mrso);
  T.kx = kproj(nx, lx);
  T.ky = kproj(ny, ly);
  T.kz = kproj(nz, lz);
  const double qmax = qmax_of_grid(T.kx, T.ky, T.kz, nz);
  T.dq = luc.dq;
  T.nq = static_cast<int>(qmax / luc.dq + 0.5) + 2;
  throw_if(luc.nq < T.nq, "Luc table has fewer q rows than the grid needs");
  const int h = mmax / mrso, W5 = 2 * h + 1, WK = 2 * mmax + 1;
  std::vector<int> hp(static_cast<size_t>(mmax + 1) * (mmax + 1) * W5 * W5 * WK, -1);
  auto hidx = [&](int m, int n, int mu2, int nu2, int khi) {
    return (((static_cast<size_t>(m) * (mmax + 1) + n) * W5 + mu2 + h) * W5 + nu2 + h) * WK + khi + mmax;
  };
  for (int i = 0; i < luc.np_luc; ++i) {
    const int* t = &luc.tuples[i * 5];
    throw_if(t[0] < 0 || t[0] > mmax || t[1] < 0 || t[1] > mmax || std::abs(t[4]) > mmax ||
                 std::abs(t[2] / mrso) > h || std::abs(t[3] / mrso) > h,
             "Luc tuple outside the angular grid");
    hp[hidx(t[0], t[1], t[2] / mrso, t[3] / mrso, t[4])] = i;
  }
  const AngularGrid& g = T.ang;
  T.ia_start.assign(g.np + 1, 0);
  for (int ip = 0; ip < g.np; ++ip) {
    int cnt = 0;
    for (int n = std::abs(g.mup[ip]); n <= mmax; ++n) cnt += 2 * (n / mrso) + 1;
    T.ia_start[ip + 1] = T.ia_start[ip] + cnt;
  }
  for (int m = 0; m <= mmax; ++m)
    for (int khi = -m; khi <= m; ++khi)
      for (int mu2 = 0; mu2 <= m / mrso; ++mu2)
        for (int n = std::abs(khi); n <= mmax; ++n)
          for (int nu2 = -n / mrso; nu2 <= n / mrso; ++nu2) {
            const int src = hp[hidx(m, n, mu2, nu2, khi)];
            throw_if(src < 0, "Luc file lacks a tuple the convolution needs");
            const int nu = nu2 * mrso;
            const double sgn = nu < 0 ? std::pow(-1.0, khi + nu) : std::pow(-1.0, n);
            for (int iq = 0; iq < T.nq; ++iq) T.ctab.push_back(sgn * luc.c[static_cast<size_t>(src) * luc.nq + iq]);
          }
  T.np_new = static_cast<int>(T.ctab.size() / T.nq);
  throw_if(T.np_new != T.ia_start[g.np], "np_new does not match the ia enumeration");
  return T;
}

}  // namespace mdft
