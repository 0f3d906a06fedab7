// Tests of the C++ facade (include/mdft/MDFT.hpp), written the way the
// reference's unit tests use the same classes (unit_test/Test_Grid.cpp,
// Test_OrientationProjectionTransform.cpp, Test_Convolution.cpp,
// Test_Solvent.cpp), with the FP64 oracle (oracle/mdft_oracle.cpp, linked as
// a library) as the expected values.  Typed over float and double like the
// reference's TYPED_TESTs.
//
//   test_facade --host-only          grids, maps and error behaviour (no GPU)
//   test_facade <luc_file> <np_luc>  everything; the Luc file is an nmax = 3
//                                    (np_luc = 252) table for mmax = 3, mrso = 2
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "mdft/MDFT.hpp"

extern "C" {
int orc_angl2proj(int nx, int ny, int nz, int mmax, int mrso, const double* o, double* p);
int orc_proj2angl(int nx, int ny, int nz, int mmax, int mrso, const double* p, double* o);
void* orc_conv_create(int nx, int ny, int nz, double lx, double ly, double lz, int mmax, int mrso, int np_luc,
                      const int* lm, const int* ln, const int* lmu, const int* lnu, const int* lkhi,
                      const double* lc, int nq_file, double dq);
void orc_conv_destroy(void* h);
int orc_conv_execute(void* h, double* p);
int orc_energy_cproj(void* h, const double* xi, double rho0, double prefactor, double* df, double* ff,
                     double* stage_s);
int orc_angular_tables(int mmax, int mrso, double* theta, double* wtheta, double* w, int* pmap, int* m, int* mup,
                       int* mu, int* mu2, double* wigner_d);
}

namespace {

using execution_space = MDFT::HIP;
using cd = std::complex<double>;

int g_failed = 0, g_run = 0;
std::string g_where;

#define EXPECT_TRUE(c)                                                                  \
  do {                                                                                  \
    if (!(c)) {                                                                         \
      std::printf("    %s:%d: expected %s (%s)\n", __FILE__, __LINE__, #c, g_where.c_str()); \
      throw std::runtime_error("expectation failed");                                   \
    }                                                                                   \
  } while (0)

template <typename F>
void expect_throw(F&& f, const std::string& message) {
  try {
    f();
  } catch (const std::runtime_error& e) {
    if (std::string(e.what()).find(message) == std::string::npos) {
      std::printf("    threw \"%s\", expected \"%s\"\n", e.what(), message.c_str());
      throw std::runtime_error("wrong message");
    }
    return;
  }
  std::printf("    no exception, expected \"%s\"\n", message.c_str());
  throw std::runtime_error("no exception");
}

void run(const char* name, const std::function<void()>& body) {
  ++g_run;
  std::printf("[ RUN      ] %s\n", name);
  try {
    body();
    std::printf("[       OK ] %s\n", name);
  } catch (const std::exception& e) {
    ++g_failed;
    std::printf("[  FAILED  ] %s: %s\n", name, e.what());
  }
}

template <typename T>
double rel_err(const std::vector<T>& a, const std::vector<T>& b) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    num = std::max(num, static_cast<double>(std::abs(a[i] - b[i])));
    den = std::max(den, static_cast<double>(std::abs(b[i])));
  }
  return num / (den > 0 ? den : 1.0);
}

template <typename T>
constexpr double tol() {
  return std::is_same<T, double>::value ? 1e-12 : 2e-5;
}

// ---- grids and maps (host) -------------------------------------------------
template <typename T>
void test_spatial_grid() {
  using SG = MDFT::SpatialGrid<execution_space, T>;
  SG grid({8, 6, 5}, {10, 9, 7.5});
  EXPECT_TRUE(grid.m_nx == 8 && grid.m_ny == 6 && grid.m_nz == 5);
  EXPECT_TRUE(std::abs(grid.m_dv - T(10.0 / 8 * 9.0 / 6 * 7.5 / 5)) <= 1e-6 * grid.m_dv);
  // kproj: i 2 pi / L up to n/2, then (i - n) 2 pi / L (common/MDFT_Grid.hpp:87-103)
  for (int i = 0; i < 8; ++i) {
    const double k = (i <= 4 ? i : i - 8) * 2.0 * M_PI / 10.0;
    EXPECT_TRUE(std::abs(grid.m_kx[i] - T(k)) <= 1e-6);
  }
}

// AngularGrid (its weights are a device view) and OrientationProjectionMap
template <typename T>
void test_angular_grid() {
  using AG = MDFT::AngularGrid<execution_space, T>;
  AG ag(5, 2);
  EXPECT_TRUE(ag.m_ntheta == 6 && ag.m_nphi == 11 && ag.m_npsi == 5 && ag.m_no == 330 && ag.m_np == 88);
  double sw = 0.0;
  for (T v : MDFT::create_mirror_view_and_copy(ag.m_w)) sw += v;
  EXPECT_TRUE(std::abs(sw - 8.0 * M_PI * M_PI / 2.0) <= 1e-5 * sw);
  MDFT::OrientationProjectionMap<execution_space, T> map(ag);
  EXPECT_TRUE(static_cast<int>(map.m().size()) == 88 && map.wigner_small_d().size() == 6u * 88u);
  // p(m, mup + mmax, mu2) = ip, enumerated m, mup, mu2 (:76-91)
  int ip = 0;
  for (int m = 0; m <= 5; ++m)
    for (int mup = -m; mup <= m; ++mup)
      for (int mu2 = 0; mu2 <= m / 2; ++mu2, ++ip) {
        EXPECT_TRUE(map.p()[(m * 11 + mup + 5) * 3 + mu2] == ip);
        EXPECT_TRUE(map.m()[ip] == m && map.mup()[ip] == mup && map.mu2()[ip] == mu2 && map.mu()[ip] == 2 * mu2);
      }
  // Wigner table against the oracle's
  std::vector<double> d(6 * 88);
  orc_angular_tables(5, 2, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, d.data());
  for (size_t i = 0; i < d.size(); ++i) EXPECT_TRUE(std::abs(map.wigner_small_d()[i] - T(d[i])) <= 1e-6);
}

template <typename T>
void test_errors() {
  using SG = MDFT::SpatialGrid<execution_space, T>;
  using AG = MDFT::AngularGrid<execution_space, T>;
  expect_throw([] { SG g({8, 8, 8}, {10, -1, 10}); }, "The supercell cannot have negative length.");
  expect_throw([] { SG g({8, 0, 8}, {10, 10, 10}); }, "The space is divided into grid nodes.");
  expect_throw([] { AG a(0, 2); }, "The grid mmax must be greater than zero.");
  expect_throw([] { AG a(3, 0); }, "The grid molrotsymorder must be greater than zero.");
}

// ---- transforms (device) ---------------------------------------------------
template <typename T>
void test_angl2proj_proj2angl(int nx, int ny, int nz, int mmax, int mrso, MDFT::Precision prec) {
  using SG = MDFT::SpatialGrid<execution_space, T>;
  using AG = MDFT::AngularGrid<execution_space, T>;
  using OT = typename std::conditional<std::is_same<T, float>::value, float, double>::type;
  SG grid({nx, ny, nz}, {1, 1, 1});
  AG ag(mmax, mrso);
  MDFT::OrientationProjectionMap<execution_space, T> map(ag);
  MDFT::OrientationProjectionTransform<execution_space, T> opt(grid, ag, map, prec);
  const int N = nx * ny * nz;
  std::mt19937_64 gen(3);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<double> o(static_cast<size_t>(N) * ag.m_no);
  for (double& v : o) v = U(gen);
  std::vector<cd> pref(static_cast<size_t>(ag.m_np) * N);
  orc_angl2proj(nx, ny, nz, mmax, mrso, o.data(), reinterpret_cast<double*>(pref.data()));
  const bool mixed = prec == MDFT::Precision::Mixed || prec == MDFT::Precision::MixedX6;
  const double tl = mixed ? 5e-5 : tol<T>();
  auto run_one = [&](auto ptag) {
    using P = decltype(ptag);
    MDFT::View<OT, 4> od("o", N, ag.m_ntheta, ag.m_nphi, ag.m_npsi);
    MDFT::deep_copy(od, std::vector<OT>(o.begin(), o.end()));
    MDFT::View<P, 4> pd("p", ag.m_np, nx, ny, nz);
    opt.angl2proj(od, pd);
    auto ph = MDFT::create_mirror_view_and_copy(pd);
    std::vector<cd> got(ph.begin(), ph.end());
    g_where = "angl2proj";
    EXPECT_TRUE(rel_err(got, pref) < tl);
    // proj2angl of the oracle's projections
    std::vector<double> oref(o.size());
    orc_proj2angl(nx, ny, nz, mmax, mrso, reinterpret_cast<const double*>(pref.data()), oref.data());
    std::vector<P> pin(pref.begin(), pref.end());
    MDFT::deep_copy(pd, pin);
    MDFT::View<OT, 4> back("o_inv", N, ag.m_ntheta, ag.m_nphi, ag.m_npsi);
    opt.proj2angl(pd, back);
    auto bh = MDFT::create_mirror_view_and_copy(back);
    g_where = "proj2angl";
    EXPECT_TRUE(rel_err(std::vector<double>(bh.begin(), bh.end()), oref) < tl);
  };
  if (std::is_same<T, float>::value || mixed)
    run_one(std::complex<float>());
  else
    run_one(std::complex<double>());
}

// Test_OrientationProjectionTransform.cpp:313-348: a constant field goes
// through angl2proj and proj2angl unchanged
template <typename T>
void test_identity() {
  using OT = T;
  MDFT::SpatialGrid<execution_space, T> grid({5, 5, 5}, {1, 1, 1});
  MDFT::AngularGrid<execution_space, T> ag(3, 2);
  MDFT::OrientationProjectionMap<execution_space, T> map(ag);
  MDFT::OrientationProjectionTransform<execution_space, T> opt(grid, ag, map);
  const int N = 125;
  MDFT::View<OT, 4> o("o", N, ag.m_ntheta, ag.m_nphi, ag.m_npsi), o_inv("o_inv", N, ag.m_ntheta, ag.m_nphi, ag.m_npsi);
  MDFT::View<std::complex<T>, 4> p("p", ag.m_np, 5, 5, 5);
  MDFT::deep_copy(o, std::vector<OT>(o.size(), OT(1.0)));
  opt.angl2proj(o, p);
  opt.proj2angl(p, o_inv);
  auto h = MDFT::create_mirror_view_and_copy(o_inv);
  for (OT v : h) EXPECT_TRUE(std::abs(v - OT(1.0)) < 10 * tol<T>());
}

struct LucFile {
  std::vector<int> tuples;  // np_luc x 5
  std::vector<cd> c;        // np_luc x nq
  int nq = 0;
  double dq = 0.0;
};

LucFile read_luc(const std::string& path, int np_luc, double qmax) {
  LucFile f;
  MDFT::Impl::check(mdft_luc_read(path.c_str(), np_luc, qmax, &f.nq, &f.dq, nullptr, nullptr));
  f.tuples.resize(static_cast<size_t>(np_luc) * 5);
  f.c.resize(static_cast<size_t>(np_luc) * f.nq);
  MDFT::Impl::check(
      mdft_luc_read(path.c_str(), np_luc, qmax, &f.nq, &f.dq, f.tuples.data(), reinterpret_cast<double*>(f.c.data())));
  return f;
}

// Test_Convolution.cpp:81-155: execute against the oracle on the same c(q)
template <typename T>
void test_convolution(const std::string& path, int np_luc, int n, double L, MDFT::Precision prec) {
  MDFT::SpatialGrid<execution_space, T> grid({n, n, n}, {T(L), T(L), T(L)});
  MDFT::AngularGrid<execution_space, T> ag(3, 2);
  MDFT::OrientationProjectionMap<execution_space, T> map(ag);
  MDFT::Convolution<execution_space, T> conv(path, grid, ag, map, np_luc, prec);
  const double qmax = std::sqrt(3.0) * M_PI * n / L;  // |k| at the grid's corner
  LucFile f = read_luc(path, np_luc, qmax);
  std::vector<int> cols[5];
  for (int j = 0; j < 5; ++j)
    for (int i = 0; i < np_luc; ++i) cols[j].push_back(f.tuples[i * 5 + j]);
  void* oc = orc_conv_create(n, n, n, L, L, L, 3, 2, np_luc, cols[0].data(), cols[1].data(), cols[2].data(),
                             cols[3].data(), cols[4].data(), reinterpret_cast<const double*>(f.c.data()), f.nq, f.dq);
  EXPECT_TRUE(oc != nullptr);
  EXPECT_TRUE(std::abs(conv.dq() - f.dq) <= 1e-15 * f.dq);
  const size_t total = static_cast<size_t>(ag.m_np) * n * n * n;
  std::mt19937_64 gen(4);
  std::normal_distribution<double> G;
  std::vector<cd> p(total);
  for (cd& v : p) v = cd(G(gen), G(gen));
  std::vector<cd> ref = p;
  orc_conv_execute(oc, reinterpret_cast<double*>(ref.data()));
  orc_conv_destroy(oc);
  const bool dbl = std::is_same<T, double>::value && (prec == MDFT::Precision::Default || prec == MDFT::Precision::FP64);
  auto run_one = [&](auto ptag) {
    using P = decltype(ptag);
    MDFT::View<P, 4> pd("deltarho_p", ag.m_np, n, n, n);
    MDFT::deep_copy(pd, std::vector<P>(p.begin(), p.end()));
    conv.execute(pd);
    auto h = MDFT::create_mirror_view_and_copy(pd);
    g_where = "Convolution::execute";
    EXPECT_TRUE(rel_err(std::vector<cd>(h.begin(), h.end()), ref) < (dbl ? 1e-11 : 2e-5));
  };
  if (dbl)
    run_one(std::complex<double>());
  else
    run_one(std::complex<float>());
  // the pair list and table accessors
  EXPECT_TRUE(conv.gamma_p_map().size() % 3 == 0 && !conv.mnmunukhi_q().empty());
}

// Test_Solvent.cpp:129-211: closed forms of get_delta_rho / get_delta_f
template <typename T>
void test_solvent_functions() {
  const int nx = 2, ny = 3, nz = 4, no = 5;
  const size_t n = static_cast<size_t>(nx) * ny * nz * no;
  std::mt19937_64 gen(12345);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<T> xi(n), vexc(n);
  std::vector<double> w(no);
  for (auto& v : xi) v = T(U(gen));
  for (auto& v : vexc) v = T(U(gen));
  for (auto& v : w) v = U(gen);
  const T rho0 = T(0.3), pref = T(0.334);
  execution_space exec;
  MDFT::View<T, 4> xd("xi", nx, ny, nz, no), vd("vexc", nx, ny, nz, no), dr("delta_rho", nx, ny, nz, no),
      dfd("delta_f", nx, ny, nz, no);
  MDFT::View<double, 1> wd("w", no);
  MDFT::deep_copy(xd, xi);
  MDFT::deep_copy(vd, vexc);
  MDFT::deep_copy(wd, w);
  MDFT::get_delta_rho(exec, xd, dr, rho0);
  T ff = 0;
  MDFT::get_delta_f(exec, xd, vd, wd, dfd, ff, rho0, pref);
  auto drh = MDFT::create_mirror_view_and_copy(dr);
  auto dfh = MDFT::create_mirror_view_and_copy(dfd);
  double ffr = 0.0;
  for (size_t i = 0; i < n; ++i) {
    EXPECT_TRUE(std::abs(drh[i] - rho0 * (xi[i] * xi[i] - T(1))) <= 10 * tol<T>());
    EXPECT_TRUE(std::abs(dfh[i] - T(2) * rho0 * xi[i] * vexc[i]) <= 10 * tol<T>());
    ffr += double(rho0) * (double(xi[i]) * xi[i] - 1.0) * double(pref) * w[i % no] * double(vexc[i]);
  }
  EXPECT_TRUE(std::abs(double(ff) - ffr) <= 100 * tol<T>() * std::abs(ffr));
  MDFT::View<T, 4> wrong("wrong", nx, ny, nz, no + 1);
  expect_throw([&] { MDFT::get_delta_rho(exec, xd, wrong, rho0); },
               "size of delta_rho must be the same as the size of xi");
}

// src/mini-mdft.hpp:137-193 through ExcessFunctional: FP64 to 1e-10 of the
// oracle, the mixed form to its stated accuracy
void test_excess_functional(const std::string& path, int np_luc) {
  const int n = 8;
  const double L = 10.0;
  MDFT::SpatialGrid<execution_space, double> grid({n, n, n}, {L, L, L});
  MDFT::AngularGrid<execution_space, double> ag(3, 2);
  MDFT::OrientationProjectionMap<execution_space, double> map(ag);
  const double n0 = 0.0332891, T = 298.15, kT = 1.380649e-23 * 6.02214076e23 * T * 1e-3;
  const double rho0 = n0 * 2 / (8 * M_PI * M_PI), pref = -2 * kT * 4 * M_PI * M_PI / 2 / n0;
  const size_t total = static_cast<size_t>(n) * n * n * ag.m_no;
  std::mt19937_64 gen(5);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::vector<double> xi(total);
  for (double& v : xi) v = 1.0 + 0.1 * U(gen);
  const double qmax = std::sqrt(3.0) * M_PI * n / L;
  LucFile f = read_luc(path, np_luc, qmax);
  std::vector<int> cols[5];
  for (int j = 0; j < 5; ++j)
    for (int i = 0; i < np_luc; ++i) cols[j].push_back(f.tuples[i * 5 + j]);
  void* oc = orc_conv_create(n, n, n, L, L, L, 3, 2, np_luc, cols[0].data(), cols[1].data(), cols[2].data(),
                             cols[3].data(), cols[4].data(), reinterpret_cast<const double*>(f.c.data()), f.nq, f.dq);
  std::vector<double> df_ref(total);
  double ff_ref = 0.0;
  orc_energy_cproj(oc, xi.data(), rho0, pref, df_ref.data(), &ff_ref, nullptr);
  orc_conv_destroy(oc);
  for (MDFT::Precision prec : {MDFT::Precision::FP64, MDFT::Precision::MixedX6}) {
    MDFT::Convolution<execution_space, double> conv(path, grid, ag, map, np_luc, prec);
    MDFT::ExcessFunctional<execution_space, double> fun(conv, rho0, pref);
    MDFT::View<double, 4> xd("xi", n, n, n, ag.m_no), dfd("df", n, n, n, ag.m_no);
    MDFT::deep_copy(xd, xi);
    double ff = 0.0;
    fun.energy_cproj_mrso(ff, xd, dfd);
    auto dfh = MDFT::create_mirror_view_and_copy(dfd);
    const bool f64 = prec == MDFT::Precision::FP64;
    g_where = f64 ? "FP64" : "MDFT_MIXED_X6";
    std::printf("    %s: F_exc = %.12g kJ/mol (oracle %.12g), |dF| = %.3g\n", g_where.c_str(), ff, ff_ref,
                std::abs(ff - ff_ref));
    // FP64: 1e-10 relative; the mixed form's bar is absolute (|dF| <= 0.01
    // kJ/mol), checked here 100x tighter
    if (f64)
      EXPECT_TRUE(std::abs(ff - ff_ref) <= 1e-10 * std::abs(ff_ref));
    else
      EXPECT_TRUE(std::abs(ff - ff_ref) <= 1e-4);
    EXPECT_TRUE(rel_err(dfh, df_ref) < (f64 ? 1e-10 : 1e-5));
  }
}


// Solver::energy_cproj_mrso (src/mini-mdft.hpp:137-193) written against the
// facade as a caller of the reference writes it: Solvents holds xi, rho0 and
// the prefactor, the stages run one by one through the reference's views
// (delta_rho and vexc (N, ntheta, nphi, npsi), delta_rho_p (np, nx, ny, nz),
// vexc re-viewed as (nx, ny, nz, no)), get_delta_f reads the angular grid's
// device weights.  FP64 to 1e-10 of the oracle; the float weight view and the
// map / grid mismatch checks.
void test_solvent_dropin(const std::string& path, int np_luc) {
  using View4DType = MDFT::View<double, 4>;
  using ComplexView4DType = MDFT::View<std::complex<double>, 4>;
  const int n = 8;
  const double L = 10.0;
  execution_space exec;
  MDFT::SpatialGrid<execution_space, double> grid({n, n, n}, {L, L, L});
  MDFT::AngularGrid<execution_space, double> ag(3, 2);
  MDFT::OrientationProjectionMap<execution_space, double> map(ag);
  MDFT::Thermo<double> thermo(298.15);
  MDFT::Solvents<execution_space, double> solvents(grid, ag, "spce", thermo);
  auto& solvent = solvents.m_solvents.at(0);
  const int no = ag.m_no, np = ag.m_np;
  const size_t total = static_cast<size_t>(n) * n * n * no;
  // Solvents' values (common/MDFT_Solvent.hpp:417-440, :366-368)
  const double kT = 1.38064852e-23 * 6.02214129e23 * 298.15 * 1e-3;
  EXPECT_TRUE(std::abs(thermo.m_kbt - kT) <= 1e-15 * kT && solvent.m_npluc[3] == np_luc);
  EXPECT_TRUE(std::abs(solvent.m_rho0 - 0.0332891 / (4 * M_PI * M_PI)) <= 1e-15 * solvent.m_rho0);
  EXPECT_TRUE(std::abs(solvents.m_prefactor + 2 * kT * 2 * M_PI * M_PI / 0.0332891) <= 1e-12 * kT / 0.0332891);
  EXPECT_TRUE(solvent.m_xi.extent(0) == static_cast<size_t>(n) && solvent.m_xi.extent(3) == static_cast<size_t>(no));
  for (double v : MDFT::create_mirror_view_and_copy(solvent.m_xi)) EXPECT_TRUE(v == 1.0);
  EXPECT_TRUE(solvent.m_site.size() == 3 && solvent.m_site[1].q == 0.4238 && solvent.m_site[2].x == -0.816495);
  std::mt19937_64 gen(9);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::vector<double> xi(total);
  for (double& v : xi) v = 1.0 + 0.1 * U(gen);
  MDFT::deep_copy(solvent.m_xi, xi);

  MDFT::OrientationProjectionTransform<execution_space, double> op_transform(grid, ag, map);
  MDFT::Convolution<execution_space, double> conv(path, grid, ag, map, solvent.m_npluc[ag.m_mmax]);
  View4DType delta_rho("delta_rho", n * n * n, ag.m_ntheta, ag.m_nphi, ag.m_npsi);
  View4DType vexc_o("vexc", n * n * n, ag.m_ntheta, ag.m_nphi, ag.m_npsi);
  View4DType df("df", n, n, n, no);
  ComplexView4DType delta_rho_p("delta_rho_p", np, n, n, n);
  MDFT::get_delta_rho(exec, solvent.m_xi, delta_rho, solvent.m_rho0);
  op_transform.angl2proj(delta_rho, delta_rho_p);
  conv.execute(delta_rho_p);
  op_transform.proj2angl(delta_rho_p, vexc_o);
  View4DType vexc(vexc_o.data(), n, n, n, no);
  double ff = 0.0;
  MDFT::get_delta_f(exec, solvent.m_xi, vexc, ag.m_w, df, ff, solvent.m_rho0, solvents.m_prefactor);
  exec.fence();

  const double qmax = std::sqrt(3.0) * M_PI * n / L;
  LucFile f = read_luc(path, np_luc, qmax);
  std::vector<int> cols[5];
  for (int j = 0; j < 5; ++j)
    for (int i = 0; i < np_luc; ++i) cols[j].push_back(f.tuples[i * 5 + j]);
  void* oc = orc_conv_create(n, n, n, L, L, L, 3, 2, np_luc, cols[0].data(), cols[1].data(), cols[2].data(),
                             cols[3].data(), cols[4].data(), reinterpret_cast<const double*>(f.c.data()), f.nq, f.dq);
  std::vector<double> df_ref(total);
  double ff_ref = 0.0;
  orc_energy_cproj(oc, xi.data(), solvent.m_rho0, solvents.m_prefactor, df_ref.data(), &ff_ref, nullptr);
  orc_conv_destroy(oc);
  std::printf("    F_exc = %.12g kJ/mol (oracle %.12g)\n", ff, ff_ref);
  EXPECT_TRUE(std::abs(ff - ff_ref) <= 1e-10 * std::abs(ff_ref));
  EXPECT_TRUE(rel_err(MDFT::create_mirror_view_and_copy(df), df_ref) < 1e-10);

  // a float weight view: w rounded to float, the sum in FP64
  MDFT::AngularGrid<execution_space, float> agf(3, 2);
  double ff32 = 0.0;
  MDFT::get_delta_f(exec, solvent.m_xi, vexc, agf.m_w, df, ff32, solvent.m_rho0, solvents.m_prefactor);
  EXPECT_TRUE(std::abs(ff32 - ff) <= 1e-6 * std::abs(ff));

  // the reference's error paths
  MDFT::AngularGrid<execution_space, double> ag5(5, 2), ag6(6, 2);
  MDFT::OrientationProjectionMap<execution_space, double> map5(ag5);
  expect_throw([&] { MDFT::OrientationProjectionTransform<execution_space, double> t(grid, ag, map5); },
               "built for another angular grid");
  expect_throw([&] { MDFT::Convolution<execution_space, double> c(path, grid, ag, map5, np_luc); },
               "built for another angular grid");
  expect_throw([&] { MDFT::Solvents<execution_space, double> s(grid, ag6, "tip3p", thermo); },
               "solvent tip3p only avail with mmax between 0 and 5");
  expect_throw([&] { MDFT::Solvents<execution_space, double> s(grid, ag, "argon", thermo); },
               "is not implemented yet");
  MDFT::Solvents<execution_space, double> half(grid, ag, "tip3p-m", thermo, 0.5);
  EXPECT_TRUE(half.m_solvents.at(0).m_name == "tip3p" && half.m_solvents.at(0).m_n0 == 0.03349459 * 0.5 &&
              half.m_solvents.at(0).m_site[0].q == -0.95);
}

// The full functional through the facade: V_ext of one ion seen by a one-site
// solvent against the closed form, F and g against host sums over the oracle's
// excess evaluation, and a few L-BFGS iterations lowering F.
void test_functional(const std::string& path, int np_luc) {
  const int n = 8;
  const double L = 10.0;
  MDFT::SpatialGrid<execution_space, double> grid({n, n, n}, {L, L, L});
  MDFT::AngularGrid<execution_space, double> ag(3, 2);
  MDFT::OrientationProjectionMap<execution_space, double> map(ag);
  execution_space exec;
  const double n0 = 0.0332891, T = 298.15, kT = 1.38064852e-23 * 6.02214129e23 * T * 1e-3;
  const double rho0 = n0 * 2 / (8 * M_PI * M_PI), pref = -2 * kT * 4 * M_PI * M_PI / 2 / n0;
  const double qfact = 1.602176565e-19 * 1.602176565e-19 * 1.0e-3 * 6.02214129e23 /
                       (4.0 * M_PI * 8.854187817e-12 * 1.0e-10);
  const int no = ag.m_no;
  const size_t N = static_cast<size_t>(n) * n * n, total = N * no;
  // one ion at (4.1, 5.3, 6.2) and a one-site solvent
  const std::vector<MDFT::Site> ion{{0.3, 2.5, 0.5, 4.1, 5.3, 6.2}};
  const std::vector<MDFT::Site> site{{-0.8, 3.15, 0.64, 0.0, 0.0, 0.0}};
  MDFT::View<double, 4> vd("vext", n, n, n, no);
  const double vmax = 50.0;
  MDFT::compute_vext(exec, grid, ag, ion, site, vmax, vd);
  auto vh = MDFT::create_mirror_view_and_copy(vd);
  const double sig = 0.5 * (2.5 + 3.15), eps4 = 4.0 * std::sqrt(0.5 * 0.64);
  double verr = 0.0;
  for (size_t r = 0; r < N; ++r) {
    const double p[3] = {double(r / (n * n)) * L / n, double((r / n) % n) * L / n, double(r % n) * L / n};
    const double c[3] = {4.1, 5.3, 6.2};
    double d2 = 0.0;
    for (int a = 0; a < 3; ++a) {
      double d = p[a] - c[a];
      d -= L * std::nearbyint(d / L);
      d2 += d * d;
    }
    const double d = std::sqrt(d2), sr6 = std::pow(sig / d, 6);
    const double ref = std::min(eps4 * (sr6 * sr6 - sr6) + qfact * 0.3 * -0.8 / d, vmax);
    for (int io = 0; io < no; ++io) verr = std::max(verr, std::abs(vh[r * no + io] - ref));
  }
  std::printf("    V_ext max |err| = %.3g kJ/mol\n", verr);
  EXPECT_TRUE(verr <= 1e-11);
  auto rot = MDFT::rotation_matrices(ag);
  EXPECT_TRUE(rot.size() == static_cast<size_t>(no) && std::abs(rot[0][8] - std::cos(ag.m_thetaofntheta[0])) < 1e-15);

  std::mt19937_64 gen(9);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::vector<double> xi(total);
  for (double& v : xi) v = 1.0 + 0.1 * U(gen);
  const double qmax = std::sqrt(3.0) * M_PI * n / L;
  LucFile f = read_luc(path, np_luc, qmax);
  std::vector<int> cols[5];
  for (int j = 0; j < 5; ++j)
    for (int i = 0; i < np_luc; ++i) cols[j].push_back(f.tuples[i * 5 + j]);
  void* oc = orc_conv_create(n, n, n, L, L, L, 3, 2, np_luc, cols[0].data(), cols[1].data(), cols[2].data(),
                             cols[3].data(), cols[4].data(), reinterpret_cast<const double*>(f.c.data()), f.nq, f.dq);
  std::vector<double> df_ref(total);
  double ff_ref = 0.0;
  orc_energy_cproj(oc, xi.data(), rho0, pref, df_ref.data(), &ff_ref, nullptr);
  orc_conv_destroy(oc);
  const double dv = L * L * L / N;
  double fid = 0.0, fext = 0.0;
  std::vector<double> g_ref(total);
  const auto wh = MDFT::create_mirror_view_and_copy(ag.m_w);
  for (size_t i = 0; i < total; ++i) {
    const double w = wh[i % no], x2 = xi[i] * xi[i];
    fid += w * (x2 * std::log(x2) - x2 + 1.0);
    fext += w * x2 * vh[i];
    g_ref[i] = dv * w * (2 * rho0 * xi[i] * (kT * std::log(x2) + vh[i]) + 2 * pref / N * df_ref[i]);
  }
  const double F_ref[4] = {kT * rho0 * dv * fid + rho0 * dv * fext + dv / N * ff_ref, kT * rho0 * dv * fid,
                           rho0 * dv * fext, dv / N * ff_ref};
  MDFT::Convolution<execution_space, double> conv(path, grid, ag, map, np_luc);
  MDFT::ExcessFunctional<execution_space, double> excess(conv, rho0, pref);
  MDFT::Functional<execution_space, double> fun(excess, vd, kT);
  MDFT::View<double, 4> xd("xi", n, n, n, no), gd("g", n, n, n, no);
  MDFT::deep_copy(xd, xi);
  const auto F = fun.eval(xd, gd);
  auto gh = MDFT::create_mirror_view_and_copy(gd);
  g_where = "Functional::eval";
  std::printf("    F = %.12g (F_id %.6g, F_ext %.6g, F_exc %.6g) kJ/mol, host sums %.12g\n", F[0], F[1], F[2], F[3],
              F_ref[0]);
  for (int k = 0; k < 4; ++k) EXPECT_TRUE(std::abs(F[k] - F_ref[k]) <= 1e-10 * std::abs(F_ref[1] + F_ref[2]));
  EXPECT_TRUE(rel_err(gh, g_ref) < 1e-10);
  // a few iterations from xi = 1 lower F
  std::vector<double> ones(total, 1.0);
  MDFT::deep_copy(xd, ones);
  const auto r = fun.minimize(xd, 5, 5, 1e-5);
  std::printf("    minimize: %d iterations, %d evaluations, status %d, F %.8g -> %.8g\n", r.iterations,
              r.evaluations, r.status, r.history.front(), r.F[0]);
  EXPECT_TRUE(r.iterations == 5 && r.status == 1 && r.F[0] < r.history.front());
  expect_throw([&] { fun.minimize(xd, 9, 5, 1e-5); }, "m must be in 1 .. 8");
}


// The server through the facade: the first solute alone and in a batch of
// three give the same F, iteration count and final xi bit for bit.
void test_server(const std::string& path, int np_luc) {
  const int n = 8;
  const double L = 10.0;
  MDFT::SpatialGrid<execution_space, double> grid({n, n, n}, {L, L, L});
  MDFT::AngularGrid<execution_space, double> ag(3, 2);
  const double T = 298.15, n0 = 0.0332891, kT = 1.38064852e-23 * 6.02214129e23 * T * 1e-3;
  const double qmax = std::sqrt(3.0) * M_PI * n / L;
  LucFile f = read_luc(path, np_luc, qmax);
  // the exchange-symmetric definite part of the file's table: (m, m, 0, 0, 0) with a negative value
  std::vector<std::complex<double>> c(f.c.size(), 0.0);
  for (int i = 0; i < np_luc; ++i) {
    const int* t = &f.tuples[i * 5];
    if (t[0] == t[1] && t[2] == 0 && t[3] == 0 && t[4] == 0)
      for (int q = 0; q < f.nq; ++q) c[i * f.nq + q] = -std::abs(f.c[i * f.nq + q].real()) / (1.0 + t[0]);
  }
  const std::vector<MDFT::Site> water{{-0.834, 3.15061, 0.636386, 0, 0, 0},
                                      {0.417, 0, 0, 0.756950, 0, 0.585882},
                                      {0.417, 0, 0, -0.756950, 0, 0.585882}};
  const std::vector<std::vector<MDFT::Site>> sol{{{0.05, 3.0, 0.4, 5.1, 4.9, 5.2}, {-0.05, 2.8, 0.3, 4.0, 5.5, 5.0}},
                                                 {{0.08, 3.3, 0.5, 5.0, 5.0, 5.0}},
                                                 {{-0.06, 2.6, 0.2, 4.5, 5.5, 4.8}}};
  const size_t total = static_cast<size_t>(n) * n * n * ag.m_no;
  MDFT::View<double, 4> alone("alone", n, n, n, ag.m_no);
  std::vector<MDFT::View<double, 4>> mix;
  for (int k = 0; k < 3; ++k) mix.emplace_back("mix", n, n, n, ag.m_no);
  MDFT::SolveResult ra{};
  {
    MDFT::Server srv(f.tuples, c, f.nq, f.dq, 2, n0, T, MDFT::Precision::FP64, 3);
    srv.submit(grid, 3, sol[0], water, 30 * kT, alone.data());
    ra = srv.run().at(0);
  }
  std::vector<MDFT::SolveResult> rm;
  {
    MDFT::Server srv(f.tuples, c, f.nq, f.dq, 2, n0, T, MDFT::Precision::FP64, 3);
    for (int k = 0; k < 3; ++k) srv.submit(grid, 3, sol[k], water, 30 * kT, mix[k].data());
    rm = srv.run();
  }
  g_where = "Server";
  EXPECT_TRUE(rm.size() == 3);
  const MDFT::SolveResult* first = nullptr;
  for (const auto& r : rm)
    if (r.id == 1) first = &r;
  EXPECT_TRUE(first != nullptr);
  std::printf("    alone: status %d, %d iterations, F %.12g; in the batch: status %d, %d iterations, F %.12g\n",
              ra.status, ra.iterations, ra.F[0], first->status, first->iterations, first->F[0]);
  EXPECT_TRUE(ra.status == 0 && first->status == 0);
  EXPECT_TRUE(ra.iterations == first->iterations && ra.F[0] == first->F[0]);
  auto ha = MDFT::create_mirror_view_and_copy(alone);
  auto hm = MDFT::create_mirror_view_and_copy(mix[0]);
  EXPECT_TRUE(ha.size() == total && std::memcmp(ha.data(), hm.data(), total * sizeof(double)) == 0);
}

}  // namespace

int main(int argc, char** argv) {
  const bool host_only = argc > 1 && std::strcmp(argv[1], "--host-only") == 0;
  run("SpatialGrid<float>", [] { test_spatial_grid<float>(); });
  run("SpatialGrid<double>", [] { test_spatial_grid<double>(); });
  run("Errors<float>", [] { test_errors<float>(); });
  run("Errors<double>", [] { test_errors<double>(); });
  if (!host_only) {
    if (argc < 3) {
      std::printf("usage: test_facade --host-only | test_facade <luc_file> <np_luc>\n");
      return 2;
    }
    const std::string path = argv[1];
    const int np_luc = std::atoi(argv[2]);
    run("AngularGrid/OrientationProjectionMap<float>", [] { test_angular_grid<float>(); });
    run("AngularGrid/OrientationProjectionMap<double>", [] { test_angular_grid<double>(); });
    run("Convolution: missing Luc file", [] {
      MDFT::SpatialGrid<execution_space, double> grid({4, 4, 4}, {5, 5, 5});
      MDFT::AngularGrid<execution_space, double> ag(3, 2);
      MDFT::OrientationProjectionMap<execution_space, double> map(ag);
      expect_throw([&] { MDFT::Convolution<execution_space, double> c("/nonexistent/c.luc", grid, ag, map, 252); },
                   "does not exist");
    });
    run("OrientationProjectionTransform<float>", [] { test_angl2proj_proj2angl<float>(3, 4, 2, 3, 2, MDFT::Precision::Default); });
    run("OrientationProjectionTransform<double>", [] { test_angl2proj_proj2angl<double>(3, 4, 2, 3, 2, MDFT::Precision::Default); });
    run("OrientationProjectionTransform<double> mmax=5", [] { test_angl2proj_proj2angl<double>(4, 3, 5, 5, 2, MDFT::Precision::Default); });
    run("OrientationProjectionTransform<double> MixedX6", [] { test_angl2proj_proj2angl<double>(3, 4, 2, 3, 2, MDFT::Precision::MixedX6); });
    run("OrientationProjectionTransform<double> Mixed", [] { test_angl2proj_proj2angl<double>(3, 4, 2, 3, 2, MDFT::Precision::Mixed); });
    run("OrientationProjection Identity<float>", [] { test_identity<float>(); });
    run("OrientationProjection Identity<double>", [] { test_identity<double>(); });
    run("Convolution Execute<double> 6^3", [&] { test_convolution<double>(path, np_luc, 6, 7.0, MDFT::Precision::Default); });
    run("Convolution Execute<float> 6^3", [&] { test_convolution<float>(path, np_luc, 6, 7.0, MDFT::Precision::Default); });
    run("Convolution Execute<double> Mixed 5^3", [&] { test_convolution<double>(path, np_luc, 5, 6.0, MDFT::Precision::Mixed); });
    run("get_delta_rho/get_delta_f<float>", [] { test_solvent_functions<float>(); });
    run("get_delta_rho/get_delta_f<double>", [] { test_solvent_functions<double>(); });
    run("ExcessFunctional energy_cproj_mrso 8^3", [&] { test_excess_functional(path, np_luc); });
    run("Solvents + stages, the reference's energy_cproj_mrso 8^3", [&] { test_solvent_dropin(path, np_luc); });
    run("Functional: V_ext, eval, minimize 8^3", [&] { test_functional(path, np_luc); });
    run("Server: alone == in a batch 8^3", [&] { test_server(path, np_luc); });
  }
  std::printf("[==========] %d tests, %d failed\n", g_run, g_failed);
  return g_failed == 0 ? 0 : 1;
}
