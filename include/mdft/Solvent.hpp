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
// xi, vexc, delta_f are (nx, ny, nz, no), w is the (no) weight view of the
// AngularGrid (device, double or float: a float view is widened to double
// first).  ff is returned after the stream has completed (FP64 accumulation,
// deterministic order).
template <typename ExecutionSpace, typename View1DType, typename View4DType, typename ScalarType>
void get_delta_f(const ExecutionSpace& exec_space, const View4DType& xi, const View4DType& vexc,
                 const View1DType& w, const View4DType& delta_f, ScalarType& ff, const ScalarType rho0,
                 const ScalarType prefactor) {
  using T = typename View4DType::non_const_value_type;
  using WT = typename View1DType::non_const_value_type;
  static_assert(std::is_same<WT, double>::value || std::is_same<WT, float>::value, "w is a double or float view");
  for (int i = 0; i < 4; i++)
    Impl::Throw_If(xi.extent(i) != vexc.extent(i), "angular grid size of xi and vexc must be the same");
  const size_t nxyz = xi.extent(0) * xi.extent(1) * xi.extent(2);
  const int no = static_cast<int>(xi.extent(3));
  Impl::Throw_If(w.size() != static_cast<size_t>(no), "w must hold one weight per orientation");
  Impl::Throw_If(delta_f.size() != xi.size(), "size of delta_f must be the same as the size of xi");
  const double* wd = nullptr;
  View<double, 1> wide;
  if constexpr (std::is_same<WT, double>::value) {
    wd = w.data();
  } else {
    exec_space.fence();  // w may have been written on the stream
    const auto h = create_mirror_view_and_copy(w);
    wide = View<double, 1>("w (double)", no);
    deep_copy(wide, std::vector<double>(h.begin(), h.end()));
    wd = wide.data();
  }
  double f = 0.0;
  Impl::check(mdft_delta_f(xi.data(), vexc.data(), wd, delta_f.data(), nxyz, no, static_cast<double>(rho0),
                           static_cast<double>(prefactor), &f, Impl::dtype_code<T>(), exec_space.handle()));
  ff = static_cast<ScalarType>(f);
}

// One site: charge (e), LJ sigma (A), epsilon (kJ/mol), position (A); solvent
// sites are in the molecular frame (Solute::read_solute,
// common/MDFT_Solute.hpp:126-191; Solvents::set_solvent,
// common/MDFT_Solvent.hpp:417-532).
struct Site {
  double q, sigma, epsilon, x, y, z;
};

// Thermo (common/MDFT_Thermo.hpp:26-66): kT in kJ/mol from the reference's
// values of Constants::Boltz and Constants::Navo
// (common/MDFT_Constants.hpp:43-50); temperature in K, pressure in bar.
template <typename ScalarType>
struct Thermo {
  ScalarType m_t, m_p, m_kbt, m_beta;
  explicit Thermo(ScalarType temperature, ScalarType pressure = 1)
      : m_t(temperature), m_p(pressure),
        m_kbt(static_cast<ScalarType>(1.38064852e-23 * 6.02214129e23 * temperature * 1e-3)),
        m_beta(static_cast<ScalarType>(1.0) / m_kbt) {}
};

// Solvent (common/MDFT_Solvent.hpp:61-300): the members the c-proj
// evaluation reads, m_xi (nx, ny, nz, no) LayoutRight on the device, m_n0,
// m_rho0 and m_npluc, with the site model and init_density (xi = 1).
template <typename ExecutionSpace, typename ScalarType>
struct Solvent {
  using IntType = int;
  using IntArrayType = std::array<IntType, 6>;
  using View4DType = View<ScalarType, 4>;

  std::string m_name;
  int m_molrotsymorder = 0;
  int m_nsite = 0;
  std::vector<Site> m_site;  // (q e, sigma A, epsilon kJ/mol, x, y, z A) in the molecular frame
  View4DType m_xi;           // xi^2 = rho / rho0
  ScalarType m_n0 = 0;       // number density of the bulk fluid, A^-3
  ScalarType m_rho0 = 0;     // number density per orientation, n0 / (8 pi^2 / mrso)
  ScalarType m_mole_fraction = 1;
  ScalarType m_relativePermittivity = 0;
  IntArrayType m_npluc{};    // Luc tuples of nmax = 0..5 (np_luc of mmax)

  void init_density(const int nx, const int ny, const int nz, const int no) {
    m_xi = View4DType("xi", nx, ny, nz, no);
    using T = typename View4DType::non_const_value_type;
    Impl::check(mdft_fill_uniform(m_xi.data(), m_xi.size(), 0, 0, 1.0, 1.0, Impl::dtype_code<T>(), nullptr));
    Impl::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  }
};

// Solvents (common/MDFT_Solvent.hpp:308-561) for one solvent species, named
// as the reference's Settings::m_solvent: the water models spce (= spce-h),
// spce-m, tip3p and tip3p-m of set_solvent (:417-532) with their charges,
// LJ site, geometry, n0 (times the mole fraction), relative permittivity,
// npluc and the mmax <= 5 limit; init_density; and m_prefactor =
// -2 kT 4 pi^2 / mrso / n0 (:366-368).  The species and mole fraction come as
// arguments instead of the Settings JSON file (the input reader is not part
// of this path).
template <typename ExecutionSpace, typename ScalarType>
struct Solvents {
  using SpatialGridType = SpatialGrid<ExecutionSpace, ScalarType>;
  using AngularGridType = AngularGrid<ExecutionSpace, ScalarType>;
  using SolventType = Solvent<ExecutionSpace, ScalarType>;
  using ThermoType = Thermo<ScalarType>;

  SpatialGridType m_spatial_grid;
  AngularGridType m_angular_grid;
  ScalarType m_prefactor;
  ThermoType m_thermo;
  std::vector<SolventType> m_solvents;

  Solvents(const SpatialGridType& spatial_grid, const AngularGridType& angular_grid, const std::string& solvent,
           const ThermoType& thermo, ScalarType mole_fraction = 1)
      : m_spatial_grid(spatial_grid), m_angular_grid(angular_grid), m_prefactor(0), m_thermo(thermo) {
    Impl::Throw_If(mole_fraction < 0 || mole_fraction > 1, "Critical errror. Mole fractions should be between 0 and 1");
    // (charge of O, O-H x, O-H z, sigma, epsilon, n0 of the pure fluid, eps_r)
    struct Model {
      const char* name;
      double qo, hx, hz, sigma, epsilon, n0, eps_r;
    };
    static const Model models[] = {
        {"spce", -0.8476, 0.816495, 0.5773525, 3.166, 0.65, 0.0332891, 71},
        {"spce-h", -0.8476, 0.816495, 0.5773525, 3.166, 0.65, 0.0332891, 71},
        {"spce-m", -1.0, 0.816495, 0.5773525, 3.166, 0.65, 0.0332891, 71},
        {"tip3p", -0.834, 0.756950, 0.585882, 3.15061, 0.636386, 0.03349459, 91},
        {"tip3p-m", -0.95, 0.756950, 0.585882, 3.15061, 0.636386, 0.03349459, 91},
    };
    const Model* md = nullptr;
    for (const Model& c : models)
      if (solvent == c.name) md = &c;
    Impl::Throw_If(md == nullptr, "Critical error. Solvent: " + solvent + " is not implemented yet");
    SolventType s;
    // "spce-h" and the "-m" variants are stored under the base model's name
    s.m_name = solvent.rfind("spce", 0) == 0 ? "spce" : "tip3p";
    s.m_nsite = 3;
    s.m_molrotsymorder = 2;
    s.m_npluc = {1, 6, 75, 252, 877, 2002};
    s.m_site = {{md->qo, md->sigma, md->epsilon, 0.0, 0.0, 0.0},
                {-md->qo / 2, 0.0, 0.0, md->hx, 0.0, md->hz},
                {-md->qo / 2, 0.0, 0.0, -md->hx, 0.0, md->hz}};
    s.m_mole_fraction = mole_fraction;
    s.m_n0 = static_cast<ScalarType>(md->n0 * mole_fraction);
    s.m_rho0 = static_cast<ScalarType>(s.m_n0 / (8 * M_PI * M_PI / s.m_molrotsymorder));
    s.m_relativePermittivity = static_cast<ScalarType>(md->eps_r);
    Impl::Throw_If(m_angular_grid.m_mmax > 5 || m_angular_grid.m_mmax < 0,
                   "solvent " + s.m_name + " only avail with mmax between 0 and 5");
    Impl::Throw_If(m_angular_grid.m_molrotsymorder != s.m_molrotsymorder,
                   "molrotsymorder of the angular grid and the solvent must be the same");
    m_solvents.push_back(s);
    init_density();
    const int mrso = m_angular_grid.m_molrotsymorder;
    m_prefactor = static_cast<ScalarType>(-2 * m_thermo.m_kbt * 4 * M_PI * M_PI / mrso / m_solvents.at(0).m_n0);
  }

  void init_density() {
    for (auto& solvent : m_solvents)
      solvent.init_density(m_spatial_grid.m_nx, m_spatial_grid.m_ny, m_spatial_grid.m_nz, m_angular_grid.m_no);
  }
};

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
