// The full functional F = F_id + F_ext + F_exc on the GPU, the solute's
// external potential and an FP64 L-BFGS minimiser (see mdft_functional.hip).
#pragma once

#include <functional>
#include <vector>

#include "mdft_device.hpp"

namespace mdft {

// (no x 9) rotation matrix of every orientation, row major
// (common/MDFT_Grid.hpp:295-318)
std::vector<double> rotation_matrices(const AngularGrid& g);

// V_ext (N x no, device, FP64) of the solute sites on the grid.  solute and
// solvent are host arrays of rows (charge e, sigma A, epsilon kJ/mol, x, y, z A);
// the solvent positions are in the molecular frame.
void compute_vext(int nx, int ny, int nz, double lx, double ly, double lz, const AngularGrid& g,
                  const double* solute, int nsolute, const double* solvent, int nsolvent, double vmax, double* vext,
                  hipStream_t s);

// xi = exp(-V / (2 kT)) on the device (n values, on stream s): the minimum of
// F_id + F_ext alone (xi^2 = exp(-V / kT)), the density guess from V_ext of
// the original MDFT (the reference fills xi with 1 only because it has no
// V_ext, common/MDFT_Solvent.hpp:287-296).  V / kT below -600 is taken as
// -600, so that xi^2 stays finite.
void xi_from_vext(const double* V, long n, double kT, double* xi, hipStream_t s);

struct FunctionalParams {
  long N;  // grid points
  int no;  // orientations
  double rho0, prefactor, kT, dv;
};

struct LbfgsOptions {
  int m = 5;  // stored pairs, 1 .. kLbfgsMaxPairs
  int max_iter = 200;
  double gtol = 1e-5;  // converged when max |h0 g| < gtol
  double c1 = 1e-4;    // Armijo constant
  int max_backtrack = 30;
  // noise allowance of the approximate Wolfe test, relative to |F_id| + |F_ext| + |F_exc|
  double eps_f = 1e-6;
};
constexpr int kLbfgsMaxPairs = 8;

struct LbfgsResult {
  double F[4] = {0, 0, 0, 0};  // F, F_id, F_ext, F_exc at the returned point
  int iterations = 0, evaluations = 0;
  // 0 converged, 1 max_iter reached, 2 no acceptable step, 4 F or g not
  // finite at the start point
  int status = 1;
  double residual = 0.0;  // max |h0 g| at the returned point
  std::vector<double> history;  // F of the start point, then after every iteration
};

// Host side of one L-BFGS solve in vector-free form: the stored pairs'
// slots and Gram entries, the two-loop recursion in coefficients, and the line
// search decisions.  A driver launches the passes: Functional::minimize for
// one field, the batch engine for many.  Pair slots are 0 .. m (m + 1 slots:
// the candidate pair of a trial never overwrites a stored one).
class LbfgsState {
 public:
  struct Trial {
    int k = 0;     // stored pairs the direction uses: slots order()[0 .. k)
    int cand = 0;  // slot receiving the candidate pair (s = x_t - x, y = g_t - g)
    // pairs order()[0 .. k0) need no Gram entries at the trial point: with a
    // full memory (k == m) the oldest is dropped when the candidate is kept,
    // and the memory restarts when it is not
    int k0 = 0;
    double cg = 1.0, gamma = 1.0, t = 1.0;
    double cs[kLbfgsMaxPairs] = {}, cq[kLbfgsMaxPairs] = {};
  };
  explicit LbfgsState(const LbfgsOptions& o);
  // results at the start point: F (4 terms), g'H g and max |h0 g|
  void start(const double F[4], double ghg, double gmax);
  // plans the next trial; false once the solve has ended (result() is final)
  bool next(Trial& tr);
  // the trial's F (4 terms) and update-pass results h (layout of a pass
  // with K pair columns, column c holding pair order()[k0 + c], K >= the
  // trial's k - k0); true when the trial is accepted (the driver then takes
  // x_t and g_t as the current point)
  bool judge(const double Fn[4], const double* h, int K);
  const std::vector<int>& order() const { return order_; }
  const LbfgsResult& result() const { return r_; }

 private:
  double& sy(int i, int j) { return SY_[i * o_.m + j]; }
  double& yhy(int i, int j) { return YHY_[i * o_.m + j]; }
  LbfgsOptions o_;
  LbfgsResult r_;
  enum Phase { kNew, kBacktrack, kDone } phase_ = kNew;
  double F_[4] = {0, 0, 0, 0};
  double ghg_ = 0.0, gmax_ = 0.0, gd_ = 0.0;
  double t_next_ = 0.5;  // the step of the next trial after a rejected one
  int it_ = 0, tries_ = 0;
  Trial cur_;
  std::vector<int> order_;
  std::vector<double> SY_, YHY_, SG_, YHG_;
};

class Functional {
 public:
  // the excess (c-proj) evaluation: fills df (device, N x no) for xi and
  // writes ff to ff_dev (device), stream ordered, without synchronising; xi
  // and df are doubles, or floats with fp32_fields
  using ExcessEval = std::function<void(const void* xi, void* df, double* ff_dev, hipStream_t s)>;

  // w: quadrature weights (no); vext: device, N x no, borrowed (kept alive by
  // the caller); fp32_pairs: store the L-BFGS pairs and gradients in FP32
  // (dot products stay FP64); fp32_fields: the excess evaluation takes an
  // FP32 copy of the point (written by the trial pass) and gives df in FP32
  // (MDFT_MIXED_X6F); the point itself and the functional's sums stay FP64
  Functional(const FunctionalParams& p, const std::vector<double>& w, const double* vext, ExcessEval excess,
             bool fp32_pairs = false, bool fp32_fields = false);
  ~Functional();
  Functional(const Functional&) = delete;
  Functional& operator=(const Functional&) = delete;

  // F[0..3] = F, F_id, F_ext, F_exc (host); g = dF/dxi (device, N x no)
  void eval(const double* xi, double* g, double F[4], hipStream_t s);

  // minimises in place from xi (device, N x no doubles)
  LbfgsResult minimize(double* xi, const LbfgsOptions& o, hipStream_t s);

  // device bytes held by the functional, and those a minimisation with m pairs adds
  size_t bytes() const;
  size_t minimize_bytes(int m) const;

 private:
  void reserve(int m);
  void release();
  const double* read_results(hipStream_t s);
  void fparts(double sid, double sext, double ff, double F[4]) const;
  template <typename TS, typename TF>
  LbfgsResult minimize_t(double* xi, const LbfgsOptions& o, hipStream_t s);

  FunctionalParams p_;
  ExcessEval excess_;
  const double* vext_;       // borrowed
  bool fp32_pairs_ = false;
  bool fp32_fields_ = false;
  long n_ = 0;               // N * no
  int nb_ = 0;               // blocks (= partials per sum) of the element passes
  int step_ = 0;             // (blocks * threads) mod no
  double* tab_ = nullptr;    // device: w (no), h0 (no)
  void* df_ = nullptr;       // device, N x no (double, or float with fp32_fields)
  float* xf_ = nullptr;      // device, N x no: the FP32 copy of the evaluated point (fp32_fields)
  double* part_ = nullptr;   // device partials, kMaxSums x nb
  double* res_ = nullptr;    // device results, kMaxSums + 1
  double* host_ = nullptr;   // pinned copy of res_
  // minimiser workspace: the trial point, g and g at the trial (FP64), then
  // m + 1 slots of s and y
  int m_alloc_ = 0;
  void* work_ = nullptr;
};

}  // namespace mdft
