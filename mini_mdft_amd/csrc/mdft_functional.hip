// The full functional, the solute's external potential and an FP64 L-BFGS
// minimiser on the GPU.
//
// The reference makes one excess evaluation (src/mini-mdft.hpp:129-133) and
// leaves the external potential and the iteration settings unused
// (common/MDFT_Solvent.hpp:160, common/MDFT_System.hpp:188-189).  The terms
// used here, per grid node r and orientation o with quadrature weight w_o:
//   F_id  = kT rho0 dv sum w (xi^2 ln xi^2 - xi^2 + 1)
//   F_ext = rho0 dv sum w xi^2 V
//   F_exc = (dv / N) ff        ff from energy_cproj; its unnormalised FFTs
//                              make ff N times the discrete integral
//   g     = dv w (2 rho0 xi (kT ln xi^2 + V) + (2 prefactor / N) df)
// with df = 2 rho0 xi vexc, the reference's gradient array
// (common/MDFT_Solvent.hpp:643-657).  xi^2 is floored at 1e-300 inside ln.
//
// The minimiser is L-BFGS with the diagonal initial inverse Hessian gamma h0,
// h0 = 1 / (2 rho0 dv w kT), a backtracking line search that accepts Armijo
// or the approximate Wolfe test of Hager and Zhang (SIAM J. Optim. 16 (2005)
// 170-192), and restarts on a non-descent direction.  Trials start at t = 1,
// or along -h0 g (no stored pair) at t = min(1, 1 / max |h0 g|), so that no
// xi moves by more than 1.  After a rejected trial t the next is the
// minimiser of the quadratic through F(0), F'(0) and F(t), kept within
// [0.1 t, 0.5 t] (Nocedal and Wright, Numerical Optimization, 2nd ed.,
// section 3.5); a trial where F is not finite halves t.  It runs in the
// vector-free form of Chen, Wang, Zhou and Gao ("Large-scale L-BFGS using
// MapReduce", NIPS 2014): the two-loop recursion is carried out on the host
// in the coefficients of the direction over the basis {s_i, y_i, g}, from the
// Gram entries s_i.y_j, y_i'H y_j, s_i.g, y_i'H g, g'H g (H = diag(h0)).
// Every iteration then makes two fused passes over HBM: one that forms the
// trial point x + t d, one (after the excess evaluation there) that forms g,
// y = g_new - g, the F_id and F_ext sums and the Gram entries the next
// direction can use.  With a full memory the oldest pair gets none: a kept
// candidate drops it, and a candidate that fails the curvature test
// (s'y <= 1e-12 |s| |y|) restarts the memory, which spares the pass that
// pair's reads (all of them at m = 1).  All sums are FP64 with fixed-order
// reductions.
#include "mdft_functional.hpp"
#include "mdft_elem.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>

namespace mdft {
namespace {

using namespace elem;

// g, F_id and F_ext sums, g'H g and max |h0 g| at x; g is stored as TG
// (the sums take the stored value)
template <typename TG, typename TF>
__global__ __launch_bounds__(kFThreads) void local_kernel(const double* __restrict__ x, const double* __restrict__ V,
                                                          const TF* __restrict__ df, TG* __restrict__ g,
                                                          Elem e, double* __restrict__ part, int nb) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  MDFT_ELEMENT_LOOP(e) {
    const double xv = ld_nt(&x[i]), v = ld_nt(&V[i]), wio = e.w[io], h = e.h0[io];
    const double dfv = static_cast<double>(ld_nt(&df[i]));
    const TG gs = static_cast<TG>(local_terms(xv, v, dfv, wio, e, acc[0], acc[1]));
    st_nt(&g[i], gs);
    const double gv = static_cast<double>(gs);
    const double hg = h * gv;
    acc[2] += gv * hg;
    acc[3] = fmax(acc[3], fabs(hg));
  }
  store_partials<4, true>(acc, part, nb);
}

// The stored pairs s_j, y_j, the candidate pair and the gradients g, g_t are
// TS (double, or float for the mixed forms, which then move half their
// bytes); every dot product and the direction are formed in FP64 from the
// stored values.
template <typename TS>
struct DirArgs {
  const double* x;
  const TS* g;
  const TS* s[kLbfgsMaxPairs];
  const TS* y[kLbfgsMaxPairs];
  double cs[kLbfgsMaxPairs], cq[kLbfgsMaxPairs];
  double cg, gamma, t;
  double* xt;  // x + t d
  float* xf;   // its FP32 copy for the excess evaluation, or null
  TS* sc;      // xt - x (the candidate s)
};

// xf = x rounded to FP32 (the evaluation's copy of the start point)
__global__ __launch_bounds__(kFThreads) void to_float_kernel(const double* __restrict__ x, float* __restrict__ xf,
                                                             long n) {
  for (long i = static_cast<long>(blockIdx.x) * kFThreads + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * kFThreads)
    xf[i] = static_cast<float>(x[i]);
}

// d = -(gamma h0 (cg g + sum cq_j y_j) + sum cs_j s_j); xt = x + t d; sc = xt - x
template <typename TS, int K>
__global__ __launch_bounds__(kFThreads) void direction_kernel(DirArgs<TS> a, Elem e) {
  MDFT_ELEMENT_LOOP(e) {
    double q = a.cg * static_cast<double>(ld_nt(&a.g[i]));
#pragma unroll
    for (int j = 0; j < K; ++j) q = fma(a.cq[j], static_cast<double>(ld_nt(&a.y[j][i])), q);
    double r = a.gamma * e.h0[io] * q;
#pragma unroll
    for (int j = 0; j < K; ++j) r = fma(a.cs[j], static_cast<double>(ld_nt(&a.s[j][i])), r);
    const double xv = ld_nt(&a.x[i]);
    const double xn = xv + a.t * (-r);
    st_nt(&a.xt[i], xn);
    if (a.xf) st_nt(&a.xf[i], static_cast<float>(xn));
    st_nt(&a.sc[i], static_cast<TS>(xn - xv));
  }
}

template <typename TS, typename TF>
struct UpdArgs {
  const double* xt;
  const double* V;
  const TF* df;
  const TS* g;  // g at x
  TS* gt;       // out: g at xt
  const TS* sc;
  TS* yn;  // out: g at xt - g at x
  const TS* s[kLbfgsMaxPairs];
  const TS* y[kLbfgsMaxPairs];
};

// At the trial point: g, y = g_t - g, the F_id and F_ext sums and the Gram
// entries of the candidate pair:
//   [0, K) sc.y_j | K sc.yn | [K+1, 2K+1) s_j.yn | 2K+1 sc.sc | 2K+2 yn.yn |
//   [2K+3, 3K+3) s_j.g_t | 3K+3 sc.g_t | [3K+4, 4K+4) yn'H y_j | 4K+4 yn'H yn |
//   [4K+5, 5K+5) y_j'H g_t | 5K+5 yn'H g_t | 5K+6 g_t'H g_t | 5K+7 F_id sum |
//   5K+8 F_ext sum | 5K+9 max |h0 g_t|
// g_t and yn enter the sums as stored (rounded to TS).
// Each trip takes two elements, i and i + stride, and requests both elements'
// loads before either's arithmetic: twice the bytes in flight per wave (one
// element per trip left this pass at 4.3 TB/s, 0.72 ms at 64^3/mmax 5, with
// 36 B per element; two: 0.59 ms).  The sums take the elements in the order
// of the one-element loop, so the results are bitwise the same.
template <typename TS, typename TF, int K>
__global__ __launch_bounds__(kFThreads) void update_kernel(UpdArgs<TS, TF> a, Elem e, double* __restrict__ part,
                                                           int nb) {
  constexpr int NA = 5 * K + 10;
  constexpr int KA = K > 0 ? K : 1;
  double acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = 0.0;
  const long stride_ = static_cast<long>(gridDim.x) * kFThreads;
  long i = static_cast<long>(blockIdx.x) * kFThreads + threadIdx.x;
  int io = static_cast<int>(i % e.no);
  auto nio = [&](int x) { return x + e.step >= e.no ? x + e.step - e.no : x + e.step; };
  for (; i < e.n; i += 2 * stride_, io = nio(nio(io))) {
    const bool two = i + stride_ < e.n;
    long ix[2] = {i, two ? i + stride_ : i};
    int iox[2] = {io, nio(io)};
    double xv[2], v[2], dfv[2], gv[2], scv[2], sj[2][KA], yj[2][KA];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      xv[u] = ld_nt(&a.xt[ix[u]]);
      v[u] = ld_nt(&a.V[ix[u]]);
      dfv[u] = static_cast<double>(ld_nt(&a.df[ix[u]]));
      gv[u] = static_cast<double>(ld_nt(&a.g[ix[u]]));
      scv[u] = static_cast<double>(ld_nt(&a.sc[ix[u]]));
#pragma unroll
      for (int j = 0; j < K; ++j) {
        sj[u][j] = static_cast<double>(ld_nt(&a.s[j][ix[u]]));
        yj[u][j] = static_cast<double>(ld_nt(&a.y[j][ix[u]]));
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (u == 1 && !two) break;
      const double wio = e.w[iox[u]], h = e.h0[iox[u]];
      const TS gs = static_cast<TS>(local_terms(xv[u], v[u], dfv[u], wio, e, acc[5 * K + 7], acc[5 * K + 8]));
      const double gn = static_cast<double>(gs);
      const TS ys = static_cast<TS>(gn - gv[u]);
      st_nt(&a.gt[ix[u]], gs);
      st_nt(&a.yn[ix[u]], ys);
      const double yv = static_cast<double>(ys);
      const double hy = h * yv, hg = h * gn;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        acc[j] += scv[u] * yj[u][j];
        acc[K + 1 + j] += sj[u][j] * yv;
        acc[2 * K + 3 + j] += sj[u][j] * gn;
        acc[3 * K + 4 + j] += hy * yj[u][j];
        acc[4 * K + 5 + j] += yj[u][j] * hg;
      }
      acc[K] += scv[u] * yv;
      acc[2 * K + 1] += scv[u] * scv[u];
      acc[2 * K + 2] += yv * yv;
      acc[3 * K + 3] += scv[u] * gn;
      acc[4 * K + 4] += hy * yv;
      acc[5 * K + 5] += yv * hg;
      acc[5 * K + 6] += gn * hg;
      acc[5 * K + 9] = fmax(acc[5 * K + 9], fabs(hg));
    }
  }
  store_partials<NA, true>(acc, part, nb);
}
#undef MDFT_ELEMENT_LOOP

// out[row] = sum (or, for row == max_row, max) of part[row * nb .. + nb), in a
// fixed order, for rows first .. first + gridDim.x
__global__ __launch_bounds__(kFThreads) void reduce_rows_kernel(const double* __restrict__ part, int nb, int first,
                                                                int max_row, double* __restrict__ out) {
  const int row = first + blockIdx.x;
  const bool is_max = row == max_row;
  double v = 0.0;
  for (int i = threadIdx.x; i < nb; i += kFThreads) {
    const double p = part[static_cast<long>(row) * nb + i];
    v = is_max ? fmax(v, p) : v + p;
  }
  __shared__ double red[kFThreads];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int st = kFThreads / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = is_max ? fmax(red[threadIdx.x], red[threadIdx.x + st])
                                                    : red[threadIdx.x] + red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[row] = red[0];
}

// ---- external potential ----------------------------------------------------
struct VextArgs {
  long n;
  int no, nyz, nz, nsolv, nsolute;
  double dx, dy, dz, lx, ly, lz, ilx, ily, ilz, vmax;
};

// V(r, o) = min(vmax, sum_k sum_j [LJ_kj(d) + qq_kj / d]) with d the
// minimum-image distance between solvent site k of orientation o at node r
// and solute site j, floored at 1e-6 A; pair[(k nsolute + j) * 3] =
// (sigma, 4 eps, qfact q q), sigma = 0 where either site has no LJ.  1/d is
// one rsqrt of d^2, clamped at 1e6 (the floor): against a square root and a
// division, 8.4 -> 5.3 ms at 64^3 / mmax 5 with 20 x 3 site pairs.  Sites at
// the molecular origin are summed once per node (vext_node_kernel); when they
// lead the site list, as the oxygen does, the sum keeps its order.

// v += sum_j [LJ_kj(d) + qq_kj / d] for solvent site k at (px, py, pz)
__device__ __forceinline__ void vext_site(const VextArgs& a, double px, double py, double pz, int k,
                                          const double* __restrict__ pair, const double* __restrict__ solute,
                                          double& v) {
  for (int j = 0; j < a.nsolute; ++j) {
    double ddx = px - solute[3 * j], ddy = py - solute[3 * j + 1], ddz = pz - solute[3 * j + 2];
    ddx -= a.lx * rint(ddx * a.ilx);
    ddy -= a.ly * rint(ddy * a.ily);
    ddz -= a.lz * rint(ddz * a.ilz);
    const double inv = fmin(rsqrt(ddx * ddx + ddy * ddy + ddz * ddz), 1e6);
    const double* pp = pair + (static_cast<long>(k) * a.nsolute + j) * 3;
    if (pp[0] > 0.0) {
      const double sr = pp[0] * inv, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2;
      v += pp[1] * (sr6 * sr6 - sr6);
    }
    v += pp[2] * inv;
  }
}

__device__ __forceinline__ void node_position(const VextArgs& a, long r, double& gx, double& gy, double& gz) {
  const int ix = static_cast<int>(r / a.nyz);
  const int rem = static_cast<int>(r - static_cast<long>(ix) * a.nyz);
  const int iy = rem / a.nz, iz = rem - iy * a.nz;
  gx = ix * a.dx;
  gy = iy * a.dy;
  gz = iz * a.dz;
}

// Vr(r) = the terms of the solvent sites at the molecular origin (the bits
// of `origin`, such as the oxygen of tip3p or spce), which no orientation
// moves: once per node instead of once per (node, orientation)
__global__ __launch_bounds__(kFThreads) void vext_node_kernel(VextArgs a, unsigned origin,
                                                              const double* __restrict__ pair,
                                                              const double* __restrict__ solute,
                                                              double* __restrict__ Vr) {
  const long nr = a.n / a.no;
  for (long r = static_cast<long>(blockIdx.x) * kFThreads + threadIdx.x; r < nr;
       r += static_cast<long>(gridDim.x) * kFThreads) {
    double gx, gy, gz;
    node_position(a, r, gx, gy, gz);
    double v = 0.0;
    for (int k = 0; k < a.nsolv; ++k)
      if ((origin >> k) & 1u) vext_site(a, gx, gy, gz, k, pair, solute, v);
    Vr[r] = v;
  }
}

// V(r, o): Vr(r) (or 0 without origin sites), then the other sites in order
__global__ __launch_bounds__(kFThreads) void vext_kernel(VextArgs a, unsigned origin, const double* __restrict__ Vr,
                                                         const double* __restrict__ off,
                                                         const double* __restrict__ pair,
                                                         const double* __restrict__ solute, double* __restrict__ V) {
  for (long idx = static_cast<long>(blockIdx.x) * kFThreads + threadIdx.x; idx < a.n;
       idx += static_cast<long>(gridDim.x) * kFThreads) {
    const long r = idx / a.no;
    const int io = static_cast<int>(idx - r * a.no);
    double gx, gy, gz;
    node_position(a, r, gx, gy, gz);
    double v = Vr ? Vr[r] : 0.0;
    for (int k = 0; k < a.nsolv; ++k) {
      if ((origin >> k) & 1u) continue;
      const double* o = off + (static_cast<long>(io) * a.nsolv + k) * 3;
      vext_site(a, gx + o[0], gy + o[1], gz + o[2], k, pair, solute, v);
    }
    V[idx] = fmin(v, a.vmax);
  }
}

__global__ __launch_bounds__(kFThreads) void xi_from_vext_kernel(const double* __restrict__ V, long n, double ikT,
                                                                 double* __restrict__ xi) {
  for (long i = static_cast<long>(blockIdx.x) * kFThreads + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * kFThreads)
    xi[i] = exp(-0.5 * fmax(V[i] * ikT, -600.0));
}

// Constants::qfact (common/MDFT_Constants.hpp:36-47): qfact q q / r in kJ/mol
constexpr double kQunit = 1.602176565e-19, kNavo = 6.02214129e23, kEps0 = 8.854187817e-12;
constexpr double kPi = 3.14159265358979323846;
const double kQfact = kQunit * kQunit * 1.0e-3 * kNavo / (4.0 * kPi * kEps0 * 1.0e-10);


}  // namespace

// ---------------------------------------------------------------------------
std::vector<double> rotation_matrices(const AngularGrid& g) {
  // dphi, dpsi and phi_j = j dphi, psi_k = k dpsi as AngularGrid sets them up
  // (common/MDFT_Grid.hpp:141-142, :181-182)
  const double dphi = 2.0 * kPi / g.nphi, dpsi = 2.0 * kPi / g.mrso / g.npsi;
  std::vector<double> R(static_cast<size_t>(g.no) * 9);
  int io = 0;
  for (int it = 0; it < g.ntheta; ++it)
    for (int ip = 0; ip < g.nphi; ++ip)
      for (int ik = 0; ik < g.npsi; ++ik, ++io) {
        const double ct = std::cos(g.theta[it]), st = std::sin(g.theta[it]);
        const double cp = std::cos(ip * dphi), sp = std::sin(ip * dphi);
        const double cs = std::cos(ik * dpsi), ss = std::sin(ik * dpsi);
        double* r = &R[static_cast<size_t>(io) * 9];
        r[0] = ct * cp * cs - sp * ss;
        r[1] = -ct * cp * ss - sp * cs;
        r[2] = st * cp;
        r[3] = ct * sp * cs + cp * ss;
        r[4] = -ct * sp * ss + cp * cs;
        r[5] = st * sp;
        r[6] = -st * cs;
        r[7] = st * ss;
        r[8] = ct;
      }
  return R;
}

void compute_vext(int nx, int ny, int nz, double lx, double ly, double lz, const AngularGrid& g,
                  const double* solute, int nsolute, const double* solvent, int nsolvent, double vmax, double* vext,
                  hipStream_t s) {
  throw_if(nx <= 0 || ny <= 0 || nz <= 0 || !(lx > 0) || !(ly > 0) || !(lz > 0), "vext: bad grid");
  throw_if(nsolute < 0 || nsolvent <= 0 || (nsolute > 0 && !solute) || !solvent || !vext, "vext: bad sites");
  const std::vector<double> R = rotation_matrices(g);
  std::vector<double> off(static_cast<size_t>(g.no) * nsolvent * 3);
  for (int io = 0; io < g.no; ++io)
    for (int k = 0; k < nsolvent; ++k)
      for (int c = 0; c < 3; ++c) {
        const double* r = &R[static_cast<size_t>(io) * 9 + 3 * c];
        const double* sv = solvent + 6 * k + 3;
        off[(static_cast<size_t>(io) * nsolvent + k) * 3 + c] = r[0] * sv[0] + r[1] * sv[1] + r[2] * sv[2];
      }
  std::vector<double> pair(static_cast<size_t>(nsolvent) * std::max(nsolute, 1) * 3, 0.0);
  std::vector<double> pos(static_cast<size_t>(std::max(nsolute, 1)) * 3, 0.0);
  for (int j = 0; j < nsolute; ++j)
    for (int c = 0; c < 3; ++c) pos[3 * j + c] = solute[6 * j + 3 + c];
  for (int k = 0; k < nsolvent; ++k)
    for (int j = 0; j < nsolute; ++j) {
      const double qk = solvent[6 * k], sk = solvent[6 * k + 1], ek = solvent[6 * k + 2];
      const double qj = solute[6 * j], sj = solute[6 * j + 1], ej = solute[6 * j + 2];
      double* p = &pair[(static_cast<size_t>(k) * nsolute + j) * 3];
      if (sk > 0 && sj > 0) {  // Lorentz-Berthelot mixing
        p[0] = 0.5 * (sk + sj);
        p[1] = 4.0 * std::sqrt(ek * ej);
      }
      p[2] = kQfact * qk * qj;
    }
  const long n = static_cast<long>(nx) * ny * nz * g.no;
  // one-time tables: plain allocations, freed on every exit path
  struct Buf {
    double* p = nullptr;
    ~Buf() {
      if (p) (void)hipFree(p);
    }
  } doff, dpair, dpos, dvr;
  MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&doff.p), off.size() * sizeof(double)));
  MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&dpair.p), pair.size() * sizeof(double)));
  MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&dpos.p), pos.size() * sizeof(double)));
  MDFT_HIP_CHECK(hipMemcpyAsync(doff.p, off.data(), off.size() * sizeof(double), hipMemcpyHostToDevice, s));
  MDFT_HIP_CHECK(hipMemcpyAsync(dpair.p, pair.data(), pair.size() * sizeof(double), hipMemcpyHostToDevice, s));
  MDFT_HIP_CHECK(hipMemcpyAsync(dpos.p, pos.data(), pos.size() * sizeof(double), hipMemcpyHostToDevice, s));
  VextArgs a{n, g.no, ny * nz, nz, nsolvent, nsolute, lx / nx, ly / ny, lz / nz, lx, ly, lz,
             1.0 / lx, 1.0 / ly, 1.0 / lz, vmax};
  // solvent sites at the molecular origin (all of the site's offsets zero)
  unsigned origin = 0;
  if (nsolvent <= 32)
    for (int k = 0; k < nsolvent; ++k)
      if (solvent[6 * k + 3] == 0.0 && solvent[6 * k + 4] == 0.0 && solvent[6 * k + 5] == 0.0) origin |= 1u << k;
  const long nr = static_cast<long>(nx) * ny * nz;
  if (origin) {
    MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&dvr.p), nr * sizeof(double)));
    hipLaunchKernelGGL(vext_node_kernel, dim3(grid_for(nr, kFThreads, 65536)), dim3(kFThreads), 0, s, a, origin,
                       dpair.p, dpos.p, dvr.p);
    MDFT_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(vext_kernel, dim3(grid_for(n, kFThreads, 65536)), dim3(kFThreads), 0, s, a, origin, dvr.p,
                     doff.p, dpair.p, dpos.p, vext);
  MDFT_HIP_CHECK(hipGetLastError());
  // the tables (host and device) are released on return
  MDFT_HIP_CHECK(hipStreamSynchronize(s));
}

void xi_from_vext(const double* V, long n, double kT, double* xi, hipStream_t s) {
  throw_if(n < 0 || (n > 0 && (!V || !xi)) || !(kT > 0), "xi_from_vext: bad arguments");
  if (n == 0) return;
  hipLaunchKernelGGL(xi_from_vext_kernel, dim3(grid_for(n, kFThreads, 65536)), dim3(kFThreads), 0, s, V, n,
                     1.0 / kT, xi);
  MDFT_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
namespace {
// rows of the partial sums and slots of the results: [0, kRows) the update
// pass (and the local pass), kFfSlot the excess evaluation's ff
constexpr int kRows = 5 * kLbfgsMaxPairs + 10;
constexpr int kFfSlot = kRows;
constexpr int kResults = kRows + 1;

void reduce_rows(const double* part, int nb, int first, int rows, int max_row, double* out, hipStream_t s) {
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(rows), dim3(kFThreads), 0, s, part, nb, first, max_row, out);
  MDFT_HIP_CHECK(hipGetLastError());
}
}  // namespace

Functional::Functional(const FunctionalParams& p, const std::vector<double>& w, const double* vext, ExcessEval excess,
                       bool fp32_pairs, bool fp32_fields)
    : p_(p), excess_(std::move(excess)), vext_(vext), fp32_pairs_(fp32_pairs), fp32_fields_(fp32_fields) {
  throw_if(p.N <= 0 || p.no <= 0 || static_cast<int>(w.size()) != p.no, "functional: bad sizes");
  throw_if(!vext, "functional: null vext");
  throw_if(!(p.kT > 0) || !(p.dv > 0) || !(p.rho0 > 0), "functional: kT, dv and rho0 must be positive");
  n_ = p.N * p.no;
  const long want = (n_ + kFThreads * 16L - 1) / (kFThreads * 16L);
  nb_ = static_cast<int>(std::min<long>(std::max<long>(want, 1), kMaxBlocks));
  step_ = static_cast<int>((static_cast<long>(nb_) * kFThreads) % p.no);
  std::vector<double> tab(2 * static_cast<size_t>(p.no));
  for (int io = 0; io < p.no; ++io) {
    tab[io] = w[io];
    tab[p.no + io] = 1.0 / (2 * p.rho0 * p.dv * w[io] * p.kT);
  }
  try {
    MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&tab_), tab.size() * sizeof(double)));
    MDFT_HIP_CHECK(hipMemcpy(tab_, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    MDFT_HIP_CHECK(hipMalloc(&df_, n_ * (fp32_fields_ ? sizeof(float) : sizeof(double))));
    if (fp32_fields_) MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&xf_), n_ * sizeof(float)));
    MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&part_), static_cast<size_t>(kRows) * nb_ * sizeof(double)));
    MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&res_), kResults * sizeof(double)));
    MDFT_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_), kResults * sizeof(double)));
  } catch (...) {
    release();
    throw;
  }
}

Functional::~Functional() { release(); }

void Functional::release() {
  if (tab_) (void)hipFree(tab_);
  if (df_) (void)hipFree(df_);
  if (xf_) (void)hipFree(xf_);
  xf_ = nullptr;
  if (part_) (void)hipFree(part_);
  if (res_) (void)hipFree(res_);
  if (host_) (void)hipHostFree(host_);
  if (work_) (void)hipFree(work_);
  tab_ = part_ = res_ = host_ = nullptr;
  df_ = nullptr;
  work_ = nullptr;
  m_alloc_ = 0;
}

size_t Functional::bytes() const {
  return (2 * static_cast<size_t>(p_.no) + static_cast<size_t>(kRows) * nb_ + kResults) * sizeof(double) +
         n_ * (fp32_fields_ ? 2 * sizeof(float) : sizeof(double)) + (m_alloc_ ? minimize_bytes(m_alloc_) : 0);
}

size_t Functional::minimize_bytes(int m) const {
  // the trial point in FP64, two gradients and m + 1 slots of s and y in TS
  const size_t ts = fp32_pairs_ ? sizeof(float) : sizeof(double);
  return static_cast<size_t>(n_) * sizeof(double) + 2 * static_cast<size_t>(m + 2) * n_ * ts;
}

void Functional::reserve(int m) {
  if (m <= m_alloc_ && work_) return;
  if (work_) (void)hipFree(work_);
  work_ = nullptr;
  m_alloc_ = 0;
  MDFT_HIP_CHECK(hipMalloc(&work_, minimize_bytes(m)));
  m_alloc_ = m;
}

const double* Functional::read_results(hipStream_t s) {
  MDFT_HIP_CHECK(hipMemcpyAsync(host_, res_, kResults * sizeof(double), hipMemcpyDeviceToHost, s));
  MDFT_HIP_CHECK(hipStreamSynchronize(s));
  return host_;
}

void Functional::fparts(double sid, double sext, double ff, double F[4]) const {
  F[1] = p_.kT * p_.rho0 * p_.dv * sid;
  F[2] = p_.rho0 * p_.dv * sext;
  F[3] = p_.dv / p_.N * ff;
  F[0] = F[1] + F[2] + F[3];
}

void Functional::eval(const double* xi, double* g, double F[4], hipStream_t s) {
  throw_if(!xi || !g || !F, "functional eval: null argument");
  const Elem e{tab_, tab_ + p_.no, p_.rho0, p_.kT, p_.dv, 2.0 * p_.prefactor / p_.N, n_, p_.no, step_};
  if (fp32_fields_) {
    hipLaunchKernelGGL(to_float_kernel, dim3(nb_), dim3(kFThreads), 0, s, xi, xf_, n_);
    excess_(xf_, df_, res_ + kFfSlot, s);
    hipLaunchKernelGGL((local_kernel<double, float>), dim3(nb_), dim3(kFThreads), 0, s, xi, vext_,
                       static_cast<const float*>(df_), g, e, part_, nb_);
  } else {
    excess_(xi, df_, res_ + kFfSlot, s);
    hipLaunchKernelGGL((local_kernel<double, double>), dim3(nb_), dim3(kFThreads), 0, s, xi, vext_,
                       static_cast<const double*>(df_), g, e, part_, nb_);
  }
  MDFT_HIP_CHECK(hipGetLastError());
  reduce_rows(part_, nb_, 0, 4, 3, res_, s);
  const double* h = read_results(s);
  fparts(h[0], h[1], h[kFfSlot], F);
}

LbfgsResult Functional::minimize(double* xi, const LbfgsOptions& o, hipStream_t s) {
  throw_if(!xi, "lbfgs: null xi");
  throw_if(o.m < 1 || o.m > kLbfgsMaxPairs, "lbfgs: m must be in 1 .. 8");
  throw_if(o.max_iter < 0 || o.max_backtrack < 1 || !(o.gtol >= 0) || !(o.c1 > 0 && o.c1 < 0.5) ||
               !(o.eps_f >= 0),
           "lbfgs: bad options");
  reserve(o.m);
  if (fp32_fields_) {
    throw_if(!fp32_pairs_, "functional: FP32 fields take FP32 pairs");
    return minimize_t<float, float>(xi, o, s);
  }
  return fp32_pairs_ ? minimize_t<float, double>(xi, o, s) : minimize_t<double, double>(xi, o, s);
}

LbfgsState::LbfgsState(const LbfgsOptions& o)
    : o_(o), SY_(o.m * o.m), YHY_(o.m * o.m), SG_(o.m), YHG_(o.m) {}

void LbfgsState::start(const double F[4], double ghg, double gmax) {
  std::memcpy(F_, F, sizeof(F_));
  ghg_ = ghg;
  gmax_ = gmax;
  r_ = LbfgsResult();
  r_.evaluations = 1;
  r_.history.assign(1, F[0]);
  order_.clear();
  it_ = 0;
  phase_ = kNew;
}

bool LbfgsState::next(Trial& tr) {
  if (phase_ == kDone) return false;
  if (phase_ == kBacktrack) {
    cur_.t = t_next_;
    tr = cur_;
    return true;
  }
  if (it_ == 0 && !(std::isfinite(F_[0]) && std::isfinite(gmax_) && std::isfinite(ghg_))) {
    // a start point where F or g is not finite (NaN or Inf in xi or V_ext)
    r_.status = 4;
    r_.iterations = 0;
    r_.residual = gmax_;
    std::memcpy(r_.F, F_, sizeof(F_));
    phase_ = kDone;
    return false;
  }
  if (gmax_ < o_.gtol || it_ == o_.max_iter) {
    r_.status = gmax_ < o_.gtol ? 0 : 1;
    r_.iterations = it_;
    r_.residual = gmax_;
    std::memcpy(r_.F, F_, sizeof(F_));
    phase_ = kDone;
    return false;
  }
  // two-loop recursion in coefficients: q = cg g + sum cq_j y_j,
  // r = gamma h0 q + sum cs_j s_j, d = -r
  Trial t;
  int k = static_cast<int>(order_.size());
  double al[kLbfgsMaxPairs] = {};
  for (int i = k - 1; i >= 0; --i) {
    double sq = t.cg * SG_[i];
    for (int j = 0; j < k; ++j) sq += t.cq[j] * sy(i, j);
    al[i] = sq / sy(i, i);
    t.cq[i] -= al[i];
  }
  t.gamma = k ? sy(k - 1, k - 1) / yhy(k - 1, k - 1) : 1.0;
  for (int i = 0; i < k; ++i) {
    double yq = t.cg * YHG_[i];
    for (int j = 0; j < k; ++j) yq += t.cq[j] * yhy(i, j);
    double yr = t.gamma * yq;
    for (int j = 0; j < k; ++j) yr += t.cs[j] * sy(j, i);
    t.cs[i] += al[i] - yr / sy(i, i);
  }
  double gq = t.cg * ghg_;
  for (int j = 0; j < k; ++j) gq += t.cq[j] * YHG_[j];
  double gd = t.gamma * gq;
  for (int j = 0; j < k; ++j) gd += t.cs[j] * SG_[j];
  gd = -gd;
  if (!(gd < 0)) {  // not a descent direction: restart from -h0 g
    order_.clear();
    k = 0;
    t = Trial();
    gd = -ghg_;
  }
  t.k = k;
  t.k0 = k > 0 && k == o_.m ? 1 : 0;
  t.cand = 0;
  while (std::find(order_.begin(), order_.end(), t.cand) != order_.end()) ++t.cand;
  // a step along -h0 g (no stored pair) moves no xi by more than 1
  t.t = k == 0 && gmax_ > 0 ? std::min(1.0, 1.0 / gmax_) : 1.0;
  gd_ = gd;
  tries_ = 0;
  cur_ = t;
  tr = t;
  return true;
}

bool LbfgsState::judge(const double Fn[4], const double* h, int K) {
  ++r_.evaluations;
  ++tries_;
  const double t = cur_.t;
  // slope at the trial point: g_t . d = (s . g_t) / t
  const double gtd = h[3 * K + 3] / t;
  const double fscale = std::fabs(F_[1]) + std::fabs(F_[2]) + std::fabs(F_[3]);
  const bool ok = std::isfinite(Fn[0]) && (Fn[0] <= F_[0] + o_.c1 * t * gd_ ||
                                           (Fn[0] <= F_[0] + o_.eps_f * fscale && gtd <= (2 * o_.c1 - 1) * gd_));
  if (!ok) {
    // quadratic interpolation: phi(0) = F, phi'(0) = gd, phi(t) = Fn; a
    // rejected finite trial fails Armijo, so den > 0 there
    const double den = 2.0 * (Fn[0] - F_[0] - gd_ * t);
    double tn = 0.5 * t;
    if (std::isfinite(Fn[0]) && den > 0) tn = std::min(std::max(-gd_ * t * t / den, 0.1 * t), 0.5 * t);
    t_next_ = tn;
    if (tries_ >= o_.max_backtrack) {
      r_.status = 2;
      r_.iterations = it_;
      r_.residual = gmax_;
      std::memcpy(r_.F, F_, sizeof(F_));
      phase_ = kDone;
    } else {
      phase_ = kBacktrack;
    }
    return false;
  }
  // the new Gram entries: pair i >= k0 in column i - k0 (columns beyond are
  // not used); pairs below k0 have none
  const int m = o_.m;
  int k = cur_.k;
  const int k0 = cur_.k0;
  const double nan = std::nan("");
  std::vector<double> sn_y(k, nan), s_yn(k, nan), yn_hy(k, nan);
  for (int i = k0; i < k; ++i) {
    const int c = i - k0;
    sn_y[i] = h[c];
    s_yn[i] = h[K + 1 + c];
    yn_hy[i] = h[3 * K + 4 + c];
    SG_[i] = h[2 * K + 3 + c];
    YHG_[i] = h[4 * K + 5 + c];
  }
  const double snyn = h[K], snsn = h[2 * K + 1], ynyn = h[2 * K + 2], sn_g = h[3 * K + 3], ynhyn = h[4 * K + 4],
               yn_hg = h[5 * K + 5];
  ghg_ = h[5 * K + 6];
  gmax_ = h[5 * K + 9];
  if (!(snyn > 1e-12 * std::sqrt(snsn * ynyn)) && k0 > 0) {
    // a full memory whose candidate fails the curvature test: the pairs below
    // k0 have no entries at the new point, so the memory restarts
    order_.clear();
  } else if (snyn > 1e-12 * std::sqrt(snsn * ynyn)) {  // keep the pair
    if (k == m) {  // drop the oldest
      order_.erase(order_.begin());
      for (int i = 0; i + 1 < m; ++i) {
        for (int j = 0; j + 1 < m; ++j) {
          sy(i, j) = sy(i + 1, j + 1);
          yhy(i, j) = yhy(i + 1, j + 1);
        }
        SG_[i] = SG_[i + 1];
        YHG_[i] = YHG_[i + 1];
      }
      sn_y.erase(sn_y.begin());
      s_yn.erase(s_yn.begin());
      yn_hy.erase(yn_hy.begin());
      k = m - 1;
    }
    order_.push_back(cur_.cand);
    for (int j = 0; j < k; ++j) {
      sy(k, j) = sn_y[j];
      sy(j, k) = s_yn[j];
      yhy(k, j) = yhy(j, k) = yn_hy[j];
    }
    sy(k, k) = snyn;
    yhy(k, k) = ynhyn;
    SG_[k] = sn_g;
    YHG_[k] = yn_hg;
  }
  std::memcpy(F_, Fn, sizeof(F_));
  r_.history.push_back(F_[0]);
  ++it_;
  phase_ = kNew;
  return true;
}

template <typename TS, typename TF>
LbfgsResult Functional::minimize_t(double* xi, const LbfgsOptions& o, hipStream_t s) {
  constexpr bool F32 = std::is_same<TF, float>::value;
  TF* df = static_cast<TF*>(df_);
  const int nslot = o.m + 1;
  const Elem e{tab_, tab_ + p_.no, p_.rho0, p_.kT, p_.dv, 2.0 * p_.prefactor / p_.N, n_, p_.no, step_};
  double* x = xi;
  double* xt = static_cast<double*>(work_);
  TS* g = reinterpret_cast<TS*>(xt + n_);
  TS* gt = g + n_;
  TS* pairs = gt + n_;
  std::vector<TS*> S(nslot), Y(nslot);
  for (int j = 0; j < nslot; ++j) {
    S[j] = pairs + 2 * static_cast<long>(j) * n_;
    Y[j] = pairs + (2 * static_cast<long>(j) + 1) * n_;
  }

  // the start point
  if (F32) {
    hipLaunchKernelGGL(to_float_kernel, dim3(nb_), dim3(kFThreads), 0, s, x, xf_, n_);
    MDFT_HIP_CHECK(hipGetLastError());
    excess_(xf_, df, res_ + kFfSlot, s);
  } else {
    excess_(x, df, res_ + kFfSlot, s);
  }
  hipLaunchKernelGGL((local_kernel<TS, TF>), dim3(nb_), dim3(kFThreads), 0, s, x, vext_, df, g, e, part_, nb_);
  MDFT_HIP_CHECK(hipGetLastError());
  reduce_rows(part_, nb_, 0, 4, 3, res_, s);
  const double* h = read_results(s);
  double F[4];
  fparts(h[0], h[1], h[kFfSlot], F);
  LbfgsState st(o);
  st.start(F, h[2], h[3]);

  // each trial: x + t d, the excess evaluation there, then g, the F_id and
  // F_ext sums and the candidate pair's Gram entries (kept if accepted)
  LbfgsState::Trial tr;
  while (st.next(tr)) {
    // the direction reads the k stored pairs, the update pass the pairs from k0
    const int K = tr.k, KU = tr.k - tr.k0;
    DirArgs<TS> da{};
    UpdArgs<TS, TF> ua{};
    for (int j = 0; j < K; ++j) {
      da.s[j] = S[st.order()[j]];
      da.y[j] = Y[st.order()[j]];
      da.cs[j] = tr.cs[j];
      da.cq[j] = tr.cq[j];
    }
    for (int j = 0; j < KU; ++j) {
      ua.s[j] = S[st.order()[tr.k0 + j]];
      ua.y[j] = Y[st.order()[tr.k0 + j]];
    }
    da.x = x;
    da.g = g;
    da.cg = tr.cg;
    da.gamma = tr.gamma;
    da.t = tr.t;
    da.xt = xt;
    da.xf = F32 ? xf_ : nullptr;
    da.sc = S[tr.cand];
    ua.xt = xt;
    ua.V = vext_;
    ua.df = df;
    ua.g = g;
    ua.gt = gt;
    ua.sc = S[tr.cand];
    ua.yn = Y[tr.cand];
    dispatch_k(K, [&](auto kc) {
      hipLaunchKernelGGL((direction_kernel<TS, decltype(kc)::value>), dim3(nb_), dim3(kFThreads), 0, s, da, e);
    });
    MDFT_HIP_CHECK(hipGetLastError());
    if (F32)
      excess_(xf_, df, res_ + kFfSlot, s);
    else
      excess_(xt, df, res_ + kFfSlot, s);
    dispatch_k(KU, [&](auto kc) {
      hipLaunchKernelGGL((update_kernel<TS, TF, decltype(kc)::value>), dim3(nb_), dim3(kFThreads), 0, s, ua, e,
                         part_, nb_);
    });
    MDFT_HIP_CHECK(hipGetLastError());
    reduce_rows(part_, nb_, 0, 5 * KU + 10, 5 * KU + 9, res_, s);
    h = read_results(s);
    double Fn[4];
    fparts(h[5 * KU + 7], h[5 * KU + 8], h[kFfSlot], Fn);
    if (st.judge(Fn, h, KU)) {
      std::swap(x, xt);
      std::swap(g, gt);
    }
  }
  if (x != xi) MDFT_HIP_CHECK(hipMemcpyAsync(xi, x, n_ * sizeof(double), hipMemcpyDeviceToDevice, s));
  MDFT_HIP_CHECK(hipStreamSynchronize(s));
  return st.result();
}

template LbfgsResult Functional::minimize_t<double, double>(double*, const LbfgsOptions&, hipStream_t);
template LbfgsResult Functional::minimize_t<float, double>(double*, const LbfgsOptions&, hipStream_t);
template LbfgsResult Functional::minimize_t<float, float>(double*, const LbfgsOptions&, hipStream_t);

}  // namespace mdft
