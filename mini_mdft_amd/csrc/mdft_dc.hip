// The r-mean (q = 0) mode of the c-proj energy in FP64, for the mixed forms.
//
// ff = rho0 prefactor sum_{r, io} w (xi^2 - 1) vexc (common/MDFT_Solvent.hpp:
// 613-660).  Split both factors into their mean over r and the rest: with
// s(io) = sum_r (xi^2 - 1) and V(io) = sum_r vexc,
//   sum_r (xi^2 - 1) vexc = s V / N + sum_r (xi^2 - 1)' vexc'
// exactly (the primed parts have zero r-mean).  The convolution maps the
// r-mean of its input to the r-mean of its output alone (q = 0 couples only to
// itself), so the first term is the q = 0 mode of the whole evaluation:
// V / N = proj2angl(OZ_{q=0}(angl2proj(rho0 s))) (the unnormalised forward
// FFT turns the r-mean into the sum over r, and the backward FFT adds the
// q = 0 value to every r).  With xi = 1 + small noise this one mode carries
// almost all of ff (99% at 64^3 for the benchmark's xi), so an FP32 pipeline's
// relative error on it, ~1e-7, is ~1e-7 of ff.  Here the mode is recomputed
// in FP64 from s and replaces s V / N: extra = sum_io w s (vdc - V / N) is
// added to the energy sum.  The rest of ff, and df, stay as the mixed path
// computed them; any error of that path that is constant over r (its own
// q = 0 value, rounding noise that leaks into the r-mean) drops out with V.
//
// The whole q = 0 route is linear in s, so it is built once per solver as an
// FP64 matrix (MeanModeFp64's constructor; its OZ step follows
// common/MDFT_Convolution.hpp:414-541 at q = 0: rotation with R(q = 0), the lab
// frame, c(|q|) at the table's first row (alpha = 0, :449-464), the OZ sum
// (:466-497), the inverse rotation, and the store rule at q = -q: the -q
// branch is stored, conjugated iff some side has n = 1, index 0 = n/2).  Per
// evaluation the column sums are reduced (fixed order) and one matrix-vector
// product gives the terms of extra, which the energy reduction adds.
#include "mdft_dc.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

#include "mdft_device.hpp"
#include "mdft_gemm.hpp"

namespace mdft {

namespace {

// One reduction pass over the per-row-block column sums: out[j][c][col] =
// sum of in[j][kFan c + i][col] over i < kFan (fixed order; rows past the
// field's count read as zero).  One wave per (chunk c, 64 columns, field j),
// all kFan loads of a thread in flight together.
constexpr int kFan = 64;
__global__ __launch_bounds__(64) void mean_colsum_kernel(const double* __restrict__ in, long nin, int width,
                                                         double* __restrict__ out, long nout) {
  const long c = blockIdx.x;
  const int col = blockIdx.y * 64 + threadIdx.x;
  const long j = blockIdx.z;
  if (col >= width) return;
  const double* src = in + j * nin * width;
  double v[kFan];
#pragma unroll
  for (int i = 0; i < kFan; ++i) {
    const long row = c * kFan + i;
    v[i] = row < nin ? src[row * width + col] : 0.0;
  }
#pragma unroll
  for (int h = kFan / 2; h > 0; h >>= 1)
#pragma unroll
    for (int i = 0; i < h; ++i) v[i] += v[i + h];
  out[(j * nout + c) * width + col] = v[0];
}

// extra[j][b][i] = s_i sum_{k in chunk b} Qt[k][i] s_k for the kKch-wide
// chunks b of k, and extra[j][nkb][i] = -w_i s_i V_i / N: summed over b and i,
// s (Q s) - w s V / N, with (s, V) the field's column sums (the last
// reduction pass left one row of 2 no values) and Qt = (diag(w) M)^T, M the
// FP64 map s -> r-mean of vexc.  One thread per (i, chunk b, field j), all
// kKch loads of a thread in flight together.
constexpr int kKch = 24;
__global__ __launch_bounds__(64) void mean_gemv_kernel(const double* __restrict__ sv, const double* __restrict__ qt,
                                                       const double* __restrict__ w, int no, int nkb, double invN,
                                                       double* __restrict__ extra) {
  const int i = blockIdx.x * 64 + threadIdx.x, b = blockIdx.y, j = blockIdx.z;
  if (i >= no) return;
  const double* s = sv + static_cast<long>(j) * 2 * no;
  double* out = extra + static_cast<long>(j) * (nkb + 1) * no;
  const double si = s[i];
  if (b == nkb) {
    out[static_cast<long>(nkb) * no + i] = -w[i] * si * s[no + i] * invN;
    return;
  }
  double v[kKch];
#pragma unroll
  for (int q = 0; q < kKch; ++q) {
    const int k = b * kKch + q;
    v[q] = k < no ? qt[static_cast<long>(k) * no + i] * s[k] : 0.0;
  }
#pragma unroll
  for (int h = 16; h > 0; h >>= 1)
#pragma unroll
    for (int q = 0; q < h; ++q)
      if (q + h < kKch) v[q] += v[q + h];
  out[static_cast<long>(b) * no + i] = si * v[0];
}

long row_blocks_per_field(long N, int nb) { return nb == 1 ? (N + 63) / 64 : N / 64; }
long chunks_after(long n) { return (n + kFan - 1) / kFan; }

template <typename T>
T* upload_vec(const std::vector<T>& h) {
  T* d = nullptr;
  MDFT_HIP_CHECK(hipMalloc(&d, std::max<size_t>(1, h.size()) * sizeof(T)));
  if (!h.empty()) MDFT_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

}  // namespace

// The map s -> r-mean of vexc, one column per unit vector s = e_i (every step
// is linear in s over the reals): pb = rho0 angl2proj(e_i) (column i of the
// dense matrix), the q = 0 rotation with R(0), the OZ sum with the first c(q)
// row, the inverse rotation of the -q branch and its store rule, and
// proj2angl.  The reference's chop of gamma (an absolute threshold of
// DBL_EPSILON, common/MDFT_Math_Utils.hpp:126-137) is not applied: it is not
// linear, and at q = 0 it would only zero components that are rounding noise
// of exact zeros (~1e-17 against values of order 1).
MeanModeFp64::MeanModeFp64(const AngularGrid& g, const ConvTables& t, double rho0) : no_(g.no) {
  throw_if(g.mmax > 6, "mean mode: mmax above 6");
  const int no = g.no, np = g.np, mmax = g.mmax, mrso = g.mrso, W = 2 * mmax + 1;
  const bool singular = t.nx == 1 || t.ny == 1 || t.nz == 1;
  std::vector<double> wa, wp;
  projection_matrices(g, wa, wp);
  std::vector<cd> R(static_cast<size_t>(mmax + 1) * W * W, cd(0.0, 0.0));
  const double q0[3] = {0.0, 0.0, 0.0};
  rotation_matrix(mmax, q0, R.data());
  auto Rq = [&](int l, int a, int b) { return R[(l * W + a + mmax) * W + b + mmax]; };
  auto c0 = [&](int ia) { return t.ctab[static_cast<size_t>(ia) * t.nq]; };
  std::vector<double> M(static_cast<size_t>(no) * no);
  std::vector<cd> pb(np), dq(np), dmq(np), gm(np), y0(np);
  for (int i = 0; i < no; ++i) {
    for (int ip = 0; ip < np; ++ip)
      pb[ip] = rho0 * cd(wa[static_cast<size_t>(2 * ip) * no + i], wa[static_cast<size_t>(2 * ip + 1) * no + i]);
    for (int ip = 0; ip < np; ++ip) {
      const int m = g.m[ip], chi = g.mup[ip], mu2 = g.mu2[ip];
      cd u(0.0, 0.0), v(0.0, 0.0);
      for (int mup = -m; mup <= m; ++mup) {
        const cd x = pb[g.P(m, mup, mu2)];
        u += x * Rq(m, mup, chi);
        v += x * Rq(m, mup, -chi);
      }
      dq[ip] = u;
      dmq[ip] = (m & 1) ? -v : v;
    }
    for (int ip = 0; ip < np; ++ip) {
      const int chi = g.mup[ip], achi = std::abs(chi);
      int ia = t.ia_start[ip];
      cd s1(0.0, 0.0);  // gamma of the -q branch
      for (int n = achi; n <= mmax; ++n)
        for (int nu2 = -n / mrso; nu2 <= n / mrso; ++nu2, ++ia)
          s1 += nu2 < 0 ? c0(ia) * dmq[g.P(n, chi, -nu2)] : c0(ia) * std::conj(dq[g.P(n, chi, nu2)]);
      gm[ip] = s1;
    }
    for (int ip = 0; ip < np; ++ip) {
      const int m = g.m[ip], mup = g.mup[ip], mu2 = g.mu2[ip];
      cd v(0.0, 0.0);
      for (int chi = -m; chi <= m; ++chi) v += gm[g.P(m, chi, mu2)] * std::conj(Rq(m, mup, -chi));
      if (m & 1) v = -v;
      y0[ip] = singular ? std::conj(v) : v;
    }
    for (int io = 0; io < no; ++io) {
      double acc = 0.0;
      for (int ip = 0; ip < np; ++ip)
        acc += y0[ip].real() * wp[static_cast<size_t>(2 * ip) * no + io] +
               y0[ip].imag() * wp[static_cast<size_t>(2 * ip + 1) * no + io];
      M[static_cast<size_t>(io) * no + i] = acc;
    }
  }
  // Qt[k][i] = w_i M[i][k]
  std::vector<double> qt(static_cast<size_t>(no) * no);
  for (int i = 0; i < no; ++i)
    for (int k = 0; k < no; ++k) qt[static_cast<size_t>(k) * no + i] = g.w[i] * M[static_cast<size_t>(i) * no + k];
  qt_ = upload_vec(qt);
  w_ = upload_vec(g.w);
}

MeanModeFp64::~MeanModeFp64() {
  if (qt_) (void)hipFree(qt_);
  if (w_) (void)hipFree(w_);
}

size_t MeanModeFp64::work_doubles(long N, int no, int nb) {
  // two ping-pong buffers of the reduction passes
  const long first = chunks_after(row_blocks_per_field(N, nb));
  return 2 * static_cast<size_t>(first) * nb * 2 * no;
}

int MeanModeFp64::extra_per_field() const { return ((no_ + kKch - 1) / kKch + 1) * no_; }

void MeanModeFp64::run(const double* colpart, long N, int nb, double* work, double* extra, hipStream_t s) const {
  throw_if(nb > 1 && N % 64 != 0, "mean mode: batched fields need N to be a multiple of 64");
  const int width = 2 * no_;
  long n = row_blocks_per_field(N, nb);
  const long first = chunks_after(n);
  double* buf[2] = {work, work + static_cast<size_t>(first) * nb * width};
  const double* in = colpart;
  int k = 0;
  do {  // reduction passes down to one row of column sums per field
    const long m = chunks_after(n);
    hipLaunchKernelGGL(mean_colsum_kernel, dim3(static_cast<unsigned>(m), (width + 63) / 64, nb), dim3(64), 0, s,
                       in, n, width, buf[k], m);
    MDFT_HIP_CHECK(hipGetLastError());
    in = buf[k];
    k ^= 1;
    n = m;
  } while (n > 1);
  const int nkb = (no_ + kKch - 1) / kKch;
  hipLaunchKernelGGL(mean_gemv_kernel, dim3((no_ + 63) / 64, nkb + 1, nb), dim3(64), 0, s, in, qt_, w_, no_, nkb,
                     1.0 / static_cast<double>(N), extra);
  MDFT_HIP_CHECK(hipGetLastError());
}

}  // namespace mdft
