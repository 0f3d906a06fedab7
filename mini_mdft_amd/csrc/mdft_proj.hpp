// Orientation <-> projection transforms and the solvent reductions (see
// mdft_proj.hip).
#pragma once

#include <vector>

#include "mdft_device.hpp"

namespace mdft {

// Device copies of the angular tables one transform needs, in FP64 and FP32.
struct ProjTables {
  int mmax, mrso, ntheta, nphi, npsi, nmu2, no, np;
  // e^{+2 pi i t/nphi} (t < nphi), e^{+2 pi i mu2 k/npsi} (nmu2 x npsi)
  double2* ephi64 = nullptr;
  double2* epsi64 = nullptr;
  float2* ephi32 = nullptr;
  float2* epsi32 = nullptr;
  // forward weights f_m w_theta d / (nphi npsi) and backward weights d f_m,
  // both ntheta x np
  double* wfwd64 = nullptr;
  double* wbwd64 = nullptr;
  float* wfwd32 = nullptr;
  float* wbwd32 = nullptr;
  int* ipinfo = nullptr;  // np: (mup + mmax) | mu2 << 8 | m << 16
  int* yidx = nullptr;    // (2mmax+1) x nmu2 x (mmax+1): ip of (m, mup, mu2) or -1
  double* w = nullptr;    // no: quadrature weights w(io)
  explicit ProjTables(const AngularGrid& g);
  ~ProjTables();
  ProjTables(const ProjTables&) = delete;
  ProjTables& operator=(const ProjTables&) = delete;
};

// The compact channel order of the fused evaluation.  The fields of the path
// are real, so a mu = 0 projection with mup < 0 is implied by its (m, -mup, 0)
// partner: p^m_{-mup,0} = (-1)^mup conj(p^m_{mup,0}) in r space, and after the
// forward FFT p^m_{-mup,0}(q) = (-1)^mup conj(p^m_{mup,0}(-q)).  angl2proj and
// the forward FFT hold the other projections only ("stored", in their
// canonical order: m, mup, mu ascending; at mmax 5, mrso 2, 73 of 88), and
// the OZ pass reads an implied projection from its partner.  On the way back
// proj2angl uses only the real part of the mu = 0 column, to which the pair
// (m, +-mup, 0) contributes through c = p^m_{mup,0} + (-1)^mup conj(p^m_{-mup,0})
// alone (the weights obey d^m_{-mup,0} = (-1)^mup d^m_{mup,0}); so the OZ pass
// stores that combination, c(q) = g_+(q) + (-1)^mup conj(g_-(-q)), in the
// partner's channel, the inverse FFT transforms the stored channels, and
// proj2angl takes the implied sums as zero.  Nothing is assumed of c(q).
__host__ __device__ constexpr bool implied_projection(int mup, int mu) { return mu == 0 && mup < 0; }
// stored projections, and the stored channel of canonical projection ip (an
// implied one: its partner's), for compile-time grids
template <int MMAX, int MRSO>
__host__ __device__ constexpr int compact_count() {
  int c = 0;
  for (int m = 0; m <= MMAX; ++m)
    for (int mup = -m; mup <= m; ++mup)
      for (int mu = 0; mu <= m; mu += MRSO) c += implied_projection(mup, mu) ? 0 : 1;
  return c;
}
template <int MMAX, int MRSO>
__host__ __device__ constexpr int compact_index(int ip) {
  int c = 0, i = 0;
  for (int m = 0; m <= MMAX; ++m)
    for (int mup = -m; mup <= m; ++mup)
      for (int mu = 0; mu <= m; mu += MRSO, ++i) {
        if (i == ip && !implied_projection(mup, mu)) return c;
        if (i == ip) return compact_index<MMAX, MRSO>(ip - 2 * mup * (m / MRSO + 1));
        c += implied_projection(mup, mu) ? 0 : 1;
      }
  return -1;
}
// the compact channel of a stored canonical projection ip = (m, mup, mu2):
// ip less the implied projections before it (m (m - 1) / 2 in the blocks of
// smaller m, and those of its own m block that precede it)
__host__ __device__ constexpr int compact_of(int ip, int m, int mup, int mu2) {
  return ip - m * (m - 1) / 2 - (mup >= 0 ? m : mup + m + (mu2 > 0 ? 1 : 0));
}
// the same as a table, for kernels that index it with unrolled loop counters
template <int MMAX, int MRSO>
struct CompactMap {
  int v[(MMAX + 1) * (MMAX + 1) * (MMAX + 2)];
  __host__ __device__ constexpr CompactMap() : v() {
    for (int i = 0; i < (MMAX + 1) * (MMAX + 1) * (MMAX + 2); ++i) v[i] = compact_index<MMAX, MRSO>(i);
  }
};
// host: the number of stored projections of a grid, and per canonical
// projection a code: bits 0-7 the stored channel it reads (an implied
// projection: its partner's), bit 16 implied, bit 17 mup odd
constexpr int kCompactImplied = 1 << 16, kCompactOdd = 1 << 17;
int compact_projections(const AngularGrid& g);
std::vector<int> compact_codes(const AngularGrid& g);

// angl2proj (common/MDFT_OrientationProjectionTransform.hpp:267-380).  With
// drho_rho0 != nullptr the input is xi and delta_rho = rho0 (xi^2 - 1) is
// formed on load (common/MDFT_Solvent.hpp:575-596 fused).
// With compact, only the stored projections of the compact channel order are
// written, at their compact channels (compile-time grids, angl2proj_compiled).
template <typename TI, typename TP, typename TC>
void angl2proj_factorized(const ProjTables& t, long N, const TI* o, cplx<TP>* p, const double* drho_rho0,
                          hipStream_t s, bool compact = false);
// whether angl2proj_factorized has a compile-time kernel for this grid with
// FP64 (fp64) or FP32 arithmetic
bool angl2proj_compiled(const ProjTables& t, bool fp64);

// proj2angl (:388-484).  With xi != nullptr the orientation output is not
// stored: get_delta_f (common/MDFT_Solvent.hpp:613-660) runs in the epilogue,
// writing df and one FP64 partial sum per block into `partials`.
// With colpart != nullptr (and xi, compile-time angular grids only) the
// epilogue also writes, per block of 64 rows rb, the column sums over its
// rows: colpart[rb][0][io] = sum_r (xi^2 - 1), colpart[rb][1][io] = sum_r vexc.
// With compact, p holds the stored channels of the compact channel order
// (compile-time grids only).
template <typename TP, typename TO, typename TC>
void proj2angl_factorized(const ProjTables& t, long N, const cplx<TP>* p, TO* o, const TO* xi, TO* df,
                          double rho0, double* partials, hipStream_t s, double* colpart = nullptr,
                          bool compact = false);
// whether proj2angl_factorized can write column sums for this grid, and the
// row blocks (64 rows each) it writes them for
bool proj2angl_has_column_sums(const ProjTables& t);
// whether proj2angl has a compile-time kernel for (mmax, mrso)
bool proj2angl_compiled(int mmax, int mrso);
long proj2angl_row_blocks(long N);
// partial sums written by the fused proj2angl over N points (one per row
// block, or per (row block, theta)); for N a multiple of 64 the partials of
// point ranges [j N, (j + 1) N) of a batch are contiguous
int proj2angl_blocks(const ProjTables& t, long N);

template <typename T>
void delta_rho(const T* xi, T* out, size_t n, double rho0, hipStream_t s);

// standalone get_delta_f; returns the number of partials written
template <typename T>
int delta_f(const T* xi, const T* vexc, const double* w, T* df, long nxyz, int no, double rho0, double* partials,
            int max_partials, hipStream_t s);
int delta_f_blocks(long nxyz);
// grid points (rows of no values) per delta_f block: one partial sum each
int delta_f_rows_per_block();

// ff = scale * sum(partials[0..count)), summed in a fixed order, into out[0];
// with nbatch > 1, field j's count partials follow field j - 1's and its sum
// goes to out[j out_stride]
// extra (device, nbatch x nextra values, may be null): field j's nextra
// values are added to its sum, before the scaling
void reduce_partials(const double* partials, int count, double scale, double* out, hipStream_t s, int nbatch = 1,
                     long out_stride = 1, const double* extra = nullptr, int nextra = 0);

template <typename T>
void fill_uniform(T* x, size_t n, unsigned long long seed, unsigned long long offset, double lo, double hi,
                  hipStream_t s);

}  // namespace mdft
