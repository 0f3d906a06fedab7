// Orientation <-> projection transforms and the solvent reductions (see
// mdft_proj.hip).
#pragma once

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

// angl2proj (common/MDFT_OrientationProjectionTransform.hpp:267-380).  With
// drho_rho0 != nullptr the input is xi and delta_rho = rho0 (xi^2 - 1) is
// formed on load (common/MDFT_Solvent.hpp:575-596 fused).
template <typename TI, typename TP, typename TC>
void angl2proj_factorized(const ProjTables& t, long N, const TI* o, cplx<TP>* p, const double* drho_rho0,
                          hipStream_t s);

// proj2angl (:388-484).  With xi != nullptr the orientation output is not
// stored: get_delta_f (common/MDFT_Solvent.hpp:613-660) runs in the epilogue,
// writing df and one FP64 partial sum per block into `partials`.
// With colpart != nullptr (and xi, compile-time angular grids only) the
// epilogue also writes, per block of 64 rows rb, the column sums over its
// rows: colpart[rb][0][io] = sum_r (xi^2 - 1), colpart[rb][1][io] = sum_r vexc.
template <typename TP, typename TO, typename TC>
void proj2angl_factorized(const ProjTables& t, long N, const cplx<TP>* p, TO* o, const TO* xi, TO* df,
                          double rho0, double* partials, hipStream_t s, double* colpart = nullptr);
// whether proj2angl_factorized can write column sums for this grid, and the
// row blocks (64 rows each) it writes them for
bool proj2angl_has_column_sums(const ProjTables& t);
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
