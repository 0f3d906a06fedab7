// Projections of the MDFT_MIXED forms as split-BF16 MFMA GEMMs (see mdft_gemm.hip).
#pragma once

#include <memory>
#include <vector>

#include "mdft_proj.hpp"

namespace mdft {

// The dense FP64 projection matrices, (2 np) x no each (rows 2 ip + (re, im)):
// angl2proj p[ip] = sum_io (wa[2ip][io] + i wa[2ip+1][io]) o[io] (the rfft2,
// unscramble and theta quadrature of
// common/MDFT_OrientationProjectionTransform.hpp:267-380 in one matrix), and
// proj2angl o[io] = sum_ip Re p[ip] wp[2ip][io] + Im p[ip] wp[2ip+1][io]
// (:388-484 with the c2r "Re" convention).
void projection_matrices(const AngularGrid& g, std::vector<double>& wa, std::vector<double>& wp);

struct ProjGemm {
  // angl2proj split (data parts, weight parts): (2, 2) BF16x3, three products
  // per term; (3, 3) BF16x6, six products, data exact to FP32 and weights to
  // ~2^-26.  p2a_mfma: proj2angl as a BF16x3 GEMM (needs (2, 2)), else the
  // factorised FP32 VALU kernel.  Grids beyond the MFMA register tiles
  // (2 np > 256 or no > 384) run the factorised FP32 projections instead.
  // compact: the stored projections of the compact channel order (the
  // implied mu = 0, mup < 0 ones left out; compact_codes), both ways; needs
  // the GEMM angl2proj and the GEMM or the compile-time FP32 proj2angl
  ProjGemm(const AngularGrid& g, int data_parts, int weight_parts, bool p2a_mfma, bool compact = false);
  ~ProjGemm();
  ProjGemm(const ProjGemm&) = delete;
  ProjGemm& operator=(const ProjGemm&) = delete;

  // o (or xi when rho0 != nullptr, delta_rho fused) -> p (np, N) complex float;
  // float o / xi (MDFT_MIXED_X6F) needs the (3, 3) split
  void angl2proj(long N, const double* o, cplx<float>* p, const double* rho0, hipStream_t s);
  void angl2proj(long N, const float* o, cplx<float>* p, const double* rho0, hipStream_t s);
  // p -> o; with xi != nullptr get_delta_f runs in the epilogue instead of
  // storing o.  Returns the number of FP64 partials written.
  // colpart: per-row-block column sums (proj2angl_factorized), FP32 VALU
  // proj2angl on compile-time grids only (has_column_sums)
  int proj2angl(long N, const cplx<float>* p, double* o, const double* xi, double* df, double rho0,
                double* partials, hipStream_t s, double* colpart = nullptr);
  // float fields (MDFT_MIXED_X6F): the FP32 VALU kernel (p2a_mfma off)
  int proj2angl(long N, const cplx<float>* p, float* o, const float* xi, float* df, double rho0,
                double* partials, hipStream_t s, double* colpart = nullptr);
  int p2a_blocks(long N) const;
  bool has_column_sums() const;
  bool compact() const;
  // whether a ProjGemm of this grid and split can take the compact order
  static bool compact_possible(const AngularGrid& g, bool p2a_mfma);

 private:
  template <typename TI>
  void angl2proj_t(long N, const TI* o, cplx<float>* p, const double* rho0, hipStream_t s);
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace mdft
