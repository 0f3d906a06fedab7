// Projections of the MDFT_MIXED forms as split-BF16 MFMA GEMMs (see mdft_gemm.hip).
#pragma once

#include <memory>

#include "mdft_proj.hpp"

namespace mdft {

struct ProjGemm {
  // angl2proj split (data parts, weight parts): (2, 2) BF16x3, three products
  // per term; (3, 3) BF16x6, six products, data exact to FP32 and weights to
  // ~2^-26.  p2a_mfma: proj2angl as a BF16x3 GEMM (needs (2, 2)), else the
  // factorised FP32 VALU kernel.  Grids beyond the MFMA register tiles
  // (2 np > 256 or no > 384) run the factorised FP32 projections instead.
  ProjGemm(const AngularGrid& g, int data_parts, int weight_parts, bool p2a_mfma);
  ~ProjGemm();
  ProjGemm(const ProjGemm&) = delete;
  ProjGemm& operator=(const ProjGemm&) = delete;

  // o (or xi when rho0 != nullptr, delta_rho fused) -> p (np, N) complex float
  void angl2proj(long N, const double* o, cplx<float>* p, const double* rho0, hipStream_t s);
  // p -> o; with xi != nullptr get_delta_f runs in the epilogue instead of
  // storing o.  Returns the number of FP64 partials written.
  int proj2angl(long N, const cplx<float>* p, double* o, const double* xi, double* df, double rho0,
                double* partials, hipStream_t s);
  int p2a_blocks(long N) const;

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace mdft
