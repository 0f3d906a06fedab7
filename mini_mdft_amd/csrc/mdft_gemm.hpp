// Projections of the MDFT_MIXED form (see mdft_gemm.hip).
#pragma once

#include <memory>

#include "mdft_proj.hpp"

namespace mdft {

struct ProjGemm {
  explicit ProjGemm(const AngularGrid& g);
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
