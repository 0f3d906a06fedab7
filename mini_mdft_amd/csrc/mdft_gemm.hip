// Projections of the MDFT_MIXED form.
#include "mdft_gemm.hpp"

namespace mdft {

struct ProjGemm::Impl {
  ProjTables tab;
  explicit Impl(const AngularGrid& g) : tab(g) {}
};

ProjGemm::ProjGemm(const AngularGrid& g) : impl_(std::make_unique<Impl>(g)) {}
ProjGemm::~ProjGemm() = default;

void ProjGemm::angl2proj(long N, const double* o, cplx<float>* p, const double* rho0, hipStream_t s) {
  angl2proj_factorized<double, float, float>(impl_->tab, N, o, p, rho0, s);
}

int ProjGemm::proj2angl(long N, const cplx<float>* p, double* o, const double* xi, double* df, double rho0,
                        double* partials, hipStream_t s) {
  proj2angl_factorized<float, double, float>(impl_->tab, N, p, o, xi, df, rho0, partials, s);
  return proj2angl_blocks(impl_->tab, N);
}

int ProjGemm::p2a_blocks(long N) const { return proj2angl_blocks(impl_->tab, N); }

}  // namespace mdft
