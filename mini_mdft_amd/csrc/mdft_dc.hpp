// The r-mean (q = 0) mode of the c-proj energy in FP64 (see mdft_dc.hip).
#pragma once

#include "mdft_tables.hpp"

#include <hip/hip_runtime.h>

namespace mdft {

// Built once per solver from its angular grid, convolution tables and rho0:
// the FP64 map M from the column sums s(io) = sum_r (xi^2 - 1) to the r-mean
// of vexc (angl2proj of rho0 s, the OZ product at q = 0, proj2angl), stored
// as Qt = (diag(w) M)^T.  Per evaluation: the column sums that proj2angl's
// epilogue wrote per block of 64 rows (colpart[rb][0][io] = sum_r (xi^2 - 1),
// colpart[rb][1][io] = sum_r vexc) are reduced per field, and extra[j][i]
// receives s_i (Q s)_i - w_i s_i V_i / N: summed over i, the term that
// replaces field j's r-mean contribution to the energy sum by the FP64 one.
class MeanModeFp64 {
 public:
  MeanModeFp64(const AngularGrid& g, const ConvTables& t, double rho0);
  ~MeanModeFp64();
  MeanModeFp64(const MeanModeFp64&) = delete;
  MeanModeFp64& operator=(const MeanModeFp64&) = delete;

  // doubles of the work buffer for nb fields of N rows
  static size_t work_doubles(long N, int no, int nb);
  // extra values per field (reduce_partials sums them after the partials)
  int extra_per_field() const;
  // colpart: the proj2angl column sums of nb fields of N rows each (rows of
  // field j in row blocks [j N / 64, (j + 1) N / 64); N a multiple of 64 when
  // nb > 1); work: work_doubles(N, no, nb) doubles; extra: nb extra_per_field()
  void run(const double* colpart, long N, int nb, double* work, double* extra, hipStream_t s) const;

 private:
  int no_;
  double* qt_ = nullptr;  // no x no: Qt[k][i] = w_i M[i][k]
  double* w_ = nullptr;   // no
};

}  // namespace mdft
