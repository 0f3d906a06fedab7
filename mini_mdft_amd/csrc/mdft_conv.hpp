// Convolution::execute on gfx950 (see mdft_conv.hip).
#pragma once

#include <memory>

#include "mdft_fft.hpp"

namespace mdft {

struct ConvDevice {
  ConvTables tab;
  Fft3dPlan fft;
  long npairs = 0;
  double* kxyz = nullptr;     // kx | ky | kz (FP64)
  int* pairs = nullptr;       // packed ix | iy << 10 | iz << 20, (ix, iy, iz)-ordered
  int* ia_start = nullptr;    // np + 1: first OZ term of each output projection
  int* ipinfo = nullptr;      // np: m | (mup + mmax) << 8 | mu2 << 16
  double2* ctab64 = nullptr;  // (nq, np_new): the OZ table, transposed
  float2* ctab32 = nullptr;
  double* rcoef = nullptr;    // rotation recursion coefficients a | b | c | d
  ConvDevice(int nx, int ny, int nz, double lx, double ly, double lz, int mmax, int mrso, const LucTable& luc);
  ~ConvDevice();
  ConvDevice(const ConvDevice&) = delete;
  ConvDevice& operator=(const ConvDevice&) = delete;
  // LDS bytes of one OZ block for elements of elem bytes (float or double)
  size_t oz_lds_bytes(size_t elem) const;
  // per-pair real rotation matrices d^l(theta_q) in T, built on first use
  // by an FP64 set-up kernel on stream s (npairs x nR entries)
  template <typename T>
  const T* dtable(hipStream_t s) const;
  mutable void* dtab64 = nullptr;
  mutable void* dtab32 = nullptr;
  // per-pair set-up values of the FP32 compile-time OZ kernel (packed index,
  // c(q) row, interpolation weights; and in *phases the phases of mup =
  // 0..mmax), built on first use on stream s
  template <typename T>
  const void* pair_meta(hipStream_t s, const void** phases) const;
  mutable void* pmeta32 = nullptr;
  mutable void* pphase32 = nullptr;
  mutable void* pmeta64 = nullptr;
  mutable void* pphase64 = nullptr;
  mutable int* dmap = nullptr;  // per-thread records: LDS places of the half d table's entries, ipinfo
  // stars of the pair list (pairs whose q differ only in component signs,
  // which share |q| and so the c(q) row): four pair indices per star, -1 pads
  int* stars = nullptr;
  long nstars = 0;
  mutable bool stars_checked32 = false, stars_checked64 = false;
  // whether the star kernel serves this grid in FP32 / FP64 (then the d
  // table holds one row per star)
  bool star32 = false, star64 = false;
  mutable int* dent = nullptr;  // LDS places codes of the half d table's entries (star kernel)
  mutable int* ccode = nullptr;  // per projection: its compact channel (compact_codes)
  // row width of the padded d layout when the compile-time OZ kernel serves
  // this (mmax, mrso) (2 mmax + 1), else 0 (compact rows, generic kernel)
  int padw = 0;

  // the per-(q, -q) rotation / OZ / inverse-rotation pass on FFT'd data;
  // gamma components below chop_eps in magnitude are set to 0.  With
  // nbatch > 1, p holds nbatch fields as (np, nbatch, nx, ny, nz).  With
  // events, the launch records its own start and end times in ev0 and ev1.
  // With compact, p holds the stored channels of the compact channel order
  // (compact_projections; the star kernel only, compact_supported):
  // implied inputs are read from their partners, and each partner's channel
  // receives the pair's combination g_+ + (-1)^mup conj(g_- at -q).
  template <typename T>
  void oz_pass(cplx<T>* p, double chop_eps, hipStream_t s, int nbatch = 1, hipEvent_t ev0 = nullptr,
               hipEvent_t ev1 = nullptr, bool compact = false) const;
  // FFT, oz_pass, inverse FFT: in place on (np, nx, ny, nz) (or (np, nbatch,
  // nx, ny, nz)); marks the FFT / OZ stage boundaries on `timer` when given
  template <typename T>
  void execute(cplx<T>* p, double chop_eps, hipStream_t s, StageTimer* timer = nullptr, int nbatch = 1,
               bool compact = false) const;
  // whether oz_pass takes the compact channel order for FP64 (fp64) or FP32
  // projections
  bool compact_supported(bool fp64) const;
};

// (ix, iy, iz) triples of the reference's (q, -q) list with the same q / -q
// designation, sorted by (ix, iy, iz) so that consecutive pairs are
// neighbours in memory.
std::vector<int> qpairs_memory_order(int nx, int ny, int nz);

// The stars of a pair list given as packed ix | iy << 10 | iz << 20: the
// pairs whose (ix, iy, iz) fold to the same (|ix|, |iy|, |iz|) (i folds to
// min(i, n - i)), as four indices into the list per star (list order within
// a star, -1 pads), the stars ordered by (|ix|, |iy|, |iz|).
std::vector<int> oz_stars(const std::vector<int>& packed, int nx, int ny, int nz);

}  // namespace mdft
