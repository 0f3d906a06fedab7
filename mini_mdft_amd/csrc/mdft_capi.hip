// The C ABI of include/mdft_hip.h: handles, argument checks, error capture
// and dispatch by precision form.  Exceptions never cross the boundary; every
// failure becomes a nonzero status plus the thread's last-error text (the
// reference raises std::runtime_error via Impl::Throw_If,
// common/MDFT_Asserts.hpp:15-28).
#include <cfloat>
#include <cstring>
#include <algorithm>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/mdft_hip.h"
#include "mdft_conv.hpp"
#include "mdft_dc.hpp"
#include "mdft_engine.hpp"
#include "mdft_functional.hpp"
#include "mdft_gemm.hpp"
#include "mdft_proj.hpp"

using namespace mdft;

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
  } catch (...) {
    g_last_error = "unknown error";
  }
  return 1;
}

void check_precision(int precision) {
  throw_if(precision != MDFT_FP64 && precision != MDFT_FP32 && precision != MDFT_MIXED &&
               precision != MDFT_MIXED_X6 && precision != MDFT_MIXED_X6F,
           "unknown precision form");
}

bool is_mixed(int precision) {
  return precision == MDFT_MIXED || precision == MDFT_MIXED_X6 || precision == MDFT_MIXED_X6F;
}

std::unique_ptr<ProjGemm> make_gemm(const AngularGrid& g, int precision, bool compact = false) {
  if (!is_mixed(precision)) return nullptr;
  return precision == MDFT_MIXED ? std::make_unique<ProjGemm>(g, 2, 2, true, compact)
                                 : std::make_unique<ProjGemm>(g, 3, 3, false, compact);
}

void check_dtype(int dtype) { throw_if(dtype != MDFT_DOUBLE && dtype != MDFT_FLOAT, "unknown element type"); }

// Chop threshold of gamma (common/MDFT_Convolution.hpp:373, :514-517): the
// epsilon of the reference's ScalarType, i.e. DBL_EPSILON for MDFT_FP64 and
// FLT_EPSILON for MDFT_FP32.  MDFT_MIXED approximates the FP64 form and keeps
// its threshold.
double chop_eps_of(int precision) { return precision == MDFT_FP32 ? FLT_EPSILON : DBL_EPSILON; }

}  // namespace

struct mdft_opt_s {
  int nx, ny, nz, precision;
  long N;
  AngularGrid grid;
  std::unique_ptr<ProjTables> tab;
  std::unique_ptr<ProjGemm> gemm;  // MDFT_MIXED only
};

struct mdft_conv_s {
  int precision;
  std::unique_ptr<ConvDevice> dev;
};

struct mdft_solver_s {
  mdft_conv_s* conv;  // borrowed
  int precision;
  double rho0, prefactor;
  long N;
  int capacity = 1;       // fields one call may evaluate together
  int field_parts = 0;    // partial sums per field
  std::unique_ptr<ProjTables> tab;
  std::unique_ptr<ProjGemm> gemm;
  // the fused evaluation in the compact channel order (the implied mu = 0,
  // mup < 0 projections neither stored nor transformed; every form where the
  // grid allows, on by default)
  bool compact = false;
  void* pbuf = nullptr;  // projections, np x N complex (compact: the first channels)
  double* partials = nullptr;
  int nparts = 0;
  double* ffdev = nullptr;
  size_t bytes = 0;
  std::unique_ptr<StageTimer> timer;  // set while stage timing is on
  // the r-mean (q = 0) mode of ff in FP64 (on by default for MDFT_MIXED_X6)
  bool mean_on = false;
  std::unique_ptr<MeanModeFp64> mean;
  double* colpart = nullptr;  // proj2angl column sums: row blocks x 2 no
  double* meanwork = nullptr;
  double* extra = nullptr;    // capacity x no terms
  size_t mean_bytes = 0;
  ~mdft_solver_s() {
    if (pbuf) (void)hipFree(pbuf);
    if (partials) (void)hipFree(partials);
    if (ffdev) (void)hipFree(ffdev);
    if (colpart) (void)hipFree(colpart);
    if (meanwork) (void)hipFree(meanwork);
    if (extra) (void)hipFree(extra);
  }
};

namespace {

// whether the fused evaluation of this form can take the compact channel
// order: the GEMM angl2proj (mixed forms) or the compile-time angl2proj
// (FP64, FP32), the GEMM or compile-time proj2angl and the star OZ kernel
bool compact_available(int precision, const mdft_conv_s* conv, const ProjTables& tab) {
  const AngularGrid& g = conv->dev->tab.ang;
  if (compact_projections(g) == g.np) return false;
  switch (precision) {
    case MDFT_MIXED:
    case MDFT_MIXED_X6:
    case MDFT_MIXED_X6F:
      return ProjGemm::compact_possible(g, precision == MDFT_MIXED) && conv->dev->compact_supported(false);
    case MDFT_FP64:
    case MDFT_FP32:
      return angl2proj_compiled(tab, precision == MDFT_FP64) && proj2angl_compiled(g.mmax, g.mrso) &&
             conv->dev->compact_supported(precision == MDFT_FP64);
    default:
      return false;
  }
}

// whether this solver's proj2angl can write the column sums of the mean mode
bool mean_mode_supported(const mdft_solver_s* h) {
  if (h->precision == MDFT_MIXED) return false;  // proj2angl on the matrix cores
  if (h->gemm) return h->gemm->has_column_sums();
  return proj2angl_has_column_sums(*h->tab);
}

void enable_mean_mode(mdft_solver_s* h) {
  if (h->mean) return;
  const ConvTables& t = h->conv->dev->tab;
  h->mean = std::make_unique<MeanModeFp64>(t.ang, t, h->rho0);
  const size_t ncol = static_cast<size_t>(proj2angl_row_blocks(h->N * h->capacity)) * 2 * t.ang.no;
  const size_t nwork = MeanModeFp64::work_doubles(h->N, t.ang.no, h->capacity);
  MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&h->colpart), ncol * sizeof(double)));
  MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&h->meanwork), nwork * sizeof(double)));
  const size_t nextra = static_cast<size_t>(h->capacity) * h->mean->extra_per_field();
  MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&h->extra), nextra * sizeof(double)));
  h->mean_bytes = (ncol + nwork + nextra) * sizeof(double);
  h->bytes += h->mean_bytes;
}

}  // namespace

extern "C" {

const char* mdft_last_error(void) { return g_last_error.c_str(); }

int mdft_version(void) { return 1; }

int mdft_gauss_legendre(int n, double* x, double* w) {
  return guarded([&] {
    throw_if(n <= 0 || !x || !w, "gauss_legendre: bad arguments");
    gauss_legendre(n, x, w);
  });
}

int mdft_wigner_small_d(int m, int mu, int mup, double theta, double* out) {
  return guarded([&] {
    throw_if(!out, "null output");
    *out = wigner_small_d(m, mu, mup, theta);
  });
}

int mdft_angular_sizes(int mmax, int mrso, int* sizes) {
  return guarded([&] {
    throw_if(!sizes, "null output");
    AngularGrid g = make_angular_grid(mmax, mrso);
    const int v[6] = {g.ntheta, g.nphi, g.npsi, g.no, g.np, g.nmu2};
    std::memcpy(sizes, v, sizeof(v));
  });
}

int mdft_angular_tables(int mmax, int mrso, double* theta, double* wtheta, double* w, int* pmap, int* m, int* mup,
                        int* mu, int* mu2, double* wigner_d) {
  return guarded([&] {
    AngularGrid g = make_angular_grid(mmax, mrso);
    auto put = [](auto* dst, const auto& src) {
      if (dst) std::copy(src.begin(), src.end(), dst);
    };
    put(theta, g.theta);
    put(wtheta, g.wtheta);
    put(w, g.w);
    put(pmap, g.pmap);
    put(m, g.m);
    put(mup, g.mup);
    put(mu, g.mu);
    put(mu2, g.mu2);
    put(wigner_d, g.wigner_d);
  });
}

int mdft_kproj(int n, double length, double* k) {
  return guarded([&] {
    throw_if(n <= 0 || !k, "kproj: bad arguments");
    auto v = kproj(n, length);
    std::copy(v.begin(), v.end(), k);
  });
}

int mdft_rotation_matrix(int mmax, const double* q, double* R) {
  return guarded([&] {
    throw_if(mmax < 0 || mmax > 6 || !q || !R, "rotation_matrix: bad arguments (0 <= mmax <= 6)");
    rotation_matrix(mmax, q, reinterpret_cast<cd*>(R));
  });
}

int mdft_qpair_count(int nx, int ny, int nz) {
  int out = -1;
  guarded([&] {
    throw_if(nx <= 0 || ny <= 0 || nz <= 0, "grid sizes must be positive");
    out = static_cast<int>(qpairs_reference(nx, ny, nz).size() / 3);
  });
  return out;
}

int mdft_qpairs(int nx, int ny, int nz, int* ixyz) {
  return guarded([&] {
    throw_if(nx <= 0 || ny <= 0 || nz <= 0 || !ixyz, "qpairs: bad arguments");
    auto v = qpairs_reference(nx, ny, nz);
    std::copy(v.begin(), v.end(), ixyz);
  });
}

int mdft_luc_read(const char* path, int np_luc, double qmaxwanted, int* nq, double* dq, int* tuples, double* c) {
  return guarded([&] {
    throw_if(!path || np_luc <= 0, "luc_read: bad arguments");
    LucTable t = read_luc(path, np_luc, qmaxwanted);
    if (nq) *nq = t.nq;
    if (dq) *dq = t.dq;
    if (tuples) std::copy(t.tuples.begin(), t.tuples.end(), tuples);
    if (c) std::memcpy(c, t.c.data(), t.c.size() * sizeof(cd));
  });
}

// ---------------------------------------------------------------------------
int mdft_opt_create(int nx, int ny, int nz, int mmax, int mrso, int precision, mdft_opt_t* out) {
  return guarded([&] {
    throw_if(!out, "null output handle");
    *out = nullptr;
    check_precision(precision);
    throw_if(nx <= 0 || ny <= 0 || nz <= 0, "The space is divided into grid nodes.");
    auto h = std::make_unique<mdft_opt_s>();
    h->nx = nx; h->ny = ny; h->nz = nz;
    h->N = static_cast<long>(nx) * ny * nz;
    h->precision = precision;
    h->grid = make_angular_grid(mmax, mrso);
    h->tab = std::make_unique<ProjTables>(h->grid);
    h->gemm = make_gemm(h->grid, precision);
    *out = h.release();
  });
}

int mdft_opt_destroy(mdft_opt_t h) {
  return guarded([&] { delete h; });
}

int mdft_angl2proj(mdft_opt_t h, const void* o, void* p, void* stream) {
  return guarded([&] {
    throw_if(!h || !o || !p, "angl2proj: null argument");
    hipStream_t s = as_stream(stream);
    switch (h->precision) {
      case MDFT_FP64:
        angl2proj_factorized<double, double, double>(*h->tab, h->N, static_cast<const double*>(o),
                                                     static_cast<cplx<double>*>(p), nullptr, s);
        break;
      case MDFT_FP32:
        angl2proj_factorized<float, float, float>(*h->tab, h->N, static_cast<const float*>(o),
                                                  static_cast<cplx<float>*>(p), nullptr, s);
        break;
      case MDFT_MIXED_X6F:
        h->gemm->angl2proj(h->N, static_cast<const float*>(o), static_cast<cplx<float>*>(p), nullptr, s);
        break;
      default:
        h->gemm->angl2proj(h->N, static_cast<const double*>(o), static_cast<cplx<float>*>(p), nullptr, s);
        break;
    }
  });
}

int mdft_proj2angl(mdft_opt_t h, const void* p, void* o, void* stream) {
  return guarded([&] {
    throw_if(!h || !o || !p, "proj2angl: null argument");
    hipStream_t s = as_stream(stream);
    switch (h->precision) {
      case MDFT_FP64:
        proj2angl_factorized<double, double, double>(*h->tab, h->N, static_cast<const cplx<double>*>(p),
                                                     static_cast<double*>(o), nullptr, nullptr, 0.0, nullptr, s);
        break;
      case MDFT_FP32:
        proj2angl_factorized<float, float, float>(*h->tab, h->N, static_cast<const cplx<float>*>(p),
                                                  static_cast<float*>(o), nullptr, nullptr, 0.0, nullptr, s);
        break;
      case MDFT_MIXED_X6F:
        h->gemm->proj2angl(h->N, static_cast<const cplx<float>*>(p), static_cast<float*>(o), nullptr, nullptr, 0.0,
                           nullptr, s);
        break;
      default:
        h->gemm->proj2angl(h->N, static_cast<const cplx<float>*>(p), static_cast<double*>(o), nullptr, nullptr,
                           0.0, nullptr, s);
        break;
    }
  });
}

// ---------------------------------------------------------------------------
static mdft_conv_s* make_conv(int nx, int ny, int nz, double lx, double ly, double lz, int mmax, int mrso,
                              int precision, const LucTable& luc) {
  check_precision(precision);
  auto h = std::make_unique<mdft_conv_s>();
  h->precision = precision;
  h->dev = std::make_unique<ConvDevice>(nx, ny, nz, lx, ly, lz, mmax, mrso, luc);
  return h.release();
}

int mdft_conv_create(const char* luc_path, int nx, int ny, int nz, double lx, double ly, double lz, int mmax,
                     int mrso, int np_luc, int precision, mdft_conv_t* out) {
  return guarded([&] {
    throw_if(!out || !luc_path, "conv_create: null argument");
    *out = nullptr;
    throw_if(nx <= 0 || ny <= 0 || nz <= 0, "The space is divided into grid nodes.");
    throw_if(lx <= 0 || ly <= 0 || lz <= 0, "The supercell cannot have negative length.");
    const auto kx = kproj(nx, lx), ky = kproj(ny, ly), kz = kproj(nz, lz);
    LucTable luc = read_luc(luc_path, np_luc, qmax_of_grid(kx, ky, kz, nz));
    *out = make_conv(nx, ny, nz, lx, ly, lz, mmax, mrso, precision, luc);
  });
}

int mdft_conv_create_from_table(int nx, int ny, int nz, double lx, double ly, double lz, int mmax, int mrso,
                                int np_luc, const int* tuples, const double* c, int nq_table, double dq,
                                int precision, mdft_conv_t* out) {
  return guarded([&] {
    throw_if(!out || !tuples || !c, "conv_create_from_table: null argument");
    *out = nullptr;
    throw_if(np_luc <= 0 || nq_table <= 0 || !(dq > 0), "conv_create_from_table: bad sizes");
    LucTable luc;
    luc.np_luc = np_luc;
    luc.nq = nq_table;
    luc.dq = dq;
    luc.tuples.assign(tuples, tuples + static_cast<size_t>(np_luc) * 5);
    luc.c.resize(static_cast<size_t>(np_luc) * nq_table);
    std::memcpy(luc.c.data(), c, luc.c.size() * sizeof(cd));
    *out = make_conv(nx, ny, nz, lx, ly, lz, mmax, mrso, precision, luc);
  });
}

int mdft_conv_info(mdft_conv_t h, int* nq, int* np_new, int* npairs, double* dq) {
  return guarded([&] {
    throw_if(!h, "null handle");
    if (nq) *nq = h->dev->tab.nq;
    if (np_new) *np_new = h->dev->tab.np_new;
    if (npairs) *npairs = static_cast<int>(h->dev->npairs);
    if (dq) *dq = h->dev->tab.dq;
  });
}

int mdft_conv_table(mdft_conv_t h, double* c) {
  return guarded([&] {
    throw_if(!h || !c, "null argument");
    std::memcpy(c, h->dev->tab.ctab.data(), h->dev->tab.ctab.size() * sizeof(cd));
  });
}

int mdft_conv_execute(mdft_conv_t h, void* p, void* stream) {
  return guarded([&] {
    throw_if(!h || !p, "conv_execute: null argument");
    if (h->precision == MDFT_FP64)
      h->dev->execute<double>(static_cast<cplx<double>*>(p), chop_eps_of(h->precision), as_stream(stream));
    else
      h->dev->execute<float>(static_cast<cplx<float>*>(p), chop_eps_of(h->precision), as_stream(stream));
  });
}

int mdft_conv_destroy(mdft_conv_t h) {
  return guarded([&] { delete h; });
}

int mdft_fft3d(void* data, int batch, int nx, int ny, int nz, int sign, int dtype, void* stream) {
  return guarded([&] {
    throw_if(!data || batch < 0 || nx <= 0 || ny <= 0 || nz <= 0, "fft3d: bad arguments");
    check_dtype(dtype);
    Fft3dPlan plan(nx, ny, nz);
    hipStream_t s = as_stream(stream);
    if (dtype == MDFT_DOUBLE)
      plan.execute<double>(static_cast<cplx<double>*>(data), batch, sign, s);
    else
      plan.execute<float>(static_cast<cplx<float>*>(data), batch, sign, s);
    // the plan's twiddle tables are freed on return: finish the work first
    MDFT_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int mdft_fft3d_config(void* data, int batch, int nx, int ny, int nz, int sign, int dtype, int plane,
                      int plane_threads, int wz, int wy, int wx, void* stream) {
  return guarded([&] {
    throw_if(!data || batch < 0 || nx <= 0 || ny <= 0 || nz <= 0, "fft3d: bad arguments");
    check_dtype(dtype);
    Fft3dPlan plan(nx, ny, nz);
    FftConfig c;
    c.plane = plane != 0;
    c.plane_threads = plane_threads;
    c.wz = wz;
    c.wy = wy;
    c.wx = wx;
    hipStream_t s = as_stream(stream);
    if (dtype == MDFT_DOUBLE)
      plan.execute_config<double>(c, static_cast<cplx<double>*>(data), batch, sign, s);
    else
      plan.execute_config<float>(c, static_cast<cplx<float>*>(data), batch, sign, s);
    MDFT_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int mdft_conv_fft_config(mdft_conv_t h, int* tuned, int* plane, int* plane_threads, int* wz, int* wy, int* wx) {
  return guarded([&] {
    throw_if(!h, "null handle");
    const Fft3dPlan& f = h->dev->fft;
    const bool f64 = h->precision == MDFT_FP64;
    const bool t = f64 ? f.tuned<double>() : f.tuned<float>();
    FftConfig c = f64 ? f.default_config<double>() : f.default_config<float>();
    if (t) c = f64 ? f.autotune<double>(0, nullptr) : f.autotune<float>(0, nullptr);  // the cached choice
    if (tuned) *tuned = t ? 1 : 0;
    if (plane) *plane = c.plane ? 1 : 0;
    if (plane_threads) *plane_threads = c.plane_threads;
    if (wz) *wz = c.wz;
    if (wy) *wy = c.wy;
    if (wx) *wx = c.wx;
  });
}

// ---------------------------------------------------------------------------
int mdft_delta_rho(const void* xi, void* delta_rho, size_t n, double rho0, int dtype, void* stream) {
  return guarded([&] {
    throw_if((!xi || !delta_rho) && n > 0, "delta_rho: null argument");
    check_dtype(dtype);
    if (dtype == MDFT_DOUBLE)
      mdft::delta_rho<double>(static_cast<const double*>(xi), static_cast<double*>(delta_rho), n, rho0,
                              as_stream(stream));
    else
      mdft::delta_rho<float>(static_cast<const float*>(xi), static_cast<float*>(delta_rho), n, rho0,
                             as_stream(stream));
  });
}

int mdft_delta_f(const void* xi, const void* vexc, const double* w, void* delta_f, size_t nxyz, int no, double rho0,
                 double prefactor, double* ff, int dtype, void* stream) {
  return guarded([&] {
    throw_if(!ff || no <= 0, "delta_f: bad arguments");
    check_dtype(dtype);
    throw_if(nxyz > 0 && (!xi || !vexc || !w || !delta_f), "delta_f: null argument");
    if (nxyz == 0) {
      *ff = 0.0;
      return;
    }
    hipStream_t s = as_stream(stream);
    const int nb = delta_f_blocks(static_cast<long>(nxyz));
    double* part = nullptr;
    MDFT_HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&part), (nb + 1) * sizeof(double), s));
    if (dtype == MDFT_DOUBLE)
      mdft::delta_f<double>(static_cast<const double*>(xi), static_cast<const double*>(vexc), w,
                            static_cast<double*>(delta_f), static_cast<long>(nxyz), no, rho0, part, nb, s);
    else
      mdft::delta_f<float>(static_cast<const float*>(xi), static_cast<const float*>(vexc), w,
                           static_cast<float*>(delta_f), static_cast<long>(nxyz), no, rho0, part, nb, s);
    reduce_partials(part, nb, rho0 * prefactor, part + nb, s);
    MDFT_HIP_CHECK(hipMemcpyAsync(ff, part + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    MDFT_HIP_CHECK(hipFreeAsync(part, s));
    MDFT_HIP_CHECK(hipStreamSynchronize(s));
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// get_delta_f without allocation: the partial sums and the reduced ff live in
// `scratch`, a file-scope device array (each GPU the library runs on holds
// its own copy until the process ends).  Its device address on each GPU comes
// from hipGetSymbolAddress, taken once per device with that device current;
// the host-side name `scratch` is not a device address and is never passed to
// a kernel or a copy.
constexpr int kNoAllocScratchDoubles = 131072;  // 1 MiB per GPU
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];
// one partial per delta_f block of grid points, plus the reduced value
static_assert(static_cast<long>(MDFT_DELTA_F_NOALLOC_MAX_POINTS) == (kNoAllocScratchDoubles - 1) * 8L,
              "the header's limit must match the scratch size");

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxScratchDevices] = {};

// the device address of `scratch` on the current device
double* scratch_of_current_device() {
  int dev = -1;
  MDFT_HIP_CHECK(hipGetDevice(&dev));
  throw_if(dev < 0 || dev >= kMaxScratchDevices, "delta_f_noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    throw_if(e != hipSuccess || !p,
             std::string("delta_f_noalloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int mdft_delta_f_noalloc(const void* xi, const void* vexc, const double* w, void* delta_f, size_t nxyz, int no,
                         double rho0, double prefactor, double* ff, int dtype, void* stream) {
  return guarded([&] {
    // every argument and size check comes before the first HIP call
    throw_if(!ff || no <= 0, "delta_f_noalloc: bad arguments");
    check_dtype(dtype);
    throw_if(nxyz > 0 && (!xi || !vexc || !w || !delta_f), "delta_f_noalloc: null argument");
    throw_if(nxyz > static_cast<size_t>(MDFT_DELTA_F_NOALLOC_MAX_POINTS),
             "delta_f_noalloc: nxyz = " + std::to_string(nxyz) + " exceeds the scratch limit of " +
                 std::to_string(MDFT_DELTA_F_NOALLOC_MAX_POINTS) + " grid points");
    throw_if(delta_f_rows_per_block() != 8, "delta_f_noalloc: scratch sized for 8 grid points per partial");
    if (nxyz == 0) {
      *ff = 0.0;
      return;
    }
    hipStream_t s = as_stream(stream);
    double* part = scratch_of_current_device();
    const int nb = delta_f_blocks(static_cast<long>(nxyz));
    // every partial [0, nb) is written by the delta_f pass and the value at
    // part[nb] by the reduction, so nothing needs clearing between calls
    if (dtype == MDFT_DOUBLE)
      mdft::delta_f<double>(static_cast<const double*>(xi), static_cast<const double*>(vexc), w,
                            static_cast<double*>(delta_f), static_cast<long>(nxyz), no, rho0, part,
                            kNoAllocScratchDoubles - 1, s);
    else
      mdft::delta_f<float>(static_cast<const float*>(xi), static_cast<const float*>(vexc), w,
                           static_cast<float*>(delta_f), static_cast<long>(nxyz), no, rho0, part,
                           kNoAllocScratchDoubles - 1, s);
    reduce_partials(part, nb, rho0 * prefactor, part + nb, s);
    MDFT_HIP_CHECK(hipMemcpyAsync(ff, part + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    MDFT_HIP_CHECK(hipStreamSynchronize(s));
  });
}

// ---------------------------------------------------------------------------
int mdft_solver_create_batch(mdft_conv_t conv, double rho0, double prefactor, int precision, int capacity,
                             mdft_solver_t* out) {
  return guarded([&] {
    throw_if(!out || !conv, "solver_create: null argument");
    *out = nullptr;
    check_precision(precision);
    throw_if(capacity < 1, "solver_create: capacity must be at least 1");
    throw_if((precision == MDFT_FP64) != (conv->precision == MDFT_FP64),
             "solver and convolution must agree on FP64 vs FP32 projections");
    auto h = std::make_unique<mdft_solver_s>();
    h->conv = conv;
    h->precision = precision;
    h->rho0 = rho0;
    h->prefactor = prefactor;
    h->capacity = capacity;
    const ConvTables& t = conv->dev->tab;
    h->N = static_cast<long>(t.nx) * t.ny * t.nz;
    throw_if(capacity > 1 && h->N % 64 != 0, "batched evaluation needs nx ny nz to be a multiple of 64");
    h->tab = std::make_unique<ProjTables>(t.ang);
    h->compact = compact_available(precision, conv, *h->tab);
    h->gemm = make_gemm(t.ang, precision, h->compact);
    h->field_parts = h->gemm ? h->gemm->p2a_blocks(h->N) : proj2angl_blocks(*h->tab, h->N);
    const size_t pbytes = static_cast<size_t>(h->N) * t.ang.np * capacity * (precision == MDFT_FP64 ? 16 : 8);
    h->nparts = h->field_parts * capacity;
    MDFT_HIP_CHECK(hipMalloc(&h->pbuf, pbytes));
    MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&h->partials), h->nparts * sizeof(double)));
    MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&h->ffdev), capacity * sizeof(double)));
    h->bytes = pbytes + (h->nparts + capacity) * sizeof(double);
    if ((precision == MDFT_MIXED_X6 || precision == MDFT_MIXED_X6F) && mean_mode_supported(h.get())) {
      enable_mean_mode(h.get());
      h->mean_on = true;
    }
    *out = h.release();
  });
}

int mdft_solver_create(mdft_conv_t conv, double rho0, double prefactor, int precision, mdft_solver_t* out) {
  return mdft_solver_create_batch(conv, rho0, prefactor, precision, 1, out);
}

int mdft_solver_workspace_bytes(mdft_solver_t h, size_t* bytes) {
  return guarded([&] {
    throw_if(!h || !bytes, "null argument");
    *bytes = h->bytes;
  });
}

int mdft_solver_destroy(mdft_solver_t h) {
  return guarded([&] { delete h; });
}

// nb fields back to back in xi and df (N no each); ff of field j into
// ffd[j ffstride].  Every stage runs once over all nb fields: the pointwise
// projections over nb N points, the FFTs over np nb transforms, the OZ pass
// over the nb fields of each pair.  Each field's arithmetic (and its partial
// sums, reduced in the same order) is the same as when evaluated alone.
static void energy_impl(mdft_solver_t h, int nb, const void* xi, void* df, double* ffd, long ffstride,
                        hipStream_t s) {
  throw_if(!h || !xi || !df || !ffd, "energy_cproj: null argument");
  throw_if(nb < 1 || nb > h->capacity, "energy_cproj: batch size exceeds the solver's capacity");
  const double rho0 = h->rho0;
  const double eps = chop_eps_of(h->precision);
  const long Nb = h->N * nb;
  double* colp = h->mean_on ? h->colpart : nullptr;
  StageTimer* tm = h->timer.get();
  if (tm) {
    tm->begin_eval();
    tm->mark(s, kStageStart);
  }
  switch (h->precision) {
    case MDFT_FP64: {
      auto* p = static_cast<cplx<double>*>(h->pbuf);
      angl2proj_factorized<double, double, double>(*h->tab, Nb, static_cast<const double*>(xi), p, &rho0, s,
                                                   h->compact);
      h->conv->dev->execute<double>(p, eps, s, tm, nb, h->compact);
      proj2angl_factorized<double, double, double>(*h->tab, Nb, p, nullptr, static_cast<const double*>(xi),
                                                   static_cast<double*>(df), rho0, h->partials, s, colp,
                                                   h->compact);
      break;
    }
    case MDFT_FP32: {
      auto* p = static_cast<cplx<float>*>(h->pbuf);
      angl2proj_factorized<float, float, float>(*h->tab, Nb, static_cast<const float*>(xi), p, &rho0, s,
                                                h->compact);
      h->conv->dev->execute<float>(p, eps, s, tm, nb, h->compact);
      proj2angl_factorized<float, float, float>(*h->tab, Nb, p, nullptr, static_cast<const float*>(xi),
                                                static_cast<float*>(df), rho0, h->partials, s, colp,
                                                h->compact);
      break;
    }
    case MDFT_MIXED_X6F: {
      auto* p = static_cast<cplx<float>*>(h->pbuf);
      h->gemm->angl2proj(Nb, static_cast<const float*>(xi), p, &rho0, s);
      h->conv->dev->execute<float>(p, eps, s, tm, nb, h->compact);
      h->gemm->proj2angl(Nb, p, nullptr, static_cast<const float*>(xi), static_cast<float*>(df), rho0, h->partials,
                         s, colp);
      break;
    }
    default: {
      auto* p = static_cast<cplx<float>*>(h->pbuf);
      h->gemm->angl2proj(Nb, static_cast<const double*>(xi), p, &rho0, s);
      h->conv->dev->execute<float>(p, eps, s, tm, nb, h->compact);
      h->gemm->proj2angl(Nb, p, nullptr, static_cast<const double*>(xi), static_cast<double*>(df), rho0, h->partials,
                         s, colp);
      break;
    }
  }
  if (colp) h->mean->run(colp, h->N, nb, h->meanwork, h->extra, s);
  reduce_partials(h->partials, h->field_parts, rho0 * h->prefactor, ffd, s, nb, ffstride, colp ? h->extra : nullptr,
                  colp ? h->mean->extra_per_field() : 0);
  if (tm) tm->mark(s, kStageEnd);
}

int mdft_solver_set_compact(mdft_solver_t h, int on) {
  return guarded([&] {
    throw_if(!h, "null handle");
    if (on)
      throw_if(!compact_available(h->precision, h->conv, *h->tab),
               "solver: no compact channel order for this form and grid");
    if ((on != 0) == h->compact) return;
    h->compact = on != 0;
    h->gemm = make_gemm(h->conv->dev->tab.ang, h->precision, h->compact);
  });
}

int mdft_solver_compact(mdft_solver_t h, int* on) {
  return guarded([&] {
    throw_if(!h || !on, "null argument");
    *on = h->compact ? 1 : 0;
  });
}

int mdft_solver_set_mean_fp64(mdft_solver_t h, int on) {
  return guarded([&] {
    throw_if(!h, "null handle");
    if (on) {
      throw_if(!mean_mode_supported(h),
               "solver: the FP64 mean mode needs proj2angl on the vector ALUs and a compile-time angular grid");
      enable_mean_mode(h);
    }
    h->mean_on = on != 0;
  });
}

int mdft_solver_set_timing(mdft_solver_t h, int enable) {
  return guarded([&] {
    throw_if(!h, "null handle");
    throw_if(enable < 0 || enable > 2, "stage timing mode must be 0, 1 or 2");
    if (enable)
      h->timer = std::make_unique<StageTimer>(enable == 2);
    else
      h->timer.reset();
  });
}

int mdft_solver_stage_ms(mdft_solver_t h, double* ms, long* count) {
  return guarded([&] {
    throw_if(!h || !ms, "null argument");
    throw_if(!h->timer, "stage timing is off (mdft_solver_set_timing)");
    h->timer->collect();
    for (int i = 0; i < kNumStages; ++i) ms[i] = h->timer->acc()[i];
    if (count) *count = h->timer->count();
  });
}

int mdft_energy_cproj(mdft_solver_t h, const void* xi, void* df, double* ff, void* stream) {
  return mdft_energy_cproj_batch(h, 1, xi, df, ff, stream);
}

int mdft_energy_cproj_async(mdft_solver_t h, const void* xi, void* df, double* ff_device, void* stream) {
  return guarded([&] { energy_impl(h, 1, xi, df, ff_device, 1, as_stream(stream)); });
}

int mdft_energy_cproj_batch(mdft_solver_t h, int nbatch, const void* xi, void* df, double* ff, void* stream) {
  return guarded([&] {
    throw_if(!ff, "energy_cproj: null ff");
    hipStream_t s = as_stream(stream);
    energy_impl(h, nbatch, xi, df, h ? h->ffdev : nullptr, 1, s);
    MDFT_HIP_CHECK(hipMemcpyAsync(ff, h->ffdev, nbatch * sizeof(double), hipMemcpyDeviceToHost, s));
    MDFT_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int mdft_energy_cproj_batch_async(mdft_solver_t h, int nbatch, const void* xi, void* df, double* ff_device,
                                  void* stream) {
  return guarded([&] { energy_impl(h, nbatch, xi, df, ff_device, 1, as_stream(stream)); });
}

int mdft_fill_uniform(void* x, size_t n, unsigned long long seed, unsigned long long offset, double lo, double hi,
                      int dtype, void* stream) {
  return guarded([&] {
    throw_if(!x && n > 0, "fill_uniform: null argument");
    check_dtype(dtype);
    if (dtype == MDFT_DOUBLE)
      fill_uniform<double>(static_cast<double*>(x), n, seed, offset, lo, hi, as_stream(stream));
    else
      fill_uniform<float>(static_cast<float*>(x), n, seed, offset, lo, hi, as_stream(stream));
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// external potential, full functional and minimiser
struct mdft_functional_s {
  mdft_solver_t solver;  // borrowed
  std::unique_ptr<Functional> f;
};

extern "C" {

int mdft_rotation_matrices(int mmax, int mrso, double* rot) {
  return guarded([&] {
    throw_if(!rot, "rotation_matrices: null output");
    const std::vector<double> R = rotation_matrices(make_angular_grid(mmax, mrso));
    std::memcpy(rot, R.data(), R.size() * sizeof(double));
  });
}

int mdft_vext(int nx, int ny, int nz, double lx, double ly, double lz, int mmax, int mrso, int nsolute,
              const double* solute, int nsolvent, const double* solvent, double vmax, double* vext, void* stream) {
  return guarded([&] {
    compute_vext(nx, ny, nz, lx, ly, lz, make_angular_grid(mmax, mrso), solute, nsolute, solvent, nsolvent, vmax,
                 vext, as_stream(stream));
  });
}

int mdft_xi_from_vext(const double* vext, long long n, double kT, double* xi, void* stream) {
  return guarded([&] {
    xi_from_vext(vext, static_cast<long>(n), kT, xi, as_stream(stream));
    MDFT_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  });
}

int mdft_functional_create(mdft_solver_t solver, const double* vext, double kT, mdft_functional_t* out) {
  return guarded([&] {
    throw_if(!out || !solver || !vext, "functional_create: null argument");
    *out = nullptr;
    throw_if(solver->precision == MDFT_FP32,
             "functional: the MDFT_FP32 form has no FP64 reductions; use MDFT_FP64 or a mixed form");
    const ConvTables& t = solver->conv->dev->tab;
    FunctionalParams p{solver->N, t.ang.no, solver->rho0, solver->prefactor, kT,
                       t.lx * t.ly * t.lz / static_cast<double>(solver->N)};
    auto h = std::make_unique<mdft_functional_s>();
    h->solver = solver;
    // the mixed forms keep their L-BFGS pairs and gradients in FP32; MDFT_MIXED_X6F
    // evaluates the excess on an FP32 copy of the FP64 iterate and returns FP32 df
    h->f = std::make_unique<Functional>(
        p, t.ang.w, vext,
        [solver](const void* xi, void* df, double* ffd, hipStream_t s) { energy_impl(solver, 1, xi, df, ffd, 1, s); },
        solver->precision != MDFT_FP64, solver->precision == MDFT_MIXED_X6F);
    *out = h.release();
  });
}

int mdft_functional_destroy(mdft_functional_t h) {
  return guarded([&] { delete h; });
}

int mdft_functional_eval(mdft_functional_t h, const double* xi, double* g, double* F, void* stream) {
  return guarded([&] {
    throw_if(!h, "functional_eval: null handle");
    h->f->eval(xi, g, F, as_stream(stream));
  });
}

int mdft_lbfgs_minimize(mdft_functional_t h, double* xi, int m, int max_iter, double gtol, double* F,
                        int* iterations, int* evaluations, int* status, double* residual, double* history,
                        void* stream) {
  return guarded([&] {
    throw_if(!h, "lbfgs: null handle");
    LbfgsOptions o;
    o.m = m;
    o.max_iter = max_iter;
    o.gtol = gtol;
    const LbfgsResult r = h->f->minimize(xi, o, as_stream(stream));
    if (F) std::memcpy(F, r.F, sizeof(r.F));
    if (iterations) *iterations = r.iterations;
    if (evaluations) *evaluations = r.evaluations;
    if (status) *status = r.status;
    if (residual) *residual = r.residual;
    if (history) std::memcpy(history, r.history.data(), r.history.size() * sizeof(double));
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched and served minimisation
namespace {

// the convolution and solver of one engine's shape; the engine's excess
// callback holds them for as long as the engine lives
struct ShapeHold {
  std::unique_ptr<mdft_conv_s> conv;
  mdft_solver_t solver = nullptr;
  ~ShapeHold() { delete solver; }
};

constexpr double kBoltz = 1.38064852e-23, kNavo = 6.02214129e23, kPiC = 3.14159265358979323846;

}  // namespace

struct mdft_server_s {
  LucTable luc;
  int mrso = 2, precision = MDFT_FP64, capacity = 1;
  double kT = 0, rho0 = 0, prefactor = 0;
  LbfgsOptions opt;
  bool graphs = true;
  bool start_vext = false;
  std::unique_ptr<Server> server;
  std::string last_poll_error;  // of the result mdft_server_poll returned last

  // the tuples an mmax needs (m, n <= mmax), in the table's order
  LucTable luc_for(int mmax) const {
    LucTable out;
    out.nq = luc.nq;
    out.dq = luc.dq;
    for (int i = 0; i < luc.np_luc; ++i) {
      const int* t = &luc.tuples[static_cast<size_t>(i) * 5];
      if (t[0] > mmax || t[1] > mmax) continue;
      out.tuples.insert(out.tuples.end(), t, t + 5);
      out.c.insert(out.c.end(), luc.c.begin() + static_cast<long>(i) * luc.nq,
                   luc.c.begin() + static_cast<long>(i + 1) * luc.nq);
      ++out.np_luc;
    }
    return out;
  }

  // device bytes of one solute, and an estimate of an engine's tables,
  // plans and slabs (d table, c(q), projections, trial and df slabs)
  std::pair<size_t, size_t> footprint(const Server::Shape& sh) const {
    const AngularGrid ag = make_angular_grid(sh.mmax, mrso);
    const size_t N = static_cast<size_t>(sh.nx) * sh.ny * sh.nz, n = N * ag.no;
    const size_t ts = precision == MDFT_FP64 ? 8 : 4;
    const size_t job = 2 * n * 8 + 2 * static_cast<size_t>(opt.m + 2) * n * ts;  // Engine::job_bytes
    const size_t proj = N * ag.np * capacity * (precision == MDFT_FP64 ? 16 : 8);
    const size_t dtab = (N / 2 + 8) * static_cast<size_t>((sh.mmax + 1) * (sh.mmax + 1)) * (2 * sh.mmax + 1) * ts;
    const size_t slabs = 2 * static_cast<size_t>(capacity) * n * 8 + n * ts;
    // the column sums of the FP64 mean mode (MDFT_MIXED_X6, MDFT_MIXED_X6F)
    const size_t colp =
        precision == MDFT_MIXED_X6 || precision == MDFT_MIXED_X6F ? (N * capacity / 64 + 1) * 2 * ag.no * 8 : 0;
    return {job, proj + dtab + slabs + colp + (size_t(64) << 20)};
  }
};

extern "C" {

int mdft_server_create(int np_luc, const int* tuples, const double* c, int nq_table, double dq, int mrso, double n0,
                       double temperature, int precision, int capacity, int m, int max_iter, double gtol,
                       double hbm_budget_bytes, mdft_server_t* out) {
  return guarded([&] {
    throw_if(!out || !tuples || !c, "server_create: null argument");
    *out = nullptr;
    check_precision(precision);
    throw_if(precision == MDFT_FP32, "server: the MDFT_FP32 form has no FP64 reductions; use MDFT_FP64 or a mixed form");
    throw_if(np_luc <= 0 || nq_table <= 0 || !(dq > 0) || mrso < 1 || !(n0 > 0) || !(temperature > 0),
             "server_create: bad arguments");
    throw_if(capacity < 1, "server_create: capacity must be at least 1");
    throw_if(m < 1 || m > kLbfgsMaxPairs, "lbfgs: m must be in 1 .. 8");
    auto h = std::make_unique<mdft_server_s>();
    h->luc.np_luc = np_luc;
    h->luc.nq = nq_table;
    h->luc.dq = dq;
    h->luc.tuples.assign(tuples, tuples + static_cast<size_t>(np_luc) * 5);
    h->luc.c.resize(static_cast<size_t>(np_luc) * nq_table);
    std::memcpy(h->luc.c.data(), c, h->luc.c.size() * sizeof(cd));
    h->mrso = mrso;
    h->precision = precision;
    h->capacity = capacity;
    // Thermo::m_kbt and the Solvent's rho0 and prefactor
    // (common/MDFT_Thermo.hpp:60, common/MDFT_Solvent.hpp:366-368, :430-431)
    h->kT = kBoltz * kNavo * temperature * 1e-3;
    h->rho0 = n0 / (8 * kPiC * kPiC / mrso);
    h->prefactor = -2 * h->kT * 4 * kPiC * kPiC / mrso / n0;
    h->opt.m = m;
    h->opt.max_iter = max_iter;
    h->opt.gtol = gtol;
    mdft_server_s* sp = h.get();
    auto make = [sp](const Server::Shape& sh, size_t& table_bytes) {
      auto hold = std::make_shared<ShapeHold>();
      hold->conv.reset(
          make_conv(sh.nx, sh.ny, sh.nz, sh.lx, sh.ly, sh.lz, sh.mmax, sp->mrso, sp->precision, sp->luc_for(sh.mmax)));
      mdft_solver_t sv = nullptr;
      if (mdft_solver_create_batch(hold->conv.get(), sp->rho0, sp->prefactor, sp->precision, sp->capacity, &sv))
        throw Error(g_last_error);
      hold->solver = sv;
      const ConvTables& t = hold->conv->dev->tab;
      FunctionalParams p{sv->N, t.ang.no, sp->rho0, sp->prefactor, sp->kT,
                         t.lx * t.ly * t.lz / static_cast<double>(sv->N)};
      auto eng = std::make_unique<Engine>(
          p, t.ang.w, sp->capacity, sp->opt, sp->precision != MDFT_FP64,
          [hold, sv](int nb, const void* x, void* df, double* ff, long st, hipStream_t s) {
            energy_impl(sv, nb, x, df, ff, st, s);
          },
          sp->precision == MDFT_MIXED_X6F);
      eng->use_graphs(sp->graphs);
      eng->start_from_vext(sp->start_vext);
      table_bytes = sp->footprint(sh).second;
      return eng;
    };
    auto fp = [sp](const Server::Shape& sh) { return sp->footprint(sh); };
    h->server = std::make_unique<Server>(make, fp, hbm_budget_bytes, mrso);
    *out = h.release();
  });
}

int mdft_server_destroy(mdft_server_t h) {
  return guarded([&] { delete h; });
}

int mdft_server_use_graphs(mdft_server_t h, int on) {
  return guarded([&] {
    throw_if(!h, "null handle");
    throw_if(h->server->engines() > 0, "server: set the graph mode before the first step");
    h->graphs = on != 0;
  });
}

int mdft_server_start_from_vext(mdft_server_t h, int on) {
  return guarded([&] {
    throw_if(!h, "null handle");
    throw_if(h->server->engines() > 0, "server: set the start point before the first step");
    h->start_vext = on != 0;
  });
}

int mdft_server_submit(mdft_server_t h, int nx, int ny, int nz, double lx, double ly, double lz, int mmax,
                       int nsolute, const double* solute, int nsolvent, const double* solvent, double vmax,
                       double* out_xi, long long* id) {
  return guarded([&] {
    throw_if(!h || !id, "server_submit: null argument");
    throw_if(nx <= 0 || ny <= 0 || nz <= 0 || !(lx > 0) || !(ly > 0) || !(lz > 0), "server_submit: bad grid");
    throw_if(h->capacity > 1 && (static_cast<long>(nx) * ny * nz) % 64 != 0,
             "batched evaluation needs nx ny nz to be a multiple of 64");
    (void)make_angular_grid(mmax, h->mrso);  // validates mmax
    // the c(q) table must cover the grid (q rows up to its qmax, every tuple
    // of mmax): the host tables of the convolution are built (and dropped)
    // here, so a request the engine could not build fails at submission
    (void)make_conv_tables(nx, ny, nz, lx, ly, lz, mmax, h->mrso, h->luc_for(mmax));
    *id = h->server->submit(Server::Shape{nx, ny, nz, lx, ly, lz, mmax}, solute, nsolute, solvent, nsolvent, vmax,
                            out_xi);
  });
}

int mdft_server_step(mdft_server_t h, int* active, int* pending) {
  return guarded([&] {
    throw_if(!h, "null handle");
    h->server->step();
    if (active) *active = h->server->active();
    if (pending) *pending = h->server->pending();
  });
}

int mdft_server_poll(mdft_server_t h, long long* id, double* F, int* iterations, int* evaluations, int* status,
                     double* residual, double* queued_s, double* total_s, int* got) {
  return guarded([&] {
    throw_if(!h || !got, "server_poll: null argument");
    Server::Result r;
    *got = h->server->poll(r) ? 1 : 0;
    if (!*got) return;
    if (id) *id = r.id;
    if (F) std::memcpy(F, r.r.F, sizeof(r.r.F));
    if (iterations) *iterations = r.r.iterations;
    if (evaluations) *evaluations = r.r.evaluations;
    if (status) *status = r.r.status;
    if (residual) *residual = r.r.residual;
    if (queued_s) *queued_s = r.queued_s;
    if (total_s) *total_s = r.total_s;
    h->last_poll_error = r.error;
  });
}

int mdft_server_poll_error(mdft_server_t h, char* buf, int buflen) {
  return guarded([&] {
    throw_if(!h || !buf || buflen <= 0, "server_poll_error: bad arguments");
    const std::string& e = h->last_poll_error;
    const size_t n = std::min(e.size(), static_cast<size_t>(buflen - 1));
    std::memcpy(buf, e.data(), n);
    buf[n] = '\0';
  });
}

int mdft_server_info(mdft_server_t h, double* bytes_in_use, int* engines, int* active, int* pending) {
  return guarded([&] {
    throw_if(!h, "null handle");
    if (bytes_in_use) *bytes_in_use = h->server->bytes_in_use();
    if (engines) *engines = h->server->engines();
    if (active) *active = h->server->active();
    if (pending) *pending = h->server->pending();
  });
}

}  // extern "C"
