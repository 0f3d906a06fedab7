// Batched and served minimisation (no counterpart in the reference, whose
// Solver runs one field once, src/mini-mdft.hpp:129-133).
//
// Engine: capacity slots of one shape.  A tick gives every active slot one
// L-BFGS trial (or its start-point evaluation):
//   1. slot_direction_kernel: for each slot, x_t = x + t d into the slot's
//      place in the trial slab (the start point is copied there);
//   2. the excess evaluation of the whole slab in one batched pass;
//   3. slot_update_kernel: for each slot, g, y and the Gram entries (or, at
//      the start point, g and its sums);
//   4. slot_reduce_kernel: the per-slot sums, in a fixed order.
// Steps 1-4 are one HIP graph per batch width (captured after one eager tick
// of that width has made every lazy table), replayed every tick.  The slots'
// scalars (L-BFGS coefficients, step, pointers) live in a device table
// uploaded before the replay.  The per-slot decisions are LbfgsState's, the
// same code that drives Functional::minimize; the passes use every pair column
// up to m, unused columns reading a shared zero field with zero coefficients,
// so each slot's arithmetic is that of a solve alone (a slot's results do not
// depend on what else is in the batch).
//
// Server: requests of any shape wait in a queue; admission is first fit in
// arrival order against an HBM budget (a solute's fields plus, for a shape
// with no engine yet, the engine's tables and slabs); engines of different
// shapes tick concurrently on their own streams.
#include "mdft_engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "mdft_elem.hpp"

namespace mdft {

double now_seconds() {
  using clock = std::chrono::steady_clock;
  return std::chrono::duration<double>(clock::now().time_since_epoch()).count();
}

namespace {

using namespace elem;

constexpr int kModeIdle = 0, kModeStart = 1, kModeTrial = 2;
// per-slot rows of the partial sums / results: the update pass's 5 m + 10
// rows for m = kLbfgsMaxPairs at most, then ff
constexpr int kSlotRows = 5 * kLbfgsMaxPairs + 10;
constexpr int kSlotRes = kSlotRows + 1;
constexpr int kSlotFf = kSlotRows;

template <typename TS>
struct SlotDesc {
  int mode;
  int k;   // stored pairs in use (columns k .. KM - 1 read the zero field)
  int k0;  // the update pass reads pairs k0 .. k0 + KM - 2 (LbfgsState::Trial::k0)
  double cg, gamma, t;
  double cs[kLbfgsMaxPairs], cq[kLbfgsMaxPairs];
  const double* src;  // current point (start: the initial field)
  double* save;       // copy of the current point, or null
  const double* V;
  TS* g;   // g at the current point (written at the start point)
  TS* gt;  // g at the trial
  const TS* s[kLbfgsMaxPairs];
  const TS* y[kLbfgsMaxPairs];
  TS* sc;
  TS* yn;
};

// grid (nb, slots): x_t of slot blockIdx.y into xslab + slot n, and with
// F32F its FP32 copy into xfslab + slot n (the excess evaluation's input)
template <typename TS, int KM, bool F32F>
// (xslab is not restrict: a slot's current point may be its slab field, read
// and overwritten in place element by element)
__global__ __launch_bounds__(kFThreads) void slot_direction_kernel(const SlotDesc<TS>* __restrict__ desc,
                                                                   double* xslab, float* __restrict__ xfslab,
                                                                   Elem e) {
  const SlotDesc<TS>& a = desc[blockIdx.y];
  if (a.mode == kModeIdle) return;
  double* xt = xslab + static_cast<long>(blockIdx.y) * e.n;
  float* xf = F32F ? xfslab + static_cast<long>(blockIdx.y) * e.n : nullptr;
  if (a.mode == kModeStart) {
    MDFT_ELEMENT_LOOP(e) {
      const double xv = ld_nt(&a.src[i]);
      st_nt(&xt[i], xv);
      if (F32F) st_nt(&xf[i], static_cast<float>(xv));
    }
    return;
  }
  MDFT_ELEMENT_LOOP(e) {
    double q = a.cg * static_cast<double>(ld_nt(&a.g[i]));
#pragma unroll
    for (int j = 0; j < KM; ++j) q = fma(a.cq[j], static_cast<double>(ld_nt(&a.y[j][i])), q);
    double r = a.gamma * e.h0[io] * q;
#pragma unroll
    for (int j = 0; j < KM; ++j) r = fma(a.cs[j], static_cast<double>(ld_nt(&a.s[j][i])), r);
    const double xv = ld_nt(&a.src[i]);
    if (a.save) st_nt(&a.save[i], xv);
    const double xn = xv + a.t * (-r);
    st_nt(&xt[i], xn);
    if (F32F) st_nt(&xf[i], static_cast<float>(xn));
    st_nt(&a.sc[i], static_cast<TS>(xn - xv));
  }
}

// grid (nb, slots): the update pass of Functional's L-BFGS for slot
// blockIdx.y (rows as in update_kernel with K = KM, KM = m - 1: the pairs
// from the trial's k0); at the start point, g and its sums into the same rows
// (5K+6 g'H g, 5K+7 F_id, 5K+8 F_ext, 5K+9 max)
template <typename TS, int KM, typename TF>
__global__ __launch_bounds__(kFThreads) void slot_update_kernel(const SlotDesc<TS>* __restrict__ desc,
                                                                const double* __restrict__ xslab,
                                                                const TF* __restrict__ dfslab, Elem e,
                                                                double* __restrict__ part, int nb) {
  constexpr int NA = 5 * KM + 10;
  constexpr int KA = KM > 0 ? KM : 1;
  const SlotDesc<TS>& a = desc[blockIdx.y];
  if (a.mode == kModeIdle) return;
  const double* xt = xslab + static_cast<long>(blockIdx.y) * e.n;
  const TF* df = dfslab + static_cast<long>(blockIdx.y) * e.n;
  part += static_cast<long>(blockIdx.y) * kSlotRows * nb;
  double acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = 0.0;
  if (a.mode == kModeStart) {
    MDFT_ELEMENT_LOOP(e) {
      const double xv = ld_nt(&xt[i]), v = ld_nt(&a.V[i]), wio = e.w[io], h = e.h0[io];
      const double dfv = static_cast<double>(ld_nt(&df[i]));
      const TS gs = static_cast<TS>(local_terms(xv, v, dfv, wio, e, acc[5 * KM + 7], acc[5 * KM + 8]));
      st_nt(&a.g[i], gs);
      const double gv = static_cast<double>(gs);
      const double hg = h * gv;
      acc[5 * KM + 6] += gv * hg;
      acc[5 * KM + 9] = fmax(acc[5 * KM + 9], fabs(hg));
    }
  } else {
    // two elements per trip, both elements' loads first (update_kernel's
    // loop; the sums take the elements in the one-element order)
    const long stride_ = static_cast<long>(gridDim.x) * kFThreads;
    long i = static_cast<long>(blockIdx.x) * kFThreads + threadIdx.x;
    int io = static_cast<int>(i % e.no);
    auto nio = [&](int x) { return x + e.step >= e.no ? x + e.step - e.no : x + e.step; };
    for (; i < e.n; i += 2 * stride_, io = nio(nio(io))) {
      const bool two = i + stride_ < e.n;
      const long ix[2] = {i, two ? i + stride_ : i};
      const int iox[2] = {io, nio(io)};
      double xv[2], v[2], dfv[2], gv[2], scv[2], sj[2][KA], yj[2][KA];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        xv[u] = ld_nt(&xt[ix[u]]);
        v[u] = ld_nt(&a.V[ix[u]]);
        dfv[u] = static_cast<double>(ld_nt(&df[ix[u]]));
        gv[u] = static_cast<double>(ld_nt(&a.g[ix[u]]));
        scv[u] = static_cast<double>(ld_nt(&a.sc[ix[u]]));
#pragma unroll
        for (int j = 0; j < KM; ++j) {
          sj[u][j] = static_cast<double>(ld_nt(&a.s[a.k0 + j][ix[u]]));
          yj[u][j] = static_cast<double>(ld_nt(&a.y[a.k0 + j][ix[u]]));
        }
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        if (u == 1 && !two) break;
        const double wio = e.w[iox[u]], h = e.h0[iox[u]];
        const TS gs =
            static_cast<TS>(local_terms(xv[u], v[u], dfv[u], wio, e, acc[5 * KM + 7], acc[5 * KM + 8]));
        const double gn = static_cast<double>(gs);
        const TS ys = static_cast<TS>(gn - gv[u]);
        st_nt(&a.gt[ix[u]], gs);
        st_nt(&a.yn[ix[u]], ys);
        const double yv = static_cast<double>(ys);
        const double hy = h * yv, hg = h * gn;
#pragma unroll
        for (int j = 0; j < KM; ++j) {
          acc[j] += scv[u] * yj[u][j];
          acc[KM + 1 + j] += sj[u][j] * yv;
          acc[2 * KM + 3 + j] += sj[u][j] * gn;
          acc[3 * KM + 4 + j] += hy * yj[u][j];
          acc[4 * KM + 5 + j] += yj[u][j] * hg;
        }
        acc[KM] += scv[u] * yv;
        acc[2 * KM + 1] += scv[u] * scv[u];
        acc[2 * KM + 2] += yv * yv;
        acc[3 * KM + 3] += scv[u] * gn;
        acc[4 * KM + 4] += hy * yv;
        acc[5 * KM + 5] += yv * hg;
        acc[5 * KM + 6] += gn * hg;
        acc[5 * KM + 9] = fmax(acc[5 * KM + 9], fabs(hg));
      }
    }
  }
  store_partials<NA, true>(acc, part, nb);
}

// grid (rows, slots): res[slot kSlotRes + row] = fixed-order sum (or max) of
// the slot's row partials
__global__ __launch_bounds__(kFThreads) void slot_reduce_kernel(const double* __restrict__ part, int nb, int max_row,
                                                                double* __restrict__ res) {
  const int row = blockIdx.x;
  const bool is_max = row == max_row;
  part += (static_cast<long>(blockIdx.y) * kSlotRows + row) * nb;
  double v = 0.0;
  for (int i = threadIdx.x; i < nb; i += kFThreads) v = is_max ? fmax(v, part[i]) : v + part[i];
  __shared__ double red[kFThreads];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int st = kFThreads / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = is_max ? fmax(red[threadIdx.x], red[threadIdx.x + st])
                                                    : red[threadIdx.x] + red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) res[static_cast<long>(blockIdx.y) * kSlotRes + row] = red[0];
}

template <typename T>
__global__ __launch_bounds__(kFThreads) void fill_kernel(T* __restrict__ x, long n, T v) {
  for (long i = static_cast<long>(blockIdx.x) * kFThreads + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * kFThreads)
    x[i] = v;
}

template <typename T>
void fill(T* x, long n, T v, hipStream_t s) {
  hipLaunchKernelGGL(fill_kernel<T>, dim3(grid_for(n, kFThreads, 65536)), dim3(kFThreads), 0, s, x, n, v);
  MDFT_HIP_CHECK(hipGetLastError());
}

}  // namespace

// ---------------------------------------------------------------------------
struct Engine::Slot {
  long long id = 0;
  bool started = false;      // start point evaluated
  bool cur_in_slab = false;  // the current point is the slot's slab field
  LbfgsState st;
  LbfgsState::Trial tr;
  bool has_trial = false;
  double* x = nullptr;  // current point when not in the slab
  double* V = nullptr;
  void* g = nullptr;      // TS
  void* gt = nullptr;     // TS
  void* pairs = nullptr;  // 2 (m + 1) fields of TS
  void* mem = nullptr;    // one allocation: x | V (FP64) | g | gt | pairs (TS)
  double* out = nullptr;  // where the final point goes (device), or null
  explicit Slot(const LbfgsOptions& o) : st(o) {}
};

Engine::Engine(const FunctionalParams& p, const std::vector<double>& w, int capacity, const LbfgsOptions& o,
               bool fp32_pairs, ExcessBatch excess, bool fp32_fields)
    : p_(p), o_(o), excess_(std::move(excess)), capacity_(capacity), fp32_(fp32_pairs), f32f_(fp32_fields) {
  throw_if(capacity < 1, "engine: capacity must be at least 1");
  throw_if(fp32_fields && !fp32_pairs, "engine: FP32 fields take FP32 pairs");
  throw_if(o.m < 1 || o.m > kLbfgsMaxPairs, "lbfgs: m must be in 1 .. 8");
  throw_if(o.max_iter < 0 || o.max_backtrack < 1 || !(o.gtol >= 0) || !(o.c1 > 0 && o.c1 < 0.5) ||
               !(o.eps_f >= 0),
           "lbfgs: bad options");
  throw_if(p.N <= 0 || p.no <= 0 || static_cast<int>(w.size()) != p.no, "engine: bad sizes");
  throw_if(!(p.kT > 0) || !(p.dv > 0) || !(p.rho0 > 0), "engine: kT, dv and rho0 must be positive");
  n_ = p.N * p.no;
  const long want = (n_ + kFThreads * 16L - 1) / (kFThreads * 16L);
  nb_ = static_cast<int>(std::min<long>(std::max<long>(want, 1), kMaxBlocks));
  step_ = static_cast<int>((static_cast<long>(nb_) * kFThreads) % p.no);
  std::vector<double> tab(2 * static_cast<size_t>(p.no));
  for (int io = 0; io < p.no; ++io) {
    tab[io] = w[io];
    tab[p.no + io] = 1.0 / (2 * p.rho0 * p.dv * w[io] * p.kT);
  }
  const size_t ts = fp32_ ? sizeof(float) : sizeof(double);
  desc_bytes_ = static_cast<size_t>(capacity) * (fp32_ ? sizeof(SlotDesc<float>) : sizeof(SlotDesc<double>));
  try {
    MDFT_HIP_CHECK(hipStreamCreate(&stream_));
    MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&tab_), tab.size() * sizeof(double)));
    MDFT_HIP_CHECK(hipMemcpy(tab_, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&xslab_), capacity * n_ * sizeof(double)));
    MDFT_HIP_CHECK(hipMalloc(&dfslab_, capacity * n_ * (f32f_ ? sizeof(float) : sizeof(double))));
    if (f32f_) MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&xfslab_), capacity * n_ * sizeof(float)));
    MDFT_HIP_CHECK(hipMalloc(&zeros_, n_ * ts));
    MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&part_),
                             static_cast<size_t>(capacity) * kSlotRows * nb_ * sizeof(double)));
    MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&res_), static_cast<size_t>(capacity) * kSlotRes * sizeof(double)));
    MDFT_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_res_),
                                 static_cast<size_t>(capacity) * kSlotRes * sizeof(double)));
    MDFT_HIP_CHECK(hipMalloc(&desc_, desc_bytes_));
    MDFT_HIP_CHECK(hipHostMalloc(&host_desc_, desc_bytes_));
    std::memset(host_desc_, 0, desc_bytes_);
    // finite contents everywhere the batched passes may read
    fill(xslab_, capacity * n_, 1.0, stream_);
    if (f32f_) {
      fill(xfslab_, capacity * n_, 1.0f, stream_);
      fill(static_cast<float*>(dfslab_), capacity * n_, 0.0f, stream_);
    } else {
      fill(static_cast<double*>(dfslab_), capacity * n_, 0.0, stream_);
    }
    MDFT_HIP_CHECK(hipMemsetAsync(zeros_, 0, n_ * ts, stream_));
    MDFT_HIP_CHECK(hipStreamSynchronize(stream_));
  } catch (...) {
    release();
    throw;
  }
  // the slabs: 16 bytes per value either way (x, df: 8 + 8; x, x32, df32: 8 + 4 + 4)
  shared_bytes_ = (2 * static_cast<size_t>(capacity) * n_ + 2 * p.no) * sizeof(double) + n_ * ts +
                  static_cast<size_t>(capacity) * (kSlotRows * nb_ + kSlotRes) * sizeof(double) + desc_bytes_;
  for (int j = 0; j < capacity; ++j) slots_.push_back(nullptr);
}

Engine::~Engine() { release(); }

void Engine::release() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& kv : graphs_by_width_) (void)hipGraphExecDestroy(kv.second);
  graphs_by_width_.clear();
  for (int j = 0; j < static_cast<int>(slots_.size()); ++j)
    if (slots_[j]) release_slot(j);
  if (tab_) (void)hipFree(tab_);
  if (xslab_) (void)hipFree(xslab_);
  if (dfslab_) (void)hipFree(dfslab_);
  if (xfslab_) (void)hipFree(xfslab_);
  if (zeros_) (void)hipFree(zeros_);
  if (part_) (void)hipFree(part_);
  if (res_) (void)hipFree(res_);
  if (host_res_) (void)hipHostFree(host_res_);
  if (desc_) (void)hipFree(desc_);
  if (host_desc_) (void)hipHostFree(host_desc_);
  if (stream_) (void)hipStreamDestroy(stream_);
  tab_ = xslab_ = part_ = res_ = host_res_ = nullptr;
  xfslab_ = nullptr;
  zeros_ = dfslab_ = desc_ = host_desc_ = nullptr;
  stream_ = nullptr;
}

int Engine::active() const {
  int a = 0;
  for (const auto& s : slots_) a += s != nullptr;
  return a;
}

size_t Engine::job_bytes() const {
  const size_t ts = fp32_ ? sizeof(float) : sizeof(double);
  return 2 * static_cast<size_t>(n_) * sizeof(double) + 2 * static_cast<size_t>(o_.m + 2) * n_ * ts;
}

int Engine::admit(long long id, double* out_xi) {
  int j = 0;
  while (j < capacity_ && slots_[j]) ++j;
  throw_if(j == capacity_, "engine: no free slot");
  auto sl = std::make_unique<Slot>(o_);
  sl->id = id;
  sl->out = out_xi;
  MDFT_HIP_CHECK(hipMalloc(&sl->mem, job_bytes()));
  double* base = static_cast<double*>(sl->mem);
  const size_t ts = fp32_ ? sizeof(float) : sizeof(double);
  sl->x = base;
  sl->V = base + n_;
  sl->g = base + 2 * n_;
  sl->gt = reinterpret_cast<char*>(sl->g) + n_ * ts;
  sl->pairs = reinterpret_cast<char*>(sl->gt) + n_ * ts;
  fill(sl->x, n_, 1.0, stream_);  // the start point xi = 1 (Solvent::init_density)
  slots_[j] = std::move(sl);
  return j;
}

void Engine::cancel(int slot) {
  if (slot < 0 || slot >= capacity_ || !slots_[slot]) return;
  (void)hipStreamSynchronize(stream_);  // the slot's set-up work has ended
  release_slot(slot);
}

void Engine::vext_ready(int slot) {
  throw_if(slot < 0 || slot >= capacity_ || !slots_[slot], "engine: no such slot");
  if (start_vext_) xi_from_vext(slots_[slot]->V, n_, p_.kT, slots_[slot]->x, stream_);
}

double* Engine::vext_of(int slot) {
  throw_if(slot < 0 || slot >= capacity_ || !slots_[slot], "engine: no such slot");
  return slots_[slot]->V;
}

void Engine::release_slot(int j) {
  if (slots_[j] && slots_[j]->mem) (void)hipFree(slots_[j]->mem);
  slots_[j].reset();
}

std::vector<Engine::Done> Engine::take_done() {
  std::vector<Done> d;
  d.swap(done_);
  return d;
}

void Engine::enqueue(int nbatch) {
  const Elem e{tab_, tab_ + p_.no, p_.rho0, p_.kT, p_.dv, 2.0 * p_.prefactor / p_.N, n_, p_.no, step_};
  // the direction reads all m pair columns, the update pass m - 1
  const int KM = o_.m, KU = o_.m - 1, NA = 5 * KU + 10;
  dispatch_k(KM, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    if (f32f_) {
      hipLaunchKernelGGL((slot_direction_kernel<float, K, true>), dim3(nb_, nbatch), dim3(kFThreads), 0, stream_,
                         static_cast<const SlotDesc<float>*>(desc_), xslab_, xfslab_, e);
    } else if (fp32_) {
      hipLaunchKernelGGL((slot_direction_kernel<float, K, false>), dim3(nb_, nbatch), dim3(kFThreads), 0, stream_,
                         static_cast<const SlotDesc<float>*>(desc_), xslab_, nullptr, e);
    } else {
      hipLaunchKernelGGL((slot_direction_kernel<double, K, false>), dim3(nb_, nbatch), dim3(kFThreads), 0,
                         stream_, static_cast<const SlotDesc<double>*>(desc_), xslab_, nullptr, e);
    }
  });
  MDFT_HIP_CHECK(hipGetLastError());
  if (f32f_)
    excess_(nbatch, xfslab_, dfslab_, res_ + kSlotFf, kSlotRes, stream_);
  else
    excess_(nbatch, xslab_, dfslab_, res_ + kSlotFf, kSlotRes, stream_);
  dispatch_k(KU, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    if (f32f_) {
      hipLaunchKernelGGL((slot_update_kernel<float, K, float>), dim3(nb_, nbatch), dim3(kFThreads), 0, stream_,
                         static_cast<const SlotDesc<float>*>(desc_), xslab_, static_cast<const float*>(dfslab_), e,
                         part_, nb_);
    } else if (fp32_) {
      hipLaunchKernelGGL((slot_update_kernel<float, K, double>), dim3(nb_, nbatch), dim3(kFThreads), 0, stream_,
                         static_cast<const SlotDesc<float>*>(desc_), xslab_, static_cast<const double*>(dfslab_),
                         e, part_, nb_);
    } else {
      hipLaunchKernelGGL((slot_update_kernel<double, K, double>), dim3(nb_, nbatch), dim3(kFThreads), 0, stream_,
                         static_cast<const SlotDesc<double>*>(desc_), xslab_, static_cast<const double*>(dfslab_),
                         e, part_, nb_);
    }
  });
  MDFT_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(slot_reduce_kernel, dim3(NA, nbatch), dim3(kFThreads), 0, stream_, part_, nb_, 5 * KU + 9,
                     res_);
  MDFT_HIP_CHECK(hipGetLastError());
}

namespace {
template <typename TS>
void fill_desc(SlotDesc<TS>& d, int mode, const LbfgsState::Trial* tr, const std::vector<int>* order,
               const double* src, double* save, const double* V, TS* g, TS* gt, TS* pairs, long n,
               const TS* zeros, int m) {
  d = SlotDesc<TS>();
  d.mode = mode;
  d.src = src;
  d.save = save;
  d.V = V;
  d.g = g;
  d.gt = gt;
  for (int j = 0; j < m; ++j) {
    d.s[j] = zeros;
    d.y[j] = zeros;
  }
  if (!tr) return;
  d.k = tr->k;
  d.k0 = tr->k0;
  d.cg = tr->cg;
  d.gamma = tr->gamma;
  d.t = tr->t;
  for (int j = 0; j < tr->k; ++j) {
    d.cs[j] = tr->cs[j];
    d.cq[j] = tr->cq[j];
    d.s[j] = pairs + 2 * static_cast<long>((*order)[j]) * n;
    d.y[j] = pairs + (2 * static_cast<long>((*order)[j]) + 1) * n;
  }
  d.sc = pairs + 2 * static_cast<long>(tr->cand) * n;
  d.yn = pairs + (2 * static_cast<long>(tr->cand) + 1) * n;
}
}  // namespace

int Engine::tick() {
  begin_tick();
  return end_tick();
}

void Engine::begin_tick() {
  int hi = 0;
  for (int j = 0; j < capacity_; ++j)
    if (slots_[j]) hi = j + 1;
  width_ = hi;
  if (hi == 0) return;
  // the slots' descriptors
  for (int j = 0; j < hi; ++j) {
    Slot* sl = slots_[j].get();
    const int mode = !sl ? kModeIdle : !sl->started ? kModeStart : kModeTrial;
    const double* src = nullptr;
    double* save = nullptr;
    if (mode == kModeStart) src = sl->x;
    if (mode == kModeTrial) {
      src = sl->cur_in_slab ? xslab_ + j * n_ : sl->x;
      save = sl->cur_in_slab ? sl->x : nullptr;
      sl->cur_in_slab = false;  // the slab now receives the trial; x holds the current point
    }
    const LbfgsState::Trial* tr = mode == kModeTrial ? &sl->tr : nullptr;
    const std::vector<int>* order = sl ? &sl->st.order() : nullptr;
    if (fp32_)
      fill_desc(static_cast<SlotDesc<float>*>(host_desc_)[j], mode, tr, order, src, save, sl ? sl->V : nullptr,
                sl ? static_cast<float*>(sl->g) : nullptr, sl ? static_cast<float*>(sl->gt) : nullptr,
                sl ? static_cast<float*>(sl->pairs) : nullptr, n_,
                static_cast<const float*>(zeros_), o_.m);
    else
      fill_desc(static_cast<SlotDesc<double>*>(host_desc_)[j], mode, tr, order, src, save, sl ? sl->V : nullptr,
                sl ? static_cast<double*>(sl->g) : nullptr, sl ? static_cast<double*>(sl->gt) : nullptr,
                sl ? static_cast<double*>(sl->pairs) : nullptr, n_,
                static_cast<const double*>(zeros_), o_.m);
  }
  const size_t dsz = fp32_ ? sizeof(SlotDesc<float>) : sizeof(SlotDesc<double>);
  MDFT_HIP_CHECK(hipMemcpyAsync(desc_, host_desc_, hi * dsz, hipMemcpyHostToDevice, stream_));
  // the passes: eager until this width has run once, then its graph
  auto g = graphs_by_width_.find(hi);
  if (graphs_ && g != graphs_by_width_.end()) {
    MDFT_HIP_CHECK(hipGraphLaunch(g->second, stream_));
  } else if (graphs_ && eager_runs_[hi] >= 1) {
    hipGraph_t graph = nullptr;
    MDFT_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      enqueue(hi);
    } catch (...) {
      (void)hipStreamEndCapture(stream_, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    MDFT_HIP_CHECK(hipStreamEndCapture(stream_, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MDFT_HIP_CHECK(err);
    graphs_by_width_[hi] = exec;
    MDFT_HIP_CHECK(hipGraphLaunch(exec, stream_));
  } else {
    enqueue(hi);
    ++eager_runs_[hi];
  }
  MDFT_HIP_CHECK(hipMemcpyAsync(host_res_, res_, static_cast<size_t>(hi) * kSlotRes * sizeof(double),
                                hipMemcpyDeviceToHost, stream_));
}

int Engine::end_tick() {
  const int hi = width_;
  width_ = 0;
  if (hi == 0) return 0;
  MDFT_HIP_CHECK(hipStreamSynchronize(stream_));

  // the decisions (update-pass rows with m - 1 pair columns)
  const int KM = o_.m - 1;
  int finished = 0;
  for (int j = 0; j < hi; ++j) {
    Slot* sl = slots_[j].get();
    if (!sl) continue;
    const double* h = host_res_ + static_cast<long>(j) * kSlotRes;
    double F[4];
    F[1] = p_.kT * p_.rho0 * p_.dv * h[5 * KM + 7];
    F[2] = p_.rho0 * p_.dv * h[5 * KM + 8];
    F[3] = p_.dv / p_.N * h[kSlotFf];
    F[0] = F[1] + F[2] + F[3];
    if (!sl->started) {
      sl->st.start(F, h[5 * KM + 6], h[5 * KM + 9]);
      sl->started = true;
    } else if (sl->st.judge(F, h, KM)) {
      std::swap(sl->g, sl->gt);
      sl->cur_in_slab = true;
    }
    sl->has_trial = sl->st.next(sl->tr);
    if (!sl->has_trial) {
      if (sl->out) {
        MDFT_HIP_CHECK(hipMemcpyAsync(sl->out, sl->cur_in_slab ? xslab_ + j * n_ : sl->x, n_ * sizeof(double),
                                      hipMemcpyDeviceToDevice, stream_));
        MDFT_HIP_CHECK(hipStreamSynchronize(stream_));
      }
      done_.push_back(Done{sl->id, sl->st.result()});
      release_slot(j);
      ++finished;
    }
  }
  return finished;
}

// ---------------------------------------------------------------------------
Server::Server(EngineFactory make, Footprint footprint, double budget_bytes, int mrso)
    : make_(std::move(make)), footprint_(std::move(footprint)), budget_(budget_bytes), mrso_(mrso) {
  throw_if(!(budget_bytes > 0), "server: the HBM budget must be positive");
}

long long Server::submit(const Shape& shape, const double* solute, int nsolute, const double* solvent, int nsolvent,
                         double vmax, double* out_xi) {
  throw_if(nsolute < 0 || nsolvent <= 0 || (nsolute > 0 && !solute) || !solvent, "server: bad sites");
  for (int i = 0; i < 6 * nsolute; ++i) throw_if(!std::isfinite(solute[i]), "server: solute sites must be finite");
  for (int i = 0; i < 6 * nsolvent; ++i) throw_if(!std::isfinite(solvent[i]), "server: solvent sites must be finite");
  throw_if(!(vmax > 0), "server: vmax must be positive (inf: no cap)");
  Request r;
  r.id = next_id_++;
  r.shape = shape;
  r.solute.assign(solute, solute + 6 * nsolute);
  r.solvent.assign(solvent, solvent + 6 * nsolvent);
  r.vmax = vmax;
  r.out_xi = out_xi;
  r.t_submit = now_seconds();
  queue_.push_back(std::move(r));
  return queue_.back().id;
}

int Server::active() const { return static_cast<int>(running_.size()); }

void Server::admit() {
  for (auto it = queue_.begin(); it != queue_.end();) {
    const Request& rq = *it;
    auto fp = footprint_(rq.shape);
    auto e = engines_.find(rq.shape);
    const bool have = e != engines_.end();
    const double need = static_cast<double>(fp.first) + (have ? 0.0 : static_cast<double>(fp.second));
    if (static_cast<double>(fp.first + fp.second) > budget_) {  // can never fit: reject
      Result res{rq.id, LbfgsResult(), now_seconds() - rq.t_submit, now_seconds() - rq.t_submit, true, ""};
      res.r.status = 3;
      results_.push_back(res);
      it = queue_.erase(it);
      continue;
    }
    if (have && !e->second.engine->has_free_slot()) {
      ++it;
      continue;
    }
    if (in_use_ + need > budget_) {
      // drop idle engines of other shapes to make room
      for (auto x = engines_.begin(); x != engines_.end() && in_use_ + need > budget_;) {
        if (x->second.engine->active() == 0 && x != e) {
          in_use_ -= static_cast<double>(x->second.table_bytes);
          x = engines_.erase(x);
        } else {
          ++x;
        }
      }
      e = engines_.find(rq.shape);
      if (in_use_ + need > budget_) {
        ++it;
        continue;
      }
    }
    // A request whose set-up fails (its engine cannot be built, its V_ext
    // cannot be computed) ends with status 5 and the error text; it leaves
    // the queue, so it cannot stall the requests behind it or the running
    // solves.
    int slot = -1;
    Engine* eng = nullptr;
    try {
      if (e == engines_.end()) {
        Live lv;
        lv.engine = make_(rq.shape, lv.table_bytes);
        in_use_ += static_cast<double>(lv.table_bytes);
        e = engines_.emplace(rq.shape, std::move(lv)).first;
      }
      eng = e->second.engine.get();
      slot = eng->admit(rq.id, rq.out_xi);
      // V_ext of the solute into the slot, on the engine's stream
      compute_vext(rq.shape.nx, rq.shape.ny, rq.shape.nz, rq.shape.lx, rq.shape.ly, rq.shape.lz,
                   make_angular_grid(rq.shape.mmax, mrso_), rq.solute.data(), static_cast<int>(rq.solute.size() / 6),
                   rq.solvent.data(), static_cast<int>(rq.solvent.size() / 6), rq.vmax, eng->vext_of(slot),
                   eng->stream());
      eng->vext_ready(slot);
    } catch (const std::exception& ex) {
      // a failed runtime call (an allocation) stays the thread's last HIP
      // error; clear it so that the next launch check does not report it
      (void)hipGetLastError();
      if (eng && slot >= 0) eng->cancel(slot);
      const double t = now_seconds();
      Result res{rq.id, LbfgsResult(), t - rq.t_submit, t - rq.t_submit, true, ex.what()};
      res.r.status = 5;
      results_.push_back(res);
      it = queue_.erase(it);
      continue;
    }
    in_use_ += static_cast<double>(eng->job_bytes());
    e->second.admitted_at[rq.id] = now_seconds();
    running_[rq.id] = rq;
    it = queue_.erase(it);
  }
}

void Server::step() {
  admit();
  // every engine's tick in flight on its own stream, then their decisions
  for (auto& kv : engines_)
    if (kv.second.engine->active() > 0) kv.second.engine->begin_tick();
  for (auto& kv : engines_) {
    Live& lv = kv.second;
    lv.engine->end_tick();
    for (auto& d : lv.engine->take_done()) {
      const Request& rq = running_[d.id];
      const double t = now_seconds();
      results_.push_back(Result{d.id, d.r, lv.admitted_at[d.id] - rq.t_submit, t - rq.t_submit, false, ""});
      lv.admitted_at.erase(d.id);
      running_.erase(d.id);
      in_use_ -= static_cast<double>(lv.engine->job_bytes());
    }
  }
}

bool Server::poll(Result& out) {
  if (results_.empty()) return false;
  out = results_.front();
  results_.pop_front();
  return true;
}

}  // namespace mdft
