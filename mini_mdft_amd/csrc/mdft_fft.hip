// Batched unnormalised 3D c2c FFT for gfx950: mixed-radix Stockham
// transforms of tiles staged in LDS, two launches per 3D transform when a
// (y, z) plane fits in LDS (one pass for z and y together, one for x),
// otherwise one launch per axis.
//
// Replaces the kokkos-fft plans of common/MDFT_Convolution.hpp:198-208 and
// their execution at :328-329 (forward) and :546-547 (backward), both with
// Normalization::none.  A tile holds W lines of length n: for the contiguous
// axis (z) the W lines are W consecutive rows; for the strided axes (y, x) they
// are W neighbouring z columns, so every global access is a run of W (or W*n)
// consecutive complex values.  Each Stockham stage reads one LDS buffer and
// writes the other (radix-4/2/3/5 butterflies, a generic odd-prime DFT up to
// 13), with the twiddles e^{-2 pi i t/n} read from an LDS copy of an FP64
// host table.
#include "mdft_fft.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <utility>

namespace mdft {

namespace {

constexpr int kFftThreads = 256;
constexpr int kMaxRadix = 13;

template <typename T>
__device__ __forceinline__ cplx<T> twiddle(const cplx<T>* tw, int t, int sign) {
  cplx<T> w = tw[t];
  if (sign > 0) w.y = -w.y;
  return w;
}

// x * w with one fused and one rounded product, written out so that every
// kernel (run-time or compile-time stages, any tiling) rounds the same way
// whatever the compiler would contract
template <typename T>
__device__ __forceinline__ cplx<T> tw_mul(cplx<T> x, cplx<T> w) {
  return mkc<T>(fma(x.x, w.x, -(x.y * w.y)), fma(x.x, w.y, x.y * w.x));
}

// in-place DFT of R values; sign -1 forward, +1 backward
template <typename T>
__device__ __forceinline__ void dft_small(cplx<T>* v, int R, int sign, const cplx<T>* tw, int n) {
  if (R == 2) {
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
  } else if (R == 4) {
    cplx<T> s02 = cadd<T>(v[0], v[2]), d02 = csub<T>(v[0], v[2]);
    cplx<T> s13 = cadd<T>(v[1], v[3]), d13 = csub<T>(v[1], v[3]);
    // W = e^{sign 2 pi i/4} = sign * i
    cplx<T> wd = mkc<T>(-static_cast<T>(sign) * d13.y, static_cast<T>(sign) * d13.x);
    v[0] = cadd<T>(s02, s13);
    v[2] = csub<T>(s02, s13);
    v[1] = cadd<T>(d02, wd);
    v[3] = csub<T>(d02, wd);
  } else if (R == 3) {
#pragma clang fp contract(off)
    const T h = static_cast<T>(0.86602540378443864676);  // sqrt(3)/2
    cplx<T> a = v[0], t = cadd<T>(v[1], v[2]), u = csub<T>(v[1], v[2]);
    cplx<T> m = mkc<T>(a.x - static_cast<T>(0.5) * t.x, a.y - static_cast<T>(0.5) * t.y);
    // sign * i * h * u
    cplx<T> r = mkc<T>(-static_cast<T>(sign) * h * u.y, static_cast<T>(sign) * h * u.x);
    v[0] = cadd<T>(a, t);
    v[1] = cadd<T>(m, r);
    v[2] = csub<T>(m, r);
  } else {
    cplx<T> out[kMaxRadix];
    const int step = n / R;
    for (int s = 0; s < R; ++s) {
      cplx<T> acc = v[0];
      for (int r = 1; r < R; ++r) acc = cfma<T>(acc, v[r], twiddle<T>(tw, ((r * s) % R) * step, sign));
      out[s] = acc;
    }
    for (int s = 0; s < R; ++s) v[s] = out[s];
  }
}

// DFT of R values held in registers (R fixed at compile time)
template <typename T, int R>
__device__ __forceinline__ void dft_fixed(cplx<T> (&v)[R], int sign, const cplx<T>* tw, int n) {
  if constexpr (R == 2 || R == 3 || R == 4) {
    dft_small<T>(v, R, sign, tw, n);
  } else {
    cplx<T> out[R];
    const int step = n / R;
#pragma unroll
    for (int s = 0; s < R; ++s) {
      cplx<T> acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) acc = cfma<T>(acc, v[r], twiddle<T>(tw, ((r * s) % R) * step, sign));
      out[s] = acc;
    }
#pragma unroll
    for (int s = 0; s < R; ++s) v[s] = out[s];
  }
}

// One Stockham stage (radix R, Ns = product of the earlier radices) over the
// wvalid lines of a tile: butterfly j of a line reads x[j + r n/R], applies the
// twiddles W_n^{r (j mod Ns) n/(Ns R)}, and writes its outputs at
// (j / Ns) Ns R + (j mod Ns) + r Ns.
template <typename T, int R>
__device__ __forceinline__ void stockham_stage(const cplx<T>* src, cplx<T>* dst, const cplx<T>* tw, int n, int ld,
                                               int Ns, int wvalid, int sign, int tid, int nthreads) {
  const int nb = n / R;
  const int tstep = n / (Ns * R);
  for (int beta = tid; beta < wvalid * nb; beta += nthreads) {
    const int w = beta / nb, j = beta - w * nb;
    const int k = j % Ns;
    const cplx<T>* in = src + w * ld + j;
    cplx<T> v[R];
    v[0] = in[0];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = tw_mul<T>(in[r * nb], twiddle<T>(tw, r * k * tstep, sign));
    dft_fixed<T, R>(v, sign, tw, n);
    cplx<T>* out = dst + w * ld + (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[r * Ns] = v[r];
  }
}

// The same stage for an odd prime radix 7..13 known only at run time (rare
// lengths): one thread per butterfly output, O(R) LDS reads each, no private
// arrays.
template <typename T>
__device__ __forceinline__ void stockham_stage_generic(const cplx<T>* src, cplx<T>* dst, const cplx<T>* tw, int n,
                                                       int ld, int Ns, int R, int wvalid, int sign, int tid,
                                                       int nthreads) {
  const int nb = n / R;
  const int tstep = n / (Ns * R);
  const int step = n / R;
  for (int it = tid; it < wvalid * nb * R; it += nthreads) {
    const int s = it % R, beta = it / R;
    const int w = beta / nb, j = beta - w * nb;
    const int k = j % Ns;
    const cplx<T>* in = src + w * ld + j;
    cplx<T> acc = in[0];
    for (int r = 1; r < R; ++r) {
      const cplx<T> x = tw_mul<T>(in[r * nb], twiddle<T>(tw, r * k * tstep, sign));
      acc = cfma<T>(acc, x, twiddle<T>(tw, ((r * s) % R) * step, sign));
    }
    dst[w * ld + (j / Ns) * Ns * R + k + s * Ns] = acc;
  }
}

// The same Stockham stage over lines of a 2D LDS tile with a line stride ls
// and an element stride es: element j of line w sits at w ls + j es.  With
// W_FAST, consecutive threads take consecutive lines (the column pass of a
// plane: lines are neighbouring columns, so the accesses of a wave are
// consecutive).
template <typename T, int R, bool W_FAST>
__device__ __forceinline__ void stockham_stage_2d(const cplx<T>* src, cplx<T>* dst, const cplx<T>* tw, int n, int ls,
                                                  int es, int Ns, int nlines, int sign, int tid, int nthreads) {
  const int nb = n / R;
  const int tstep = n / (Ns * R);
  for (int beta = tid; beta < nlines * nb; beta += nthreads) {
    int w, j;
    if (W_FAST) {
      j = beta / nlines;
      w = beta - j * nlines;
    } else {
      w = beta / nb;
      j = beta - w * nb;
    }
    const int k = j % Ns;
    const cplx<T>* in = src + w * ls + j * es;
    cplx<T> v[R];
    v[0] = in[0];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = tw_mul<T>(in[r * nb * es], twiddle<T>(tw, r * k * tstep, sign));
    dft_fixed<T, R>(v, sign, tw, n);
    cplx<T>* out = dst + w * ls + ((j / Ns) * Ns * R + k) * es;
#pragma unroll
    for (int r = 0; r < R; ++r) out[r * Ns * es] = v[r];
  }
}

// the stages of one axis over a 2D LDS tile (radices 2, 3, 4, 5), ping-pong
// between a and b; returns the buffer holding the result
template <typename T, bool W_FAST>
__device__ cplx<T>* stockham_2d(cplx<T>* a, cplx<T>* b, const cplx<T>* tw, const FftStages& st, int ls, int es,
                                int nlines, int sign, int tid, int nthreads) {
  int Ns = 1;
  for (int s = 0; s < st.nstages; ++s) {
    const int R = st.radix[s];
    switch (R) {
      case 4: stockham_stage_2d<T, 4, W_FAST>(a, b, tw, st.n, ls, es, Ns, nlines, sign, tid, nthreads); break;
      case 2: stockham_stage_2d<T, 2, W_FAST>(a, b, tw, st.n, ls, es, Ns, nlines, sign, tid, nthreads); break;
      case 3: stockham_stage_2d<T, 3, W_FAST>(a, b, tw, st.n, ls, es, Ns, nlines, sign, tid, nthreads); break;
      default: stockham_stage_2d<T, 5, W_FAST>(a, b, tw, st.n, ls, es, Ns, nlines, sign, tid, nthreads); break;
    }
    __syncthreads();
    cplx<T>* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

constexpr int kPlaneThreads = 512;

// z and y transforms of one (batch, x) plane of ny x nz contiguous values in
// one pass: the plane is staged in LDS (rows padded to nz + 1), transformed
// along its rows, then along its columns, and written back.  Radices 2-5 only.
template <typename T>
__global__ __launch_bounds__(kPlaneThreads) void fft_plane_kernel(cplx<T>* __restrict__ data, FftStages sz,
                                                                  FftStages sy, const cplx<T>* __restrict__ twz_g,
                                                                  const cplx<T>* __restrict__ twy_g, int sign) {
  extern __shared__ unsigned char lds_raw[];
  const int nz = sz.n, ny = sy.n, ld = nz + 1;
  cplx<T>* bufA = reinterpret_cast<cplx<T>*>(lds_raw);
  cplx<T>* bufB = bufA + ny * ld;
  cplx<T>* twz = bufB + ny * ld;
  cplx<T>* twy = twz + nz;
  const int tid = threadIdx.x;
  const int nth = blockDim.x;
  for (int i = tid; i < nz; i += nth) twz[i] = twz_g[i];
  for (int i = tid; i < ny; i += nth) twy[i] = twy_g[i];
  cplx<T>* plane = data + static_cast<long>(blockIdx.x) * ny * nz;
  for (int t = tid; t < ny * nz; t += nth) {
    const int y = t / nz, z = t - y * nz;
    bufA[y * ld + z] = plane[t];
  }
  __syncthreads();
  cplx<T>* a = stockham_2d<T, false>(bufA, bufB, twz, sz, ld, 1, ny, sign, tid, nth);
  cplx<T>* b = a == bufA ? bufB : bufA;
  a = stockham_2d<T, true>(a, b, twy, sy, 1, ld, nz, sign, tid, nth);
  for (int t = tid; t < ny * nz; t += nth) {
    const int y = t / nz, z = t - y * nz;
    plane[t] = a[y * ld + z];
  }
}

// one fused z/y pass over `planes` planes: the compile-time kernel for square
// planes of the compile-time lengths up to 96 x 96, else fft_plane_kernel
template <typename T>
void launch_plane(cplx<T>* data, const FftPlan1D& pz, const FftPlan1D& py, long planes, int threads, int sign,
                  hipStream_t s);

bool small_radices(const FftStages& st) {
  for (int s = 0; s < st.nstages; ++s)
    if (st.radix[s] > 5) return false;
  return true;
}

// LDS of the plane kernel: two padded planes and both twiddle tables
template <typename T>
size_t plane_lds(int ny, int nz) {
  return (2 * static_cast<size_t>(ny) * (nz + 1) + ny + nz) * sizeof(cplx<T>);
}
// the fused z/y pass is used untuned when two planes fit in half of the
// 160 KB LDS (at least two blocks of 8 waves per CU); the tuner also tries it
// up to the whole LDS
constexpr size_t kPlaneLdsMax = 80 * 1024;
constexpr size_t kLdsMax = 160 * 1024;

// One block transforms a tile of W lines of length n.
//   STRIDED = false: line l occupies data[l*n .. l*n+n); the tile is lines
//                    [tile*W, tile*W + W) (clipped at nlines).
//   STRIDED = true : element e of line (o, i) sits at data[(o*n + e)*S + i];
//                    tile = (o, chunk) covers i in [chunk*W, chunk*W + W).
template <typename T, bool STRIDED>
__global__ __launch_bounds__(kFftThreads) void fft_axis_kernel(cplx<T>* __restrict__ data, FftStages st,
                                                               const cplx<T>* __restrict__ tw_g, int sign,
                                                               long nlines_or_outer, long S, int W) {
  extern __shared__ unsigned char lds_raw[];
  const int n = st.n;
  const int ld = n + 1;  // padded row: avoids bank conflicts on column access
  cplx<T>* bufA = reinterpret_cast<cplx<T>*>(lds_raw);
  cplx<T>* bufB = bufA + W * ld;
  cplx<T>* tw = bufB + W * ld;
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += blockDim.x) tw[i] = tw_g[i];

  long base;
  int wvalid;
  if (!STRIDED) {
    const long l0 = static_cast<long>(blockIdx.x) * W;
    wvalid = static_cast<int>(min(static_cast<long>(W), nlines_or_outer - l0));
    base = l0 * n;
    for (int t = tid; t < wvalid * n; t += blockDim.x) {
      const int w = t / n, e = t - w * n;
      bufA[w * ld + e] = data[base + t];
    }
  } else {
    const long nchunk = (S + W - 1) / W;
    const long o = blockIdx.x / nchunk;
    const long i0 = (blockIdx.x - o * nchunk) * W;
    wvalid = static_cast<int>(min(static_cast<long>(W), S - i0));
    base = o * n * S + i0;
    for (int t = tid; t < wvalid * n; t += blockDim.x) {
      const int e = t / wvalid, w = t - e * wvalid;
      bufA[w * ld + e] = data[base + static_cast<long>(e) * S + w];
    }
  }
  __syncthreads();

  cplx<T>* src = bufA;
  cplx<T>* dst = bufB;
  int Ns = 1;
  for (int s = 0; s < st.nstages; ++s) {
    const int R = st.radix[s];
    switch (R) {
      case 4:
        stockham_stage<T, 4>(src, dst, tw, n, ld, Ns, wvalid, sign, tid, blockDim.x);
        break;
      case 2:
        stockham_stage<T, 2>(src, dst, tw, n, ld, Ns, wvalid, sign, tid, blockDim.x);
        break;
      case 3:
        stockham_stage<T, 3>(src, dst, tw, n, ld, Ns, wvalid, sign, tid, blockDim.x);
        break;
      case 5:
        stockham_stage<T, 5>(src, dst, tw, n, ld, Ns, wvalid, sign, tid, blockDim.x);
        break;
      default:
        stockham_stage_generic<T>(src, dst, tw, n, ld, Ns, R, wvalid, sign, tid, blockDim.x);
        break;
    }
    __syncthreads();
    cplx<T>* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }

  if (!STRIDED) {
    for (int t = tid; t < wvalid * n; t += blockDim.x) {
      const int w = t / n, e = t - w * n;
      data[base + t] = src[w * ld + e];
    }
  } else {
    for (int t = tid; t < wvalid * n; t += blockDim.x) {
      const int e = t / wvalid, w = t - e * wvalid;
      data[base + static_cast<long>(e) * S + w] = src[w * ld + e];
    }
  }
}

// Compile-time kernels for the lengths the grids of the served mixes use
// (16, 32, 48, 64, 80, 96, 128), with the radix schedule of FftPlan1D (4s
// first, then 2, 3, 5).  Each stage is stockham_stage_2d with the length,
// strides, line count and thread count known at compile time, so the index
// arithmetic reduces to shifts, masks and constant divisions and the
// butterfly loop unrolls; the arithmetic of every butterfly (LDS twiddles,
// run-time sign, explicitly rounded products, operation order) is that of the
// run-time stages, so results are bitwise the same.
constexpr bool ct_length(int n) {
  return n == 16 || n == 32 || n == 48 || n == 64 || n == 80 || n == 96 || n == 128;
}
template <int N>
constexpr int ct_nstages() {
  constexpr int ps[4] = {4, 2, 3, 5};
  int r = N, c = 0;
  for (int p : ps)
    while (r % p == 0) {
      r /= p;
      ++c;
    }
  return c;
}
template <int N>
constexpr int ct_radix(int s) {
  constexpr int ps[4] = {4, 2, 3, 5};
  int r = N, c = 0;
  for (int p : ps)
    while (r % p == 0) {
      if (c == s) return p;
      r /= p;
      ++c;
    }
  return 0;
}
// the tile widths the FFT tuner tries for length N (values per tile 256,
// 512, 1024, 2048; the default is the third)
template <int N>
constexpr int tuner_w(int i) {
  constexpr int vals[4] = {256, 512, 1024, 2048};
  const int w = vals[i] / N;
  return w < 1 ? 1 : (w > 64 ? 64 : w);
}

template <typename T, int R, int N, int NS, int LS, int ES, int NLINES, int NTH, bool W_FAST>
__device__ __forceinline__ void ct_stage_2d(const cplx<T>* src, cplx<T>* dst, const cplx<T>* tw, int sign, int tid) {
  constexpr int NB = N / R, TSTEP = N / (NS * R), TOTAL = NLINES * NB;
  constexpr int ITERS = (TOTAL + NTH - 1) / NTH;
#pragma unroll
  for (int i = 0; i < ITERS; ++i) {
    const int beta = tid + i * NTH;
    if (TOTAL % NTH == 0 || beta < TOTAL) {
      int w, j;
      if constexpr (W_FAST) {
        j = beta / NLINES;
        w = beta - j * NLINES;
      } else {
        w = beta / NB;
        j = beta - w * NB;
      }
      const int k = j % NS;
      const cplx<T>* in = src + w * LS + j * ES;
      cplx<T> v[R];
      v[0] = in[0];
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = tw_mul<T>(in[r * NB * ES], twiddle<T>(tw, r * k * TSTEP, sign));
      dft_fixed<T, R>(v, sign, tw, N);
      cplx<T>* out = dst + w * LS + ((j / NS) * NS * R + k) * ES;
#pragma unroll
      for (int r = 0; r < R; ++r) out[r * NS * ES] = v[r];
    }
  }
}

// stages S.. of one axis over an LDS tile, ping-pong between a and b; the
// result ends in a when the remaining stage count is even, else in b
template <typename T, int N, int LS, int ES, int NLINES, int NTH, bool W_FAST, int S = 0, int NS = 1>
__device__ __forceinline__ void ct_stockham_2d(cplx<T>* a, cplx<T>* b, const cplx<T>* tw, int sign, int tid) {
  if constexpr (S < ct_nstages<N>()) {
    constexpr int R = ct_radix<N>(S);
    ct_stage_2d<T, R, N, NS, LS, ES, NLINES, NTH, W_FAST>(a, b, tw, sign, tid);
    __syncthreads();
    ct_stockham_2d<T, N, LS, ES, NLINES, NTH, W_FAST, S + 1, NS * R>(b, a, tw, sign, tid);
  }
}

// fft_plane_kernel for an N x N plane with NTH threads
template <typename T, int N, int NTH>
__global__ __launch_bounds__(NTH) void fft_plane_ct_kernel(cplx<T>* __restrict__ data, const cplx<T>* __restrict__ twz_g,
                                                           const cplx<T>* __restrict__ twy_g, int sign) {
  extern __shared__ unsigned char lds_raw[];
  constexpr int LD = N + 1, NN = N * N, ITERS = (NN + NTH - 1) / NTH;
  cplx<T>* bufA = reinterpret_cast<cplx<T>*>(lds_raw);
  cplx<T>* bufB = bufA + N * LD;
  cplx<T>* twz = bufB + N * LD;
  cplx<T>* twy = twz + N;
  const int tid = threadIdx.x;
  if (tid < N) {
    twz[tid] = twz_g[tid];
    twy[tid] = twy_g[tid];
  }
  cplx<T>* plane = data + static_cast<long>(blockIdx.x) * NN;
#pragma unroll
  for (int i = 0; i < ITERS; ++i) {
    const int t = tid + i * NTH;
    if (NN % NTH == 0 || t < NN) bufA[(t / N) * LD + t % N] = plane[t];
  }
  __syncthreads();
  // z along the rows, then y along the columns
  ct_stockham_2d<T, N, LD, 1, N, NTH, false>(bufA, bufB, twz, sign, tid);
  constexpr bool z_in_a = ct_nstages<N>() % 2 == 0;
  cplx<T>* a = z_in_a ? bufA : bufB;
  cplx<T>* b = z_in_a ? bufB : bufA;
  ct_stockham_2d<T, N, 1, LD, N, NTH, true>(a, b, twy, sign, tid);
  const cplx<T>* res = bufA;  // two axes of the same stage count end where they began
#pragma unroll
  for (int i = 0; i < ITERS; ++i) {
    const int t = tid + i * NTH;
    if (NN % NTH == 0 || t < NN) plane[t] = res[(t / N) * LD + t % N];
  }
}

// fft_axis_kernel for length N and full tiles of W lines (W one of the tuner's widths)
template <typename T, bool STRIDED, int N, int W>
__global__ __launch_bounds__(kFftThreads) void fft_axis_ct_kernel(cplx<T>* __restrict__ data,
                                                                  const cplx<T>* __restrict__ tw_g, int sign, long S) {
  extern __shared__ unsigned char lds_raw[];
  constexpr int LD = N + 1, NTH = kFftThreads, WN = W * N, ITERS = (WN + NTH - 1) / NTH;
  cplx<T>* bufA = reinterpret_cast<cplx<T>*>(lds_raw);
  cplx<T>* bufB = bufA + W * LD;
  cplx<T>* tw = bufB + W * LD;
  const int tid = threadIdx.x;
  if (tid < N) tw[tid] = tw_g[tid];
  long base;
  if constexpr (!STRIDED) {
    base = static_cast<long>(blockIdx.x) * (W * N);
#pragma unroll
    for (int i = 0; i < ITERS; ++i) {
      const int t = tid + i * NTH;
      if (WN % NTH == 0 || t < WN) bufA[(t / N) * LD + t % N] = data[base + t];
    }
  } else {
    const long nchunk = S / W;
    const long o = blockIdx.x / nchunk;
    base = o * N * S + (blockIdx.x - o * nchunk) * W;
#pragma unroll
    for (int i = 0; i < ITERS; ++i) {
      const int t = tid + i * NTH;
      const int e = t / W, w = t % W;
      if (WN % NTH == 0 || t < WN) bufA[w * LD + e] = data[base + static_cast<long>(e) * S + w];
    }
  }
  __syncthreads();
  ct_stockham_2d<T, N, LD, 1, W, NTH, false>(bufA, bufB, tw, sign, tid);
  const cplx<T>* res = ct_nstages<N>() % 2 == 0 ? bufA : bufB;
  if constexpr (!STRIDED) {
#pragma unroll
    for (int i = 0; i < ITERS; ++i) {
      const int t = tid + i * NTH;
      if (WN % NTH == 0 || t < WN) data[base + t] = res[(t / N) * LD + t % N];
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITERS; ++i) {
      const int t = tid + i * NTH;
      const int e = t / W, w = t % W;
      if (WN % NTH == 0 || t < WN) data[base + static_cast<long>(e) * S + w] = res[w * LD + e];
    }
  }
}

// the compile-time kernel of (n, W), if any: f(integral_constant N, W)
template <typename F>
bool dispatch_axis_ct(int n, int W, F&& f) {
  auto by_w = [&](auto nn) {
    constexpr int N = decltype(nn)::value;
    if (W == tuner_w<N>(0)) f(nn, std::integral_constant<int, tuner_w<N>(0)>{});
    else if (W == tuner_w<N>(1)) f(nn, std::integral_constant<int, tuner_w<N>(1)>{});
    else if (W == tuner_w<N>(2)) f(nn, std::integral_constant<int, tuner_w<N>(2)>{});
    else if (W == tuner_w<N>(3)) f(nn, std::integral_constant<int, tuner_w<N>(3)>{});
    else return false;
    return true;
  };
  switch (n) {
    case 16: return by_w(std::integral_constant<int, 16>{});
    case 32: return by_w(std::integral_constant<int, 32>{});
    case 48: return by_w(std::integral_constant<int, 48>{});
    case 64: return by_w(std::integral_constant<int, 64>{});
    case 80: return by_w(std::integral_constant<int, 80>{});
    case 96: return by_w(std::integral_constant<int, 96>{});
    case 128: return by_w(std::integral_constant<int, 128>{});
    default: return false;
  }
}

int choose_tile_lines(int n) {
  int W = 1024 / n;
  if (W < 1) W = 1;
  if (W > 64) W = 64;
  return W;
}

template <typename T>
size_t axis_lds(int n, int W) {
  return (2 * static_cast<size_t>(W) * (n + 1) + n) * sizeof(cplx<T>);
}

template <typename T>
void launch_plane(cplx<T>* data, const FftPlan1D& pz, const FftPlan1D& py, long planes, int threads, int sign,
                  hipStream_t s) {
  const int n = pz.st.n;
  const size_t lds = plane_lds<T>(py.st.n, n);
  const dim3 grid(static_cast<unsigned>(planes));
  if (py.st.n == n && ct_length(n) && n <= 96) {
    auto go = [&](auto nn) {
      constexpr int N = decltype(nn)::value;
      if (threads == 256) {
        hipLaunchKernelGGL((fft_plane_ct_kernel<T, N, 256>), grid, dim3(256), lds, s, data, pz.twiddles<T>(),
                           py.twiddles<T>(), sign);
      } else {
        hipLaunchKernelGGL((fft_plane_ct_kernel<T, N, 512>), grid, dim3(512), lds, s, data, pz.twiddles<T>(),
                           py.twiddles<T>(), sign);
      }
    };
    switch (n) {
      case 16: go(std::integral_constant<int, 16>{}); break;
      case 32: go(std::integral_constant<int, 32>{}); break;
      case 48: go(std::integral_constant<int, 48>{}); break;
      case 64: go(std::integral_constant<int, 64>{}); break;
      case 80: go(std::integral_constant<int, 80>{}); break;
      default: go(std::integral_constant<int, 96>{}); break;
    }
  } else {
    hipLaunchKernelGGL((fft_plane_kernel<T>), grid, dim3(threads), lds, s, data, pz.st, py.st, pz.twiddles<T>(),
                       py.twiddles<T>(), sign);
  }
  MDFT_HIP_CHECK(hipGetLastError());
}

// W = 0: the default tile (1024 / n lines)
template <typename T>
void launch_axis(cplx<T>* data, const FftPlan1D& p, int sign, long nlines_or_outer, long S, bool strided,
                 hipStream_t stream, int W = 0) {
  if (p.st.n == 1 || nlines_or_outer == 0) return;
  if (W <= 0) W = choose_tile_lines(p.st.n);
  if (strided && S < W) W = static_cast<int>(S);
  const size_t lds = axis_lds<T>(p.st.n, W);
  throw_if(lds > kLdsMax, "FFT tile does not fit in LDS");
  const cplx<T>* tw = p.twiddles<T>();
  const bool full = strided ? S % W == 0 : nlines_or_outer % W == 0;
  if (full && ct_length(p.st.n)) {
    const long nblk = strided ? nlines_or_outer * (S / W) : nlines_or_outer / W;
    throw_if(nblk > 0x7fffffffL, "too many FFT tiles");
    const bool done = dispatch_axis_ct(p.st.n, W, [&](auto nn, auto ww) {
      constexpr int N = decltype(nn)::value, WW = decltype(ww)::value;
      if (strided) {
        hipLaunchKernelGGL((fft_axis_ct_kernel<T, true, N, WW>), dim3(static_cast<unsigned>(nblk)), dim3(kFftThreads),
                           lds, stream, data, tw, sign, S);
      } else {
        hipLaunchKernelGGL((fft_axis_ct_kernel<T, false, N, WW>), dim3(static_cast<unsigned>(nblk)),
                           dim3(kFftThreads), lds, stream, data, tw, sign, S);
      }
    });
    if (done) {
      MDFT_HIP_CHECK(hipGetLastError());
      return;
    }
  }
  if (!strided) {
    const long nblk = (nlines_or_outer + W - 1) / W;
    throw_if(nblk > 0x7fffffffL, "too many FFT tiles");
    hipLaunchKernelGGL((fft_axis_kernel<T, false>), dim3(static_cast<unsigned>(nblk)), dim3(kFftThreads), lds,
                       stream, data, p.st, tw, sign, nlines_or_outer, S, W);
  } else {
    const long nblk = nlines_or_outer * ((S + W - 1) / W);
    throw_if(nblk > 0x7fffffffL, "too many FFT tiles");
    hipLaunchKernelGGL((fft_axis_kernel<T, true>), dim3(static_cast<unsigned>(nblk)), dim3(kFftThreads), lds,
                       stream, data, p.st, tw, sign, nlines_or_outer, S, W);
  }
  MDFT_HIP_CHECK(hipGetLastError());
}

}  // namespace

FftPlan1D::FftPlan1D(int n) {
  throw_if(n <= 0, "FFT length must be positive");
  throw_if(n > 4096, "FFT length above 4096 is not supported");
  st.n = n;
  st.nstages = 0;
  int r = n;
  for (int p : {4, 2, 3, 5, 7, 11, 13}) {
    while (r % p == 0) {
      throw_if(st.nstages >= kFftMaxStages, "too many FFT stages");
      st.radix[st.nstages++] = p;
      r /= p;
    }
  }
  throw_if(r != 1, "FFT length must factor into primes <= 13");
  std::vector<double2> h64(n);
  std::vector<float2> h32(n);
  for (int t = 0; t < n; ++t) {
    const double a = 2.0 * M_PI * static_cast<double>(t) / static_cast<double>(n);
    h64[t].x = std::cos(a);
    h64[t].y = -std::sin(a);
    h32[t].x = static_cast<float>(h64[t].x);
    h32[t].y = static_cast<float>(h64[t].y);
  }
  MDFT_HIP_CHECK(hipMalloc(&tw64_, n * sizeof(double2)));
  MDFT_HIP_CHECK(hipMalloc(&tw32_, n * sizeof(float2)));
  MDFT_HIP_CHECK(hipMemcpy(tw64_, h64.data(), n * sizeof(double2), hipMemcpyHostToDevice));
  MDFT_HIP_CHECK(hipMemcpy(tw32_, h32.data(), n * sizeof(float2), hipMemcpyHostToDevice));
}

FftPlan1D::~FftPlan1D() {
  if (tw64_) (void)hipFree(tw64_);
  if (tw32_) (void)hipFree(tw32_);
}

template <>
const cplx<double>* FftPlan1D::twiddles<double>() const {
  return tw64_;
}
template <>
const cplx<float>* FftPlan1D::twiddles<float>() const {
  return tw32_;
}

Fft3dPlan::Fft3dPlan(int nx_, int ny_, int nz_) : nx(nx_), ny(ny_), nz(nz_), px(nx_), py(ny_), pz(nz_) {}

bool Fft3dPlan::plane_ok(size_t elem, int threads) const {
  (void)threads;
  return ny > 1 && nz > 1 && small_radices(py.st) && small_radices(pz.st) &&
         (2 * static_cast<size_t>(ny) * (nz + 1) + ny + nz) * elem <= kLdsMax;
}

template <typename T>
void Fft3dPlan::execute_config(const FftConfig& c, cplx<T>* data, long batch, int sign, hipStream_t stream) const {
  throw_if(sign != -1 && sign != 1, "FFT sign must be -1 or +1");
  const long nxl = nx, nyl = ny, nzl = nz;
  const long planes = batch * nxl;
  if (c.plane) {
    throw_if(!plane_ok(sizeof(cplx<T>), c.plane_threads) || planes > 0x7fffffffL ||
                 (c.plane_threads != 256 && c.plane_threads != 512),
             "FFT: the fused plane pass does not apply to this shape");
    // z and y in one pass over each contiguous (y, z) plane
    launch_plane<T>(data, pz, py, planes, c.plane_threads, sign, stream);
  } else {
    launch_axis<T>(data, pz, sign, batch * nxl * nyl, 1, false, stream, c.wz);  // z: contiguous lines
    launch_axis<T>(data, py, sign, batch * nxl, nzl, true, stream, c.wy);       // y: stride nz
  }
  launch_axis<T>(data, px, sign, batch, nyl * nzl, true, stream, c.wx);  // x: stride ny*nz
}

template <typename T>
FftConfig Fft3dPlan::default_config() const {
  FftConfig c;
  c.plane = plane_ok(sizeof(cplx<T>), 512) && plane_lds<T>(ny, nz) <= kPlaneLdsMax;
  c.plane_threads = 512;
  return c;
}

template <typename T>
void Fft3dPlan::execute(cplx<T>* data, long batch, int sign, hipStream_t stream) const {
  const int k = std::is_same<T, double>::value ? 1 : 0;
  execute_config<T>(tuned_[k] ? cfg_[k] : default_config<T>(), data, batch, sign, stream);
}

// Times the candidate configurations on a scratch array of `batch`
// transforms and keeps the fastest: the tile width of each per-axis pass
// (256, 512, 1024, 2048 values per tile) and, where a (y, z) plane fits in
// LDS, the fused z/y pass with 256 or 512 threads.  Every candidate runs the
// same Stockham stages on the same lines, so the choice changes timing only.
template <typename T>
FftConfig Fft3dPlan::autotune(long batch, hipStream_t s) const {
  const int k = std::is_same<T, double>::value ? 1 : 0;
  if (tuned_[k]) return cfg_[k];
  const long N = static_cast<long>(nx) * ny * nz;
  batch = std::max<long>(1, batch);
  cplx<T>* buf = nullptr;
  MDFT_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&buf), batch * N * sizeof(cplx<T>)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  FftConfig best = default_config<T>();
  try {
    MDFT_HIP_CHECK(hipMemsetAsync(buf, 0, batch * N * sizeof(cplx<T>), s));
    MDFT_HIP_CHECK(hipEventCreate(&e0));
    MDFT_HIP_CHECK(hipEventCreate(&e1));
    auto time_ms = [&](auto&& run) {
      run();
      run();
      float best_ms = 1e30f;
      for (int r = 0; r < 3; ++r) {
        MDFT_HIP_CHECK(hipEventRecord(e0, s));
        run();
        MDFT_HIP_CHECK(hipEventRecord(e1, s));
        MDFT_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        MDFT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        best_ms = std::min(best_ms, ms);
      }
      return best_ms;
    };
    const long nxl = nx, nyl = ny, nzl = nz;
    // per-axis tile widths
    auto best_w = [&](const FftPlan1D& p, long lines, long S, bool strided) {
      int bw = 0;
      float bt = 1e30f;
      for (int vals : {256, 512, 1024, 2048}) {
        int W = std::max(1, std::min(64, vals / std::max(1, p.st.n)));
        if (strided && S < W) W = static_cast<int>(S);
        if (axis_lds<T>(p.st.n, W) > kLdsMax) continue;
        const float t = time_ms([&] { launch_axis<T>(buf, p, -1, lines, S, strided, s, W); });
        if (t < bt) {
          bt = t;
          bw = W;
        }
      }
      return std::make_pair(bw, bt);
    };
    const auto z = best_w(pz, batch * nxl * nyl, 1, false);
    const auto y = best_w(py, batch * nxl, nzl, true);
    const auto x = best_w(px, batch, nyl * nzl, true);
    FftConfig axes;
    axes.plane = false;
    axes.wz = z.first;
    axes.wy = y.first;
    axes.wx = x.first;
    best = axes;
    float best_t = z.second + y.second + x.second;
    if (plane_ok(sizeof(cplx<T>), 512) && batch * nxl <= 0x7fffffffL) {
      for (int th : {256, 512}) {
        FftConfig c = axes;
        c.plane = true;
        c.plane_threads = th;
        const float t = time_ms([&] {
          launch_plane<T>(buf, pz, py, batch * nxl, th, -1, s);
        }) + x.second;
        if (t < best_t) {
          best_t = t;
          best = c;
        }
      }
    }
  } catch (...) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  MDFT_HIP_CHECK(hipStreamSynchronize(s));
  MDFT_HIP_CHECK(hipFree(buf));
  cfg_[k] = best;
  tuned_[k] = true;
  return best;
}

template void Fft3dPlan::execute<double>(cplx<double>*, long, int, hipStream_t) const;
template void Fft3dPlan::execute<float>(cplx<float>*, long, int, hipStream_t) const;
template void Fft3dPlan::execute_config<double>(const FftConfig&, cplx<double>*, long, int, hipStream_t) const;
template void Fft3dPlan::execute_config<float>(const FftConfig&, cplx<float>*, long, int, hipStream_t) const;
template FftConfig Fft3dPlan::autotune<double>(long, hipStream_t) const;
template FftConfig Fft3dPlan::autotune<float>(long, hipStream_t) const;
template FftConfig Fft3dPlan::default_config<double>() const;
template FftConfig Fft3dPlan::default_config<float>() const;

}  // namespace mdft
