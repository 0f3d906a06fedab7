// Orientation <-> projection transforms in the factorised form, for the FP64
// and FP32 precision forms, plus delta_rho / delta_f and their deterministic
// FP64 reductions (gfx950).
//
// angl2proj (common/MDFT_OrientationProjectionTransform.hpp:267-380) is an
// unnormalised rfft2 over (phi, psi), a conjugation / (nphi npsi), the mup ->
// kphi unscramble and a theta quadrature.  Because the field is real, that
// equals
//   p[ip](r) = sum_theta W[theta][ip] Y[theta][mup][mu2],
//   Y = sum_j e^{+2 pi i mup j/nphi} sum_k e^{+2 pi i mu2 k/npsi} o[theta][j][k],
// W = f_m w_theta d^m_{mup,mu}(theta) / (nphi npsi).  One block handles RT
// consecutive r: the RT x no input tile is staged in LDS (coalesced), then the
// psi sums X (one thread per (r, theta, j, mu2)), the phi sums Y (one thread
// per (r, theta, mup, mu2)) and the theta sums (one thread per (ip, r), r
// fastest so every ip row is written as a run of RT values).
//
// proj2angl (:388-484) runs the same three steps backwards: S = conj(sum_m p d
// f_m), Z = sum_mup S e^{+2 pi i mup j/nphi}, and the c2r over psi with only
// Re of the DC bin and twice Re of the others (npsi = 2 mmax/mrso + 1 is odd,
// so there is no Nyquist bin).  Optionally get_delta_f
// (common/MDFT_Solvent.hpp:613-660) is fused into the last step.
#include "mdft_proj.hpp"

#include <algorithm>
#include <cmath>

namespace mdft {

namespace {

constexpr int kThreads = 256;
constexpr int kRT = 8;  // r values per block

struct TileDims {
  int ntheta, nphi, npsi, nmu2, no, np, mmax;
};

template <typename TC>
struct TabPtrs {
  const cplx<TC>* ephi;
  const cplx<TC>* epsi;
  const TC* wts;
  const int* ipinfo;
  const int* yidx;
};

__device__ __forceinline__ int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

template <typename TC>
__device__ __forceinline__ void block_sum_store(double v, double* out) {
  __shared__ double red[kThreads];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

template <typename TI, typename TP, typename TC>
__global__ __launch_bounds__(kThreads) void angl2proj_kernel(const TI* __restrict__ o, cplx<TP>* __restrict__ p,
                                                             long N, TileDims d, TabPtrs<TC> tb, int fuse,
                                                             double rho0) {
  extern __shared__ unsigned char lds_raw[];
  const int cells = kRT * d.ntheta * d.nphi * d.nmu2;  // X / Y entries per tile
  const int regA = max(cells * 2, kRT * d.no);          // in TC units
  TC* A = reinterpret_cast<TC*>(lds_raw);               // o tile, then Y
  cplx<TC>* X = reinterpret_cast<cplx<TC>*>(A + regA);  // psi sums
  cplx<TC>* Y = reinterpret_cast<cplx<TC>*>(A);
  cplx<TC>* ephi = X + cells;
  cplx<TC>* epsi = ephi + d.nphi;
  TC* wts = reinterpret_cast<TC*>(epsi + d.nmu2 * d.npsi);

  const int tid = threadIdx.x;
  const long r0 = static_cast<long>(blockIdx.x) * kRT;
  const int rv = static_cast<int>(min(static_cast<long>(kRT), N - r0));

  for (int i = tid; i < d.nphi; i += kThreads) ephi[i] = tb.ephi[i];
  for (int i = tid; i < d.nmu2 * d.npsi; i += kThreads) epsi[i] = tb.epsi[i];
  for (int i = tid; i < d.ntheta * d.np; i += kThreads) wts[i] = tb.wts[i];
  const TI* src = o + r0 * d.no;
  for (int i = tid; i < rv * d.no; i += kThreads) {
    if (fuse) {
      const double x = static_cast<double>(src[i]);
      A[i] = static_cast<TC>(rho0 * (x * x - 1.0));
    } else {
      A[i] = static_cast<TC>(src[i]);
    }
  }
  __syncthreads();

  // X[r][theta][j][mu2] = sum_k o[r][theta][j][k] e^{+2 pi i mu2 k/npsi}
  const int nx_items = rv * d.ntheta * d.nphi * d.nmu2;
  for (int it = tid; it < nx_items; it += kThreads) {
    const int mu2 = it % d.nmu2;
    const int rtj = it / d.nmu2;  // (r*ntheta + theta)*nphi + j
    const TC* row = A + (rtj / (d.ntheta * d.nphi)) * d.no + (rtj % (d.ntheta * d.nphi)) * d.npsi;
    const cplx<TC>* e = epsi + mu2 * d.npsi;
    TC re = 0, im = 0;
    for (int k = 0; k < d.npsi; ++k) {
      re = fma(row[k], e[k].x, re);
      im = fma(row[k], e[k].y, im);
    }
    X[it] = mkc<TC>(re, im);
  }
  __syncthreads();

  // Y[r][theta][mup][mu2] = sum_j X[r][theta][j][mu2] e^{+2 pi i mup j/nphi}
  for (int it = tid; it < nx_items; it += kThreads) {
    const int mu2 = it % d.nmu2;
    const int rtm = it / d.nmu2;
    const int imup = rtm % d.nphi;
    const int rt = rtm / d.nphi;
    const int mup = imup - d.mmax;
    const cplx<TC>* xs = X + rt * d.nphi * d.nmu2 + mu2;
    cplx<TC> acc = mkc<TC>(0, 0);
    for (int j = 0; j < d.nphi; ++j) acc = cfma<TC>(acc, xs[j * d.nmu2], ephi[pmod(mup * j, d.nphi)]);
    Y[it] = acc;
  }
  __syncthreads();

  // p[ip][r] = sum_theta W[theta][ip] Y[r][theta][mup(ip)][mu2(ip)]
  for (int it = tid; it < d.np * rv; it += kThreads) {
    const int ip = it / rv, r = it - ip * rv;
    const int info = tb.ipinfo[ip];
    const int imup = info & 0xff, mu2 = (info >> 8) & 0xff;
    const cplx<TC>* ys = Y + (r * d.ntheta * d.nphi + imup) * d.nmu2 + mu2;
    TC re = 0, im = 0;
    for (int th = 0; th < d.ntheta; ++th) {
      const TC wv = wts[th * d.np + ip];
      const cplx<TC> y = ys[th * d.nphi * d.nmu2];
      re = fma(wv, y.x, re);
      im = fma(wv, y.y, im);
    }
    p[static_cast<long>(ip) * N + r0 + r] = mkc<TP>(static_cast<TP>(re), static_cast<TP>(im));
  }
}

// angl2proj_kernel for an angular grid fixed at compile time (the grids of
// dispatch_angular): the same three steps with every extent a constant, so
// that the index arithmetic folds and the short loops unroll.
template <int MMAX, int MRSO, int RT>
struct A2pSizes {
  static constexpr int kNth = MMAX + 1, kNphi = 2 * MMAX + 1, kNmu2 = MMAX / MRSO + 1;
  static constexpr int kNpsi = 2 * (MMAX / MRSO) + 1, kNo = kNth * kNphi * kNpsi;
  static constexpr int kNp() {
    int n = 0;
    for (int m = 0; m <= MMAX; ++m) n += (2 * m + 1) * (m / MRSO + 1);
    return n;
  }
  static constexpr int kCells = RT * kNth * kNphi * kNmu2;
  static constexpr int kRegA = (2 * kCells > RT * kNo) ? 2 * kCells : RT * kNo;
};

// rows per block of the compile-time angl2proj: 4 in FP64, whose tile holds
// twice the bytes (more blocks per CU), 8 in FP32
template <typename TC>
constexpr int a2p_ct_rows() {
  return sizeof(TC) == 8 ? 4 : 8;
}

template <int MMAX, int MRSO, typename TI, typename TP, typename TC, bool CP>
__global__ __launch_bounds__(kThreads) void angl2proj_ct_kernel(const TI* __restrict__ o, cplx<TP>* __restrict__ p,
                                                                long N, TabPtrs<TC> tb, int fuse, double rho0) {
  constexpr int kRT = a2p_ct_rows<TC>();
  using Z = A2pSizes<MMAX, MRSO, kRT>;
  constexpr int NTH = Z::kNth, NPHI = Z::kNphi, NMU2 = Z::kNmu2, NPSI = Z::kNpsi, NO = Z::kNo, NP = Z::kNp();
  constexpr int CELLS = Z::kCells;
  __shared__ TC A[Z::kRegA];  // o tile, then Y
  __shared__ cplx<TC> X[CELLS];
  __shared__ cplx<TC> ephi[NPHI];
  __shared__ cplx<TC> epsi[NMU2 * NPSI];
  __shared__ TC wts[NTH * NP];
  cplx<TC>* Y = reinterpret_cast<cplx<TC>*>(A);

  const int tid = threadIdx.x;
  const long r0 = static_cast<long>(blockIdx.x) * kRT;
  const int rv = static_cast<int>(min(static_cast<long>(kRT), N - r0));

  if (tid < NPHI) ephi[tid] = tb.ephi[tid];
  if (tid < NMU2 * NPSI) epsi[tid] = tb.epsi[tid];
  for (int i = tid; i < NTH * NP; i += kThreads) wts[i] = tb.wts[i];
  const TI* src = o + r0 * NO;
  if (rv == kRT) {  // a full tile: every load of the thread in flight at once
    constexpr int kPer = (kRT * NO + kThreads - 1) / kThreads;
    double v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = tid + u * kThreads;
      v[u] = static_cast<double>(src[i < kRT * NO ? i : 0]);
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int i = tid + u * kThreads;
      if (i < kRT * NO) A[i] = static_cast<TC>(fuse ? rho0 * (v[u] * v[u] - 1.0) : v[u]);
    }
  } else {
    for (int i = tid; i < rv * NO; i += kThreads) {
      const double x = static_cast<double>(src[i]);
      A[i] = static_cast<TC>(fuse ? rho0 * (x * x - 1.0) : x);
    }
  }
  __syncthreads();

  // X[r][theta][j][mu2] = sum_k o[r][theta][j][k] e^{+2 pi i mu2 k/npsi}
  const int nx_items = rv * NTH * NPHI * NMU2;
  for (int it = tid; it < nx_items; it += kThreads) {
    const int mu2 = it % NMU2;
    const int rtj = it / NMU2;  // (r*ntheta + theta)*nphi + j
    const TC* row = A + (rtj / (NTH * NPHI)) * NO + (rtj % (NTH * NPHI)) * NPSI;
    const cplx<TC>* e = epsi + mu2 * NPSI;
    TC re = 0, im = 0;
#pragma unroll
    for (int k = 0; k < NPSI; ++k) {
      re = fma(row[k], e[k].x, re);
      im = fma(row[k], e[k].y, im);
    }
    X[it] = mkc<TC>(re, im);
  }
  __syncthreads();

  // Y[r][theta][mup][mu2] = sum_j X[r][theta][j][mu2] e^{+2 pi i mup j/nphi}
  for (int it = tid; it < nx_items; it += kThreads) {
    const int mu2 = it % NMU2;
    const int rtm = it / NMU2;
    const int imup = rtm % NPHI;
    const int rt = rtm / NPHI;
    const int mup = imup - MMAX;
    const cplx<TC>* xs = X + rt * NPHI * NMU2 + mu2;
    cplx<TC> acc = mkc<TC>(0, 0);
    int t = 0;  // mup j mod nphi
    const int step = mup < 0 ? mup + NPHI : mup;
#pragma unroll
    for (int j = 0; j < NPHI; ++j) {
      acc = cfma<TC>(acc, xs[j * NMU2], ephi[t]);
      t += step;
      if (t >= NPHI) t -= NPHI;
    }
    Y[it] = acc;
  }
  __syncthreads();

  // p[ip][r] = sum_theta W[theta][ip] Y[r][theta][mup(ip)][mu2(ip)] (the
  // compact order: the stored projections, at their compact channels)
  for (int it = tid; it < NP * rv; it += kThreads) {
    const int ip = it / rv, r = it - ip * rv;
    const int info = tb.ipinfo[ip];
    const int imup = info & 0xff, mu2 = (info >> 8) & 0xff;
    if (CP && implied_projection(imup - MMAX, mu2)) continue;
    const cplx<TC>* ys = Y + (r * NTH * NPHI + imup) * NMU2 + mu2;
    TC re = 0, im = 0;
#pragma unroll
    for (int th = 0; th < NTH; ++th) {
      const TC wv = wts[th * NP + ip];
      const cplx<TC> y = ys[th * NPHI * NMU2];
      re = fma(wv, y.x, re);
      im = fma(wv, y.y, im);
    }
    const int ch = CP ? compact_of(ip, (info >> 16) & 0xff, imup - MMAX, mu2) : ip;
    p[static_cast<long>(ch) * N + r0 + r] = mkc<TP>(static_cast<TP>(re), static_cast<TP>(im));
  }
}

template <typename TP, typename TO, typename TC>
__global__ __launch_bounds__(kThreads) void proj2angl_kernel(const cplx<TP>* __restrict__ p, TO* __restrict__ o,
                                                             const TO* __restrict__ xi, TO* __restrict__ df,
                                                             const double* __restrict__ wq, long N, TileDims d,
                                                             TabPtrs<TC> tb, double rho0,
                                                             double* __restrict__ partials) {
  extern __shared__ unsigned char lds_raw[];
  const int cells = kRT * d.ntheta * d.nphi * d.nmu2;
  const int regA = max(cells, kRT * d.np);  // complex entries
  cplx<TC>* A = reinterpret_cast<cplx<TC>*>(lds_raw);  // p tile, then Z
  cplx<TC>* S = A + regA;
  cplx<TC>* ephi = S + cells;
  cplx<TC>* epsi = ephi + d.nphi;
  TC* wts = reinterpret_cast<TC*>(epsi + d.nmu2 * d.npsi);

  const int tid = threadIdx.x;
  const long r0 = static_cast<long>(blockIdx.x) * kRT;
  const int rv = static_cast<int>(min(static_cast<long>(kRT), N - r0));

  for (int i = tid; i < d.nphi; i += kThreads) ephi[i] = tb.ephi[i];
  for (int i = tid; i < d.nmu2 * d.npsi; i += kThreads) epsi[i] = tb.epsi[i];
  for (int i = tid; i < d.ntheta * d.np; i += kThreads) wts[i] = tb.wts[i];
  for (int it = tid; it < d.np * rv; it += kThreads) {
    const int ip = it / rv, r = it - ip * rv;
    const cplx<TP> v = p[static_cast<long>(ip) * N + r0 + r];
    A[ip * kRT + r] = mkc<TC>(static_cast<TC>(v.x), static_cast<TC>(v.y));
  }
  __syncthreads();

  // S[r][theta][mup][mu2] = conj(sum_{m >= max(|mup|, mrso mu2)} p[P(m,mup,mu2)] d f_m)
  const int n_items = rv * d.ntheta * d.nphi * d.nmu2;
  for (int it = tid; it < n_items; it += kThreads) {
    const int mu2 = it % d.nmu2;
    const int rtm = it / d.nmu2;
    const int imup = rtm % d.nphi;
    const int rt = rtm / d.nphi;
    const int th = rt % d.ntheta, r = rt / d.ntheta;
    const int* ys = tb.yidx + (imup * d.nmu2 + mu2) * (d.mmax + 1);
    TC re = 0, im = 0;
    for (int m = 0; m <= d.mmax; ++m) {
      const int ip = ys[m];
      if (ip < 0) continue;
      const TC wv = wts[th * d.np + ip];
      const cplx<TC> v = A[ip * kRT + r];
      re = fma(v.x, wv, re);
      im = fma(v.y, wv, im);
    }
    S[it] = mkc<TC>(re, -im);
  }
  __syncthreads();

  // Z[r][theta][j][mu2] = sum_mup S[r][theta][mup][mu2] e^{+2 pi i mup j/nphi}
  cplx<TC>* Z = A;
  for (int it = tid; it < n_items; it += kThreads) {
    const int mu2 = it % d.nmu2;
    const int rtj = it / d.nmu2;
    const int j = rtj % d.nphi;
    const int rt = rtj / d.nphi;
    const cplx<TC>* ss = S + rt * d.nphi * d.nmu2 + mu2;
    cplx<TC> acc = mkc<TC>(0, 0);
    for (int imup = 0; imup < d.nphi; ++imup)
      acc = cfma<TC>(acc, ss[imup * d.nmu2], ephi[pmod((imup - d.mmax) * j, d.nphi)]);
    Z[it] = acc;
  }
  __syncthreads();

  // o[r][theta][j][k] = Re Z[..][0] + 2 sum_{mu2 >= 1} Re(Z[..][mu2] e^{+2 pi i mu2 k/npsi})
  double acc_ff = 0.0;
  for (int it = tid; it < rv * d.no; it += kThreads) {
    const int io = it % d.no, r = it / d.no;
    const int k = io % d.npsi, tj = io / d.npsi;
    const cplx<TC>* zs = Z + (r * d.ntheta * d.nphi + tj) * d.nmu2;
    TC v = zs[0].x;
    for (int mu2 = 1; mu2 < d.nmu2; ++mu2) {
      const cplx<TC> e = epsi[mu2 * d.npsi + k];
      const cplx<TC> z = zs[mu2];
      v += static_cast<TC>(2) * (z.x * e.x - z.y * e.y);
    }
    const long g = (r0 + r) * d.no + io;
    if (xi == nullptr) {
      o[g] = static_cast<TO>(v);
    } else {
      const double x = static_cast<double>(xi[g]);
      const double vx = static_cast<double>(static_cast<TO>(v));
      acc_ff += vx * wq[io] * (x * x - 1.0);
      df[g] = static_cast<TO>(2.0 * rho0 * x * vx);
    }
  }
  if (xi != nullptr) block_sum_store<TC>(acc_ff, partials + blockIdx.x);
}

// proj2angl for angular grids with compile-time sizes: one 64-lane block per
// (64 consecutive r, theta).  Each lane keeps its S[mup][mu2] in registers
// (the m sums in the same ascending-m order as proj2angl_kernel), runs the
// phi sums on the (mup, -mup) pairs, Z = S_0 + sum_{mup > 0} (A cos + i B sin)
// with A = S_mup + S_-mup and B = S_mup - S_-mup, then the psi c2r.  All
// indices and twiddle offsets are constants after unrolling (scalar table
// loads).  The 64 x (nphi npsi) outputs are staged in LDS and written (with
// the fused get_delta_f) in runs of consecutive io, the layout of o/xi/df.
// Blocks are ordered so that the ntheta blocks of one r range run on the same
// XCD (block b runs on XCD b mod 8) and share their projection reads in L2.
template <int MMAX, int MRSO>
struct AngSizes {
  static constexpr int kNth = MMAX + 1, kNphi = 2 * MMAX + 1, kNmu2 = MMAX / MRSO + 1;
  static constexpr int kNpsi = 2 * (MMAX / MRSO) + 1, kNjk = kNphi * kNpsi, kNo = kNth * kNjk;
  static constexpr int np() {
    int n = 0;
    for (int m = 0; m <= MMAX; ++m) n += (2 * m + 1) * (m / MRSO + 1);
    return n;
  }
};

constexpr int kWaveRows = 64;  // r values per block of the compile-time kernel
constexpr int kXcds = 8;

template <int MMAX, int MRSO, typename TP, typename TO, typename TC, bool CP>
__global__ __launch_bounds__(64) void proj2angl_wave_kernel(const cplx<TP>* __restrict__ p, TO* __restrict__ o,
                                                            const TO* __restrict__ xi, TO* __restrict__ df,
                                                            const double* __restrict__ wq, long N, long nrb,
                                                            TabPtrs<TC> tb, double rho0,
                                                            double* __restrict__ partials,
                                                            double* __restrict__ colpart) {
  using A = AngSizes<MMAX, MRSO>;
  constexpr int NP = A::np();
  // FP64 outputs are staged half the phi values at a time (the full stage of
  // 64 x nphi npsi doubles would hold a CU to five blocks)
  constexpr int kChunks = sizeof(TC) == 8 ? 2 : 1;
  constexpr int kJc = (A::kNphi + kChunks - 1) / kChunks;
  __shared__ TC stage[kWaveRows * kJc * A::kNpsi];
  const int lane = threadIdx.x;
  const int q = blockIdx.x / kXcds;
  const int th = q % A::kNth;
  const long rb = static_cast<long>(q / A::kNth) * kXcds + blockIdx.x % kXcds;
  if (rb >= nrb) return;  // padding block of the XCD ordering
  const long r0 = rb * kWaveRows;
  const long r = min(r0 + lane, N - 1);
  const TC* w = tb.wts + th * NP;  // d f_m of this theta
  const double wth = wq[th * A::kNjk];  // the angular weight of every (phi, psi) of this theta

  TC sr[A::kNphi][A::kNmu2], si[A::kNphi][A::kNmu2];
#pragma unroll
  for (int a = 0; a < A::kNphi; ++a)
#pragma unroll
    for (int b = 0; b < A::kNmu2; ++b) sr[a][b] = si[a][b] = TC(0);
  {
    constexpr CompactMap<MMAX, MRSO> kCm{};
    int ip = 0;
#pragma unroll
    for (int m = 0; m <= MMAX; ++m)
#pragma unroll
      for (int mup = -m; mup <= m; ++mup)
#pragma unroll
        for (int mu = 0; mu <= m; mu += MRSO, ++ip) {
          // compact order: the stored channel (mup, 0) holds the pair's
          // combination, and the implied sums (-mup, 0) stay zero
          if (CP && implied_projection(mup, mu)) continue;
          const cplx<TP> v = p[static_cast<long>(CP ? kCm.v[ip] : ip) * N + r];
          const TC wt = w[ip];
          sr[mup + MMAX][mu / MRSO] = fma(static_cast<TC>(v.x), wt, sr[mup + MMAX][mu / MRSO]);
          si[mup + MMAX][mu / MRSO] = fma(static_cast<TC>(v.y), wt, si[mup + MMAX][mu / MRSO]);
        }
  }
  // S = conj(sum); sr/si of +-mup become A = S+ + S- and B = S+ - S-
#pragma unroll
  for (int b = 0; b < A::kNmu2; ++b) {
    si[MMAX][b] = -si[MMAX][b];
#pragma unroll
    for (int mup = 1; mup <= MMAX; ++mup) {
      const TC pr = sr[MMAX + mup][b], pi = -si[MMAX + mup][b];
      const TC nr = sr[MMAX - mup][b], ni = -si[MMAX - mup][b];
      sr[MMAX + mup][b] = pr + nr;
      si[MMAX + mup][b] = pi + ni;
      sr[MMAX - mup][b] = pr - nr;
      si[MMAX - mup][b] = pi - ni;
    }
  }
  double acc_ff = 0.0;
  const long base = r0 * A::kNo + th * A::kNjk;
  // the (phi, psi) outputs of chunk c: phi in [c kJc, c kJc + jn), their nphi
  // npsi columns [c0, c0 + W) of the row's run, staged as [64 rows][W]
  const auto chunk = [&](auto cc) {
    constexpr int c = decltype(cc)::value;
    constexpr int j0 = c * kJc, jn = A::kNphi - j0 < kJc ? A::kNphi - j0 : kJc;
    constexpr int W = jn * A::kNpsi, c0 = j0 * A::kNpsi;
    if (c > 0) __syncthreads();  // the previous chunk's epilogue has read the stage
    TC* out = stage + lane * W;
#pragma unroll
    for (int jj = 0; jj < jn; ++jj) {
      const int j = j0 + jj;
      TC zr[A::kNmu2], zi[A::kNmu2];
#pragma unroll
      for (int b = 0; b < A::kNmu2; ++b) {
        zr[b] = sr[MMAX][b];
        zi[b] = si[MMAX][b];
#pragma unroll
        for (int mup = 1; mup <= MMAX; ++mup) {
          const cplx<TC> e = tb.ephi[(mup * j) % A::kNphi];
          zr[b] = fma(sr[MMAX + mup][b], e.x, fma(-si[MMAX - mup][b], e.y, zr[b]));
          zi[b] = fma(si[MMAX + mup][b], e.x, fma(sr[MMAX - mup][b], e.y, zi[b]));
        }
      }
#pragma unroll
      for (int k = 0; k < A::kNpsi; ++k) {
        TC v = zr[0];
#pragma unroll
        for (int b = 1; b < A::kNmu2; ++b) {
          const cplx<TC> e = tb.epsi[b * A::kNpsi + k];
          v += static_cast<TC>(2) * (zr[b] * e.x - zi[b] * e.y);
        }
        out[jj * A::kNpsi + k] = v;
      }
    }
    __syncthreads();

    constexpr int kE = W;  // elements per lane (64 rows x W columns)
    if (colpart != nullptr && xi != nullptr) {
      // get_delta_f with the column sums: lane jk owns orientation th * nphi
      // npsi + c0 + jk and walks the block's rows, so that its sums over r of
      // xi^2 - 1 and of vexc build up in registers (each row's W values are
      // one contiguous run of xi and df; for the mixed forms this walk also
      // measured faster than the row-major one below, 0.358 against 0.371 ms
      // at 64^3, and in FP64 the same, 1.11 against 1.10 ms)
      constexpr int kRB = 32;  // rows per batch of loads in flight
      const int nrow = static_cast<int>(min(static_cast<long>(kWaveRows), N - r0));
      const bool on = lane < W;
      const int jk = on ? lane : 0;
      double cs = 0.0, cv = 0.0;
      for (int rl0 = 0; rl0 < kWaveRows; rl0 += kRB) {
        double xv[kRB];
#pragma unroll
        for (int b = 0; b < kRB; ++b) {
          const int rl = min(rl0 + b, nrow - 1);
          xv[b] = static_cast<double>(xi[base + static_cast<long>(rl) * A::kNo + c0 + jk]);
        }
#pragma unroll
        for (int b = 0; b < kRB; ++b) {
          const int rl = rl0 + b;
          if (on && rl < nrow) {
            const double x = xv[b];
            const double vx = static_cast<double>(static_cast<TO>(stage[rl * W + jk]));
            const double x2m1 = x * x - 1.0;
            acc_ff += vx * wth * x2m1;
            df[base + static_cast<long>(rl) * A::kNo + c0 + jk] = static_cast<TO>(2.0 * rho0 * x * vx);
            cs += x2m1;
            cv += vx;
          }
        }
      }
      if (on) {
        // colpart[rb][0][io] = sum_r (xi^2 - 1), colpart[rb][1][io] = sum_r vexc
        double* cp = colpart + rb * (2 * A::kNo) + th * A::kNjk + c0 + jk;
        cp[0] = cs;
        cp[A::kNo] = cv;
      }
    } else if (xi != nullptr && r0 + kWaveRows <= N) {
      // full block: batches of kB elements per lane with all their xi loads
      // in flight before the arithmetic.  The mixed forms (FP32 arithmetic,
      // FP64 fields) take all kE at once: at mmax 5 that waits out one HBM
      // latency instead of five and measured 3% faster, at the same three
      // waves per SIMD.  The FP64 and FP32 forms keep batches of 11 (one
      // batch would nearly double their registers at mmax 5).
      constexpr bool kMixed = std::is_same<TC, float>::value && std::is_same<TO, double>::value;
      constexpr int kB = kMixed && kE <= 64 ? kE : 11;
      for (int i0 = 0; i0 < kE; i0 += kB) {
        double xv[kB];
#pragma unroll
        for (int b = 0; b < kB; ++b) {
          const int e = lane + 64 * min(i0 + b, kE - 1);
          const int rl = e / W, jk = e - rl * W;
          xv[b] = static_cast<double>(xi[base + static_cast<long>(rl) * A::kNo + c0 + jk]);
        }
#pragma unroll
        for (int b = 0; b < kB; ++b) {
          if (i0 + b < kE) {
            const int e = lane + 64 * (i0 + b);
            const int rl = e / W, jk = e - rl * W;
            const long g = base + static_cast<long>(rl) * A::kNo + c0 + jk;
            const double x = xv[b];
            const double vx = static_cast<double>(static_cast<TO>(stage[e]));
            acc_ff += vx * wth * (x * x - 1.0);
            df[g] = static_cast<TO>(2.0 * rho0 * x * vx);
          }
        }
      }
    } else {
#pragma unroll 5
      for (int e = lane; e < kWaveRows * W; e += 64) {
        const int rl = e / W, jk = e - rl * W;
        if (r0 + rl < N) {
          const long g = base + static_cast<long>(rl) * A::kNo + c0 + jk;
          const TC v = stage[e];
          if (xi == nullptr) {
            o[g] = static_cast<TO>(v);
          } else {
            const double x = static_cast<double>(xi[g]);
            const double vx = static_cast<double>(static_cast<TO>(v));
            acc_ff += vx * wth * (x * x - 1.0);
            df[g] = static_cast<TO>(2.0 * rho0 * x * vx);
          }
        }
      }
    }
  };
  chunk(std::integral_constant<int, 0>{});
  if constexpr (kChunks > 1) chunk(std::integral_constant<int, 1>{});
  if (xi != nullptr) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc_ff += __shfl_xor(acc_ff, off);
    // indexed by the data (row block, theta), not by the launch order, so
    // that the partials of consecutive fields of a batch are contiguous
    if (lane == 0) partials[rb * A::kNth + th] = acc_ff;
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void delta_rho_kernel(const T* __restrict__ xi, T* __restrict__ out, size_t n,
                                                             T rho0) {
  for (size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    const T x = xi[i];
    out[i] = rho0 * (x * x - static_cast<T>(1.0));
  }
}

constexpr int kDfRows = 8;

template <typename T>
__global__ __launch_bounds__(kThreads) void delta_f_kernel(const T* __restrict__ xi, const T* __restrict__ vexc,
                                                           const double* __restrict__ w, T* __restrict__ df,
                                                           long nxyz, int no, double rho0,
                                                           double* __restrict__ partials) {
  const long r0 = static_cast<long>(blockIdx.x) * kDfRows;
  const long rv = min(static_cast<long>(kDfRows), nxyz - r0);
  const long base = r0 * no;
  double acc = 0.0;
  for (long i = threadIdx.x; i < rv * no; i += kThreads) {
    const int io = static_cast<int>(i % no);
    const double x = static_cast<double>(xi[base + i]);
    const double v = static_cast<double>(vexc[base + i]);
    acc += v * w[io] * (x * x - 1.0);
    df[base + i] = static_cast<T>(2.0 * rho0 * x * v);
  }
  block_sum_store<double>(acc, partials + blockIdx.x);
}

// block j: out[j out_stride] = scale * sum of partials[j count .. (j + 1) count)
// in a fixed order: thread t sums t, t + T, ... in four interleaved chains
// (all four loads of a round in flight together), then the block combines
// the threads' sums as a tree
constexpr int kRedThreads = 1024;
__global__ __launch_bounds__(kRedThreads) void reduce_partials_kernel(const double* __restrict__ partials, int count,
                                                                      double scale, double* __restrict__ out,
                                                                      long out_stride,
                                                                      const double* __restrict__ extra,
                                                                      int nextra) {
  partials += static_cast<long>(blockIdx.x) * count;
  out += blockIdx.x * out_stride;
  constexpr int T = kRedThreads;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int i = threadIdx.x;
  for (; i + 3 * T < count; i += 4 * T) {
    const double v0 = partials[i], v1 = partials[i + T], v2 = partials[i + 2 * T], v3 = partials[i + 3 * T];
    a0 += v0;
    a1 += v1;
    a2 += v2;
    a3 += v3;
  }
  for (; i < count; i += T) a0 += partials[i];
  if (extra)  // the extra terms of this field, after the partials, in the same fixed order
    for (int e = threadIdx.x; e < nextra; e += T) a1 += extra[static_cast<long>(blockIdx.x) * nextra + e];
  __shared__ double red[T];
  red[threadIdx.x] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  for (int s = T / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = scale * red[0];
}

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void fill_uniform_kernel(T* __restrict__ x, size_t n, unsigned long long key,
                                                                unsigned long long offset, double lo, double hi) {
#pragma clang fp contract(off)  // lo + (hi - lo) * u rounded twice, as numpy does
  for (size_t i = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    const unsigned long long z = splitmix64(key + offset + i);
    const double u = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
    x[i] = static_cast<T>(lo + (hi - lo) * u);
  }
}

unsigned long long splitmix64_host(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

TileDims dims_of(const ProjTables& t) {
  return TileDims{t.ntheta, t.nphi, t.npsi, t.nmu2, t.no, t.np, t.mmax};
}

template <typename TC>
TabPtrs<TC> tabs_of(const ProjTables& t, bool forward);
template <>
TabPtrs<double> tabs_of<double>(const ProjTables& t, bool forward) {
  return TabPtrs<double>{t.ephi64, t.epsi64, forward ? t.wfwd64 : t.wbwd64, t.ipinfo, t.yidx};
}
template <>
TabPtrs<float> tabs_of<float>(const ProjTables& t, bool forward) {
  return TabPtrs<float>{t.ephi32, t.epsi32, forward ? t.wfwd32 : t.wbwd32, t.ipinfo, t.yidx};
}

template <typename TC>
size_t a2p_lds(const ProjTables& t) {
  const size_t cells = static_cast<size_t>(kRT) * t.ntheta * t.nphi * t.nmu2;
  const size_t regA = std::max(cells * 2, static_cast<size_t>(kRT) * t.no);
  return regA * sizeof(TC) + (cells + t.nphi + t.nmu2 * t.npsi) * sizeof(cplx<TC>) +
         static_cast<size_t>(t.ntheta) * t.np * sizeof(TC);
}

template <typename TC>
size_t p2a_lds(const ProjTables& t) {
  const size_t cells = static_cast<size_t>(kRT) * t.ntheta * t.nphi * t.nmu2;
  const size_t regA = std::max(cells, static_cast<size_t>(kRT) * t.np);
  return (regA + cells + t.nphi + t.nmu2 * t.npsi) * sizeof(cplx<TC>) +
         static_cast<size_t>(t.ntheta) * t.np * sizeof(TC);
}

template <typename T>
T* upload(const std::vector<T>& h) {
  T* d = nullptr;
  MDFT_HIP_CHECK(hipMalloc(&d, h.size() * sizeof(T)));
  MDFT_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

}  // namespace

int compact_projections(const AngularGrid& g) {
  int c = 0;
  for (int ip = 0; ip < g.np; ++ip) c += implied_projection(g.mup[ip], g.mu[ip]) ? 0 : 1;
  return c;
}

std::vector<int> compact_codes(const AngularGrid& g) {
  throw_if(g.np > 256, "compact channel order: at most 256 projections");
  std::vector<int> idx(g.np, -1), code(g.np, 0);
  int c = 0;
  for (int ip = 0; ip < g.np; ++ip)
    if (!implied_projection(g.mup[ip], g.mu[ip])) idx[ip] = c++;
  for (int ip = 0; ip < g.np; ++ip) {
    if (idx[ip] >= 0) {
      code[ip] = idx[ip];
      continue;
    }
    const int partner = g.P(g.m[ip], -g.mup[ip], 0);
    throw_if(partner < 0 || idx[partner] < 0 || g.mu[partner] != 0, "internal error: compact projection partner");
    code[ip] = idx[partner] | kCompactImplied | ((g.mup[ip] & 1) ? kCompactOdd : 0);
  }
  return code;
}

ProjTables::ProjTables(const AngularGrid& g)
    : mmax(g.mmax), mrso(g.mrso), ntheta(g.ntheta), nphi(g.nphi), npsi(g.npsi), nmu2(g.nmu2), no(g.no), np(g.np) {
  throw_if(npsi % 2 == 0, "npsi must be odd");
  throw_if(nphi > 255 || nmu2 > 255, "angular grid too large for the packed index table");
  std::vector<double2> eph(nphi), eps(static_cast<size_t>(nmu2) * npsi);
  for (int t = 0; t < nphi; ++t) {
    const double a = 2.0 * M_PI * t / nphi;
    eph[t].x = std::cos(a);
    eph[t].y = std::sin(a);
  }
  for (int m2 = 0; m2 < nmu2; ++m2)
    for (int k = 0; k < npsi; ++k) {
      const double a = 2.0 * M_PI * static_cast<double>((m2 * k) % npsi) / npsi;
      eps[m2 * npsi + k].x = std::cos(a);
      eps[m2 * npsi + k].y = std::sin(a);
    }
  std::vector<double> wf(static_cast<size_t>(ntheta) * np), wb(static_cast<size_t>(ntheta) * np);
  const double inv = 1.0 / static_cast<double>(nphi * npsi);
  for (int th = 0; th < ntheta; ++th)
    for (int ip = 0; ip < np; ++ip) {
      const double dv = g.wigner_d[th * np + ip], fm = g.fm[g.m[ip]];
      wf[th * np + ip] = fm * g.wtheta[th] * dv * inv;
      wb[th * np + ip] = dv * fm;
    }
  std::vector<int> info(np), yi(static_cast<size_t>(nphi) * nmu2 * (mmax + 1), -1);
  for (int ip = 0; ip < np; ++ip) {
    info[ip] = (g.mup[ip] + mmax) | (g.mu2[ip] << 8) | (g.m[ip] << 16);
    yi[((g.mup[ip] + mmax) * nmu2 + g.mu2[ip]) * (mmax + 1) + g.m[ip]] = ip;
  }
  std::vector<float2> eph32(eph.size()), eps32(eps.size());
  for (size_t i = 0; i < eph.size(); ++i) eph32[i] = make_float2(static_cast<float>(eph[i].x), static_cast<float>(eph[i].y));
  for (size_t i = 0; i < eps.size(); ++i) eps32[i] = make_float2(static_cast<float>(eps[i].x), static_cast<float>(eps[i].y));
  std::vector<float> wf32(wf.begin(), wf.end()), wb32(wb.begin(), wb.end());
  ephi64 = upload(eph);
  epsi64 = upload(eps);
  ephi32 = upload(eph32);
  epsi32 = upload(eps32);
  wfwd64 = upload(wf);
  wbwd64 = upload(wb);
  wfwd32 = upload(wf32);
  wbwd32 = upload(wb32);
  ipinfo = upload(info);
  yidx = upload(yi);
  // the weights are w_theta times constants: every io of one theta has the
  // same weight, which the compile-time proj2angl kernel reads once per theta
  for (int io = 0; io < no; ++io)
    throw_if(g.w[io] != g.w[(io / (nphi * npsi)) * nphi * npsi], "internal error: angular weights depend on phi or psi");
  w = upload(g.w);
}

ProjTables::~ProjTables() {
  for (void* q : {static_cast<void*>(ephi64), static_cast<void*>(epsi64), static_cast<void*>(ephi32),
                  static_cast<void*>(epsi32), static_cast<void*>(wfwd64), static_cast<void*>(wbwd64),
                  static_cast<void*>(wfwd32), static_cast<void*>(wbwd32), static_cast<void*>(ipinfo),
                  static_cast<void*>(yidx), static_cast<void*>(w)})
    if (q) (void)hipFree(q);
}


namespace {

// the angular grids with a compile-time proj2angl kernel: mrso = 2 with mmax
// 1..6 and mrso = 1 with mmax 1..3 (larger mrso = 1 grids hold too many S
// registers per lane and take the generic kernel)
template <int MMAX = 1, int MRSO = 1, typename F>
bool dispatch_angular(int mmax, int mrso, F&& f) {
  if constexpr (MMAX > 6) {
    return false;
  } else if constexpr (MRSO > 2 || (MRSO == 1 && MMAX > 3)) {
    return dispatch_angular<MMAX + (MRSO > 2 ? 1 : 0), MRSO + (MRSO > 2 ? -2 : 1)>(mmax, mrso, f);
  } else {
    if (mmax == MMAX && mrso == MRSO) {
      f(std::integral_constant<int, MMAX>{}, std::integral_constant<int, MRSO>{});
      return true;
    }
    return dispatch_angular<MMAX, MRSO + 1>(mmax, mrso, f);
  }
}

long wave_row_blocks(long N) { return (N + kWaveRows - 1) / kWaveRows; }

long wave_blocks(const ProjTables& t, long N) {
  const long nrbx = (wave_row_blocks(N) + kXcds - 1) / kXcds;
  return nrbx * kXcds * t.ntheta;
}

}  // namespace

template <typename TI, typename TP, typename TC>
void angl2proj_factorized(const ProjTables& t, long N, const TI* o, cplx<TP>* p, const double* drho_rho0,
                          hipStream_t s, bool compact) {
  if (N == 0) return;
  const long nblk = (N + kRT - 1) / kRT;
  throw_if(nblk > 0x7fffffffL, "grid too large");
  bool compiled = false;
  dispatch_angular(t.mmax, t.mrso, [&](auto mm, auto mr) {
    constexpr int MMAX = decltype(mm)::value, MRSO = decltype(mr)::value, RT = a2p_ct_rows<TC>();
    using Z = A2pSizes<MMAX, MRSO, RT>;
    const long nblk_ct = (N + RT - 1) / RT;
    constexpr size_t kLds = (Z::kRegA + Z::kNth * Z::kNp()) * sizeof(TC) +
                            (Z::kCells + Z::kNphi + Z::kNmu2 * Z::kNpsi) * sizeof(cplx<TC>);
    if constexpr (kLds <= 64 * 1024) {  // else the run-time kernel below
      if (compact)
        hipLaunchKernelGGL((angl2proj_ct_kernel<MMAX, MRSO, TI, TP, TC, true>), dim3(static_cast<unsigned>(nblk_ct)),
                           dim3(kThreads), 0, s, o, p, N, tabs_of<TC>(t, true), drho_rho0 ? 1 : 0,
                           drho_rho0 ? *drho_rho0 : 0.0);
      else
        hipLaunchKernelGGL((angl2proj_ct_kernel<MMAX, MRSO, TI, TP, TC, false>), dim3(static_cast<unsigned>(nblk_ct)),
                           dim3(kThreads), 0, s, o, p, N, tabs_of<TC>(t, true), drho_rho0 ? 1 : 0,
                           drho_rho0 ? *drho_rho0 : 0.0);
      compiled = true;
    }
  });
  if (compiled) {
    MDFT_HIP_CHECK(hipGetLastError());
    return;
  }
  throw_if(compact, "angl2proj: the compact channel order needs a compile-time angular grid");
  const size_t lds = a2p_lds<TC>(t);
  throw_if(lds > 160 * 1024, "angular grid too large for the projection tile");
  hipLaunchKernelGGL((angl2proj_kernel<TI, TP, TC>), dim3(static_cast<unsigned>(nblk)), dim3(kThreads), lds, s, o, p,
                     N, dims_of(t), tabs_of<TC>(t, true), drho_rho0 ? 1 : 0, drho_rho0 ? *drho_rho0 : 0.0);
  MDFT_HIP_CHECK(hipGetLastError());
}

int proj2angl_blocks(const ProjTables& t, long N) {
  if (dispatch_angular(t.mmax, t.mrso, [](auto, auto) {})) return static_cast<int>(wave_row_blocks(N) * t.ntheta);
  return static_cast<int>((N + kRT - 1) / kRT);
}

bool proj2angl_has_column_sums(const ProjTables& t) { return dispatch_angular(t.mmax, t.mrso, [](auto, auto) {}); }

bool proj2angl_compiled(int mmax, int mrso) { return dispatch_angular(mmax, mrso, [](auto, auto) {}); }

bool angl2proj_compiled(const ProjTables& t, bool fp64) {
  bool compiled = false;
  auto probe = [&](auto tc) {
    using TC = decltype(tc);
    dispatch_angular(t.mmax, t.mrso, [&](auto mm, auto mr) {
      constexpr int MMAX = decltype(mm)::value, MRSO = decltype(mr)::value, RT = a2p_ct_rows<TC>();
      using Z = A2pSizes<MMAX, MRSO, RT>;
      constexpr size_t kLds = (Z::kRegA + Z::kNth * Z::kNp()) * sizeof(TC) +
                              (Z::kCells + Z::kNphi + Z::kNmu2 * Z::kNpsi) * sizeof(cplx<TC>);
      compiled = kLds <= 64 * 1024;
    });
  };
  if (fp64)
    probe(double{});
  else
    probe(float{});
  return compiled;
}

long proj2angl_row_blocks(long N) { return wave_row_blocks(N); }

template <typename TP, typename TO, typename TC>
void proj2angl_factorized(const ProjTables& t, long N, const cplx<TP>* p, TO* o, const TO* xi, TO* df,
                          double rho0, double* partials, hipStream_t s, double* colpart, bool compact) {
  if (N == 0) return;
  const long nwb = wave_blocks(t, N);
  throw_if(nwb > 0x7fffffffL, "grid too large");
  const bool compiled = dispatch_angular(t.mmax, t.mrso, [&](auto mm, auto mr) {
    constexpr int MM = decltype(mm)::value, MR = decltype(mr)::value;
    if (compact)
      hipLaunchKernelGGL((proj2angl_wave_kernel<MM, MR, TP, TO, TC, true>), dim3(static_cast<unsigned>(nwb)),
                         dim3(64), 0, s, p, o, xi, df, t.w, N, wave_row_blocks(N), tabs_of<TC>(t, false), rho0,
                         partials, colpart);
    else
      hipLaunchKernelGGL((proj2angl_wave_kernel<MM, MR, TP, TO, TC, false>), dim3(static_cast<unsigned>(nwb)),
                         dim3(64), 0, s, p, o, xi, df, t.w, N, wave_row_blocks(N), tabs_of<TC>(t, false), rho0,
                         partials, colpart);
  });
  if (compiled) {
    MDFT_HIP_CHECK(hipGetLastError());
    return;
  }
  throw_if(compact, "proj2angl: the compact channel order needs a compile-time angular grid");
  throw_if(colpart != nullptr, "proj2angl: column sums need a compile-time angular grid");
  const long nblk = (N + kRT - 1) / kRT;
  throw_if(nblk > 0x7fffffffL, "grid too large");
  const size_t lds = p2a_lds<TC>(t);
  throw_if(lds > 160 * 1024, "angular grid too large for the projection tile");
  hipLaunchKernelGGL((proj2angl_kernel<TP, TO, TC>), dim3(static_cast<unsigned>(nblk)), dim3(kThreads), lds, s, p, o,
                     xi, df, t.w, N, dims_of(t), tabs_of<TC>(t, false), rho0, partials);
  MDFT_HIP_CHECK(hipGetLastError());
}

template <typename T>
void delta_rho(const T* xi, T* out, size_t n, double rho0, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL((delta_rho_kernel<T>), dim3(grid_for(n, kThreads, 65536)), dim3(kThreads), 0, s, xi, out, n,
                     static_cast<T>(rho0));
  MDFT_HIP_CHECK(hipGetLastError());
}

int delta_f_blocks(long nxyz) { return static_cast<int>((nxyz + kDfRows - 1) / kDfRows); }
int delta_f_rows_per_block() { return kDfRows; }

template <typename T>
int delta_f(const T* xi, const T* vexc, const double* w, T* df, long nxyz, int no, double rho0, double* partials,
            int max_partials, hipStream_t s) {
  const int nb = delta_f_blocks(nxyz);
  throw_if(nb > max_partials, "partials buffer too small");
  if (nb == 0) return 0;
  hipLaunchKernelGGL((delta_f_kernel<T>), dim3(nb), dim3(kThreads), 0, s, xi, vexc, w, df, nxyz, no, rho0, partials);
  MDFT_HIP_CHECK(hipGetLastError());
  return nb;
}

void reduce_partials(const double* partials, int count, double scale, double* out, hipStream_t s, int nbatch,
                     long out_stride, const double* extra, int nextra) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(nbatch), dim3(kRedThreads), 0, s, partials, count, scale, out,
                     out_stride, extra, nextra);
  MDFT_HIP_CHECK(hipGetLastError());
}

template <typename T>
void fill_uniform(T* x, size_t n, unsigned long long seed, unsigned long long offset, double lo, double hi,
                  hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL((fill_uniform_kernel<T>), dim3(grid_for(n, kThreads, 65536)), dim3(kThreads), 0, s, x, n,
                     splitmix64_host(seed), offset, lo, hi);
  MDFT_HIP_CHECK(hipGetLastError());
}

template void angl2proj_factorized<double, double, double>(const ProjTables&, long, const double*, cplx<double>*,
                                                           const double*, hipStream_t, bool);
template void angl2proj_factorized<float, float, float>(const ProjTables&, long, const float*, cplx<float>*,
                                                        const double*, hipStream_t, bool);
template void angl2proj_factorized<double, float, float>(const ProjTables&, long, const double*, cplx<float>*,
                                                         const double*, hipStream_t, bool);
template void proj2angl_factorized<double, double, double>(const ProjTables&, long, const cplx<double>*, double*,
                                                           const double*, double*, double, double*, hipStream_t,
                                                           double*, bool);
template void proj2angl_factorized<float, float, float>(const ProjTables&, long, const cplx<float>*, float*,
                                                        const float*, float*, double, double*, hipStream_t,
                                                        double*, bool);
template void proj2angl_factorized<float, double, float>(const ProjTables&, long, const cplx<float>*, double*,
                                                         const double*, double*, double, double*, hipStream_t,
                                                         double*, bool);
template void delta_rho<double>(const double*, double*, size_t, double, hipStream_t);
template void delta_rho<float>(const float*, float*, size_t, double, hipStream_t);
template int delta_f<double>(const double*, const double*, const double*, double*, long, int, double, double*, int,
                             hipStream_t);
template int delta_f<float>(const float*, const float*, const double*, float*, long, int, double, double*, int,
                            hipStream_t);
template void fill_uniform<double>(double*, size_t, unsigned long long, unsigned long long, double, double,
                                   hipStream_t);
template void fill_uniform<float>(float*, size_t, unsigned long long, unsigned long long, double, double,
                                  hipStream_t);

}  // namespace mdft
