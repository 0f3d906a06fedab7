// Projections of the MDFT_MIXED forms as dense GEMMs on the gfx950 matrix
// cores (v_mfma_f32_16x16x32_bf16), split-BF16 operands, FP32 accumulation.
//
// angl2proj (common/MDFT_OrientationProjectionTransform.hpp:267-380) is the
// real contraction P[2ip + c][r] = sum_io Wa[2ip + c][io] delta_rho[r][io] with
//   Wa[2ip][io] + i Wa[2ip+1][io] = f_m w_theta d^m_{mup,mu}(theta)
//                                   e^{i(mup phi_j + mu psi_k)} / (nphi npsi),
// i.e. the rfft2 / unscramble / theta quadrature of the reference folded into
// one [2np x no] matrix (computed once in FP64 on the host).  proj2angl
// (:388-484) is vexc[r][io] = sum_k P[r][k] Wp[k][io] with
//   Wp[2ip][io] = c_mu2 d f_m cos(alpha), Wp[2ip+1][io] = c_mu2 d f_m sin(alpha),
//   alpha = mup phi_j + 2 pi mu2 k / npsi, c_0 = 1, c_{mu2>0} = 2
// (the c2r convention: Re of the DC psi bin, twice Re of the others).
//
// Split precision: the data x (FP32) is a sum of PX BF16 parts and the
// weights w (FP64 on the host) a sum of PW parts, each part the nearest BF16
// of the remaining residual; the product keeps the terms x_i w_j with
// i + j < max(PX, PW) (0-based):
//   PX = PW = 2 (BF16x3): x1 w1 + x1 w2 + x2 w1     -- about 16-bit operands
//   PX = PW = 3 (BF16x6): + x1 w3 + x2 w2 + x3 w1   -- data exact to FP32,
//                                                      weights to ~2^-26
// Every MFMA accumulates in FP32.  The weights are the same numbers for every
// grid point, so their rounding error does not average out over the grid:
// weights rounded to FP32 alone move F by ~eps_FP32 |F|, which is why they
// are split from FP64.  proj2angl on the matrix cores is compiled for BF16x3
// only: with exact operands its F error stays at ~3 eps_FP32 |F| (the MFMA's
// internal sum of 32 products is not an unbiased round-to-nearest, and vexc
// feeds F directly), three times that of the FP32 VALU kernel the BF16x6
// form uses.
//
// Tiling: 4 waves per block; MT output tiles of 16 (a compile-time bucket, the
// weight matrix is zero-padded to it).  The weight matrix is pre-swizzled on
// the host into the per-lane fragment order of the MFMA operand it feeds; each
// K-step's slice is staged once per block in LDS: it is fetched into
// registers while the previous slice feeds the MFMAs, then written to LDS
// between two barriers.  The data operand is read straight from HBM by the
// lanes that own it, one K-step ahead (delta_rho = rho0 (xi^2 - 1) is formed
// in the load for the fused evaluation; get_delta_f runs in the proj2angl
// epilogue).  Loads are unconditional: row and K indices are clamped into the
// arrays and the padded K range meets zero weights.  angl2proj computes the
// transposed product so that each lane holds (re, im) of one projection at
// consecutive r (coalesced stores); proj2angl keeps r as the row index so that
// xi / df are accessed in runs of consecutive io.
#include "mdft_gemm.hpp"

#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

namespace mdft {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// a native vector type: arrays of HIP's uint4 class are kept in scratch
// memory instead of registers
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;  // waves per block: rows sharing one weight slice
constexpr int kThreads = 64 * kWaves;
constexpr int kK = 32;  // K per MFMA
// angl2proj: row groups of 16 per wave.  Two groups share each LDS weight
// fragment over twice the rows; with the 10 output tiles of the compact
// channel order (mmax 5) they measured 0.163 against 0.182 ms for one group
// and 0.171 / 0.209 ms for three / four.  With the full order's 12 tiles two
// groups needed 228 registers and measured 4% slower, so wider tile counts
// keep one group.
template <int MT>
constexpr int a2p_groups() {
  return MT <= 10 ? 2 : 1;
}
constexpr int kP2aRowsPerWave = 16;
constexpr int kMaxA2pTiles = 16;  // 2 np <= 256
constexpr int kMaxP2aTiles = 24;  // no <= 384

// weight fragment of one (k-step, tile, lane): PW parts of 8 x bf16
template <int PW>
struct Frag {
  u32x4 p[PW];
};

// output tiles rounded up to the compiled buckets (even counts)
inline int tile_bucket(int tiles) { return (tiles + 1) & ~1; }

__host__ __device__ inline uint32_t bf16_bits_rne(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}
__host__ __device__ inline float bf16_value(uint32_t b) { return __builtin_bit_cast(float, b << 16); }

// x -> x1 + ... + xP, each part the nearest BF16 of the remaining residual
// (the residual is exact in the type of x)
template <int P, typename T>
__host__ __device__ inline void split_bf16(T x, uint32_t (&part)[P]) {
  T r = x;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    part[i] = bf16_bits_rne(static_cast<float>(r));
    r -= static_cast<T>(bf16_value(part[i]));
  }
}

__device__ inline bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

// acc + x * w over the kept terms, smallest first
template <int PX, int PW, bool W_IS_A>
__device__ inline f32x4 mfma_split(const u32x4 (&x)[PX], const Frag<PW>& w, f32x4 acc) {
  constexpr int kMaxLvl = (PX > PW ? PX : PW) - 1;
#pragma unroll
  for (int lvl = kMaxLvl; lvl >= 0; --lvl)
#pragma unroll
    for (int i = 0; i < PX; ++i) {
      const int j = lvl - i;
      if (j >= 0 && j < PW) {
        if (W_IS_A)
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(w.p[j]), as_bf16x8(x[i]), acc, 0, 0, 0);
        else
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(x[i]), as_bf16x8(w.p[j]), acc, 0, 0, 0);
      }
    }
  return acc;
}

// pack 8 floats into PX BF16 fragments
template <int PX>
__device__ inline void pack8(const float (&v)[8], u32x4 (&out)[PX]) {
  uint32_t b[PX][8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    uint32_t pr[PX];
    split_bf16<PX>(v[j], pr);
#pragma unroll
    for (int q = 0; q < PX; ++q) b[q][j] = pr[q];
  }
#pragma unroll
  for (int q = 0; q < PX; ++q)
    out[q] = u32x4{b[q][0] | (b[q][1] << 16), b[q][2] | (b[q][3] << 16), b[q][4] | (b[q][5] << 16),
                   b[q][6] | (b[q][7] << 16)};
}

// pack 8 floats into three BF16 fragments by truncation: hi = the top 16
// bits of x (sign, exponent, 8 significant bits), mid = the top 16 bits of
// x - hi, lo = x - hi - mid.  Both subtractions are exact, and an FP32
// significand has 24 bits, so lo has at most 8 significant bits and
// x = hi + mid + lo exactly: the same value as the nearest-BF16 split of
// pack8<3>, for about a third of its VALU work (no rounding arithmetic;
// v_perm_b32 joins the halves).
__device__ inline void pack8_exact3(const float (&v)[8], u32x4 (&out)[3]) {
  uint32_t h[8], m[8], l[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t u = __builtin_bit_cast(uint32_t, v[j]);
    h[j] = u & 0xffff0000u;
    const float r1 = v[j] - __builtin_bit_cast(float, h[j]);
    m[j] = __builtin_bit_cast(uint32_t, r1) & 0xffff0000u;
    l[j] = __builtin_bit_cast(uint32_t, r1 - __builtin_bit_cast(float, m[j]));
  }
  // result = (high half of the odd element) << 16 | (high half of the even one)
  constexpr uint32_t kSel = 0x07060302u;
  out[0] = u32x4{__builtin_amdgcn_perm(h[1], h[0], kSel), __builtin_amdgcn_perm(h[3], h[2], kSel),
                 __builtin_amdgcn_perm(h[5], h[4], kSel), __builtin_amdgcn_perm(h[7], h[6], kSel)};
  out[1] = u32x4{__builtin_amdgcn_perm(m[1], m[0], kSel), __builtin_amdgcn_perm(m[3], m[2], kSel),
                 __builtin_amdgcn_perm(m[5], m[4], kSel), __builtin_amdgcn_perm(m[7], m[6], kSel)};
  out[2] = u32x4{__builtin_amdgcn_perm(l[1], l[0], kSel), __builtin_amdgcn_perm(l[3], l[2], kSel),
                 __builtin_amdgcn_perm(l[5], l[4], kSel), __builtin_amdgcn_perm(l[7], l[6], kSel)};
}

struct GemmDims {
  long N;
  int no, np, kt;  // kt: k-steps of 32
};

// One K-step slice of swizzled weights (MT tiles x 64 lanes x PW x 16 bytes,
// contiguous in global memory) staged through registers into LDS.  The
// global array is padded so that the last slice may be over-read; the LDS
// slice holds kPer * kThreads u32x4.
template <int PW, int MT>
struct Stage {
  static constexpr int kPer = (MT * 64 * PW + kThreads - 1) / kThreads;
  u32x4 r[kPer];
  __device__ inline void load(const u32x4* __restrict__ src, int tid) {
#pragma unroll
    for (int i = 0; i < kPer; ++i) r[i] = src[tid + i * kThreads];
  }
  __device__ inline void store(u32x4* dst, int tid) const {
#pragma unroll
    for (int i = 0; i < kPer; ++i) dst[tid + i * kThreads] = r[i];
  }
};

// P^T = Wa . delta_rho^T.  Block: 16 G kWaves consecutive r, G = a2p_groups<MT>;
// wave w: r0 + 16 G w + (0 .. 16 G - 1), in groups u of 16.
// TI: the element type of o / xi (double, or float for the FP32 fields of
// MDFT_MIXED_X6F; delta_rho is formed in FP64 either way).
// EVEN: no is even, so a row's K pairs are aligned two-element loads.
template <typename TI, int PX, int PW, int MT, bool FUSE, bool EVEN>
__global__ __launch_bounds__(kThreads) void a2p_gemm_kernel(const TI* __restrict__ o, float2* __restrict__ p,
                                                            const u32x4* __restrict__ wsw, GemmDims d, double rho0) {
  using St = Stage<PW, MT>;
  constexpr int kA2pGroups = a2p_groups<MT>();
  constexpr int kRows = 16 * kA2pGroups;
  __shared__ u32x4 lds[St::kPer * kThreads];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long r0 = static_cast<long>(blockIdx.x) * (kWaves * kRows) + wave * kRows;
  const int slice = MT * 64 * PW;  // u32x4 per K-step in global memory
  f32x4 acc[MT][kA2pGroups];
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int u = 0; u < kA2pGroups; ++u) acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};

  // rows r0 + 16 u + lane%16 (clamped), k = 32 s + 8 (lane/16) + (0..7) (clamped)
  const TI* rowp[kA2pGroups];
#pragma unroll
  for (int u = 0; u < kA2pGroups; ++u) {
    const long r = r0 + 16 * u + (lane & 15);
    rowp[u] = o + (r < d.N ? r : d.N - 1) * d.no;
  }
  auto load_data = [&](int s, TI (&raw)[kA2pGroups][8]) {
    // operand slot j of lane group g = lane/16 holds k = 32 s + 8 (j/2) + 2 g + j%2
    // (a2p_k_of): load q of the four lanes of a row then reads 64 contiguous bytes
    const int k0 = s * kK + 2 * (lane >> 4);
#pragma unroll
    for (int u = 0; u < kA2pGroups; ++u) {
      if (EVEN) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int k = min(k0 + 8 * q, d.no - 2);
          using V2 = typename cvec<TI>::type;
          const V2 v = *reinterpret_cast<const V2*>(rowp[u] + k);
          raw[u][2 * q] = v.x;
          raw[u][2 * q + 1] = v.y;
        }
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) raw[u][j] = rowp[u][min(k0 + 8 * (j >> 1) + (j & 1), d.no - 1)];
      }
    }
  };
  // delta_rho = rho0 (xi^2 - 1): in FP64 for double xi; for float xi as
  // fma(xi, xi, -1) (one rounding) times rho0 split in two floats (rho0 to
  // ~2^-48: rounding rho0 to FP32 would scale every projection alike)
  const float rho0_hi = static_cast<float>(rho0);
  const float rho0_lo = static_cast<float>(rho0 - static_cast<double>(rho0_hi));
  auto pack_data = [&](const TI (&raw)[kA2pGroups][8], u32x4 (&x)[kA2pGroups][PX]) {
#pragma unroll
    for (int u = 0; u < kA2pGroups; ++u) {
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if constexpr (std::is_same<TI, float>::value) {
          const float xv = raw[u][j];
          if (FUSE) {
            const float t = fmaf(xv, xv, -1.0f);
            v[j] = fmaf(t, rho0_hi, t * rho0_lo);
          } else {
            v[j] = xv;
          }
        } else {
          double val = raw[u][j];
          if (FUSE) val = rho0 * (val * val - 1.0);
          v[j] = static_cast<float>(val);
        }
      }
      // float fields: the truncation split (their delta_rho costs no FP64
      // work, so the split is most of the VALU work); double fields keep the
      // nearest-BF16 split (no faster with the other, their FP64 work dominates)
      if constexpr (PX == 3 && std::is_same<TI, float>::value)
        pack8_exact3(v, x[u]);
      else
        pack8<PX>(v, x[u]);
    }
  };

  St st;
  TI raw[kA2pGroups][8];
  u32x4 x[kA2pGroups][PX];
  st.load(wsw, tid);
  load_data(0, raw);
  st.store(lds, tid);
  pack_data(raw, x);
  __syncthreads();
  for (int s = 0; s < d.kt; ++s) {
    const bool more = s + 1 < d.kt;
    if (more) {
      st.load(wsw + static_cast<long>(s + 1) * slice, tid);
      load_data(s + 1, raw);
    }
    const Frag<PW>* wl = reinterpret_cast<const Frag<PW>*>(lds);
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      const Frag<PW> f = wl[t * 64 + lane];
#pragma unroll
      for (int u = 0; u < kA2pGroups; ++u) acc[t][u] = mfma_split<PX, PW, true>(x[u], f, acc[t][u]);
    }
    if (more) {
      __syncthreads();
      st.store(lds, tid);
      pack_data(raw, x);
      __syncthreads();
    }
  }
  // acc[t][u][j] = P^T[m = 16 t + 4 (lane/16) + j][r = r0 + 16 u + lane%16];
  // rows m = 2 ip + (re, im): j = 0, 1 -> ip = 8 t + 2 (lane/16); j = 2, 3 -> ip + 1
#pragma unroll
  for (int t = 0; t < MT; ++t) {
#pragma unroll
    for (int u = 0; u < kA2pGroups; ++u) {
      const long r = r0 + 16 * u + (lane & 15);
      const int ip = 8 * t + 2 * (lane >> 4);
      if (r < d.N) {
        if (ip < d.np) p[static_cast<long>(ip) * d.N + r] = make_float2(acc[t][u][0], acc[t][u][1]);
        if (ip + 1 < d.np) p[static_cast<long>(ip + 1) * d.N + r] = make_float2(acc[t][u][2], acc[t][u][3]);
      }
    }
  }
}

// vexc = P . Wp.  Block: 16 kWaves consecutive r; wave w: r0 + 16 w + (0..15).
template <int PX, int PW, int MT>
__global__ __launch_bounds__(kThreads) void p2a_gemm_kernel(const float2* __restrict__ p, double* __restrict__ o,
                                                            const double* __restrict__ xi, double* __restrict__ df,
                                                            const double* __restrict__ wq,
                                                            const u32x4* __restrict__ wsw, GemmDims d, double rho0,
                                                            double* __restrict__ partials) {
  using St = Stage<PW, MT>;
  __shared__ u32x4 lds[St::kPer * kThreads];
  __shared__ double red[kThreads];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long r0 = static_cast<long>(blockIdx.x) * (kWaves * kP2aRowsPerWave) + wave * kP2aRowsPerWave;
  const int slice = MT * 64 * PW;
  f32x4 acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};

  const long rl0 = r0 + (lane & 15);
  const long rl = rl0 < d.N ? rl0 : d.N - 1;
  // k-step s: row rl, k = 32 s + 8 (lane/16) + (0..7) -> projections ip0 .. ip0+3 (re, im)
  auto load_data = [&](int s, float2 (&raw)[4]) {
    const int ip0 = (s * kK + 8 * (lane >> 4)) / 2;
#pragma unroll
    for (int j = 0; j < 4; ++j) raw[j] = p[static_cast<long>(min(ip0 + j, d.np - 1)) * d.N + rl];
  };
  auto pack_data = [&](const float2 (&raw)[4], u32x4 (&x)[PX]) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[2 * j] = raw[j].x;
      v[2 * j + 1] = raw[j].y;
    }
    pack8<PX>(v, x);
  };

  St st;
  float2 raw[4];
  u32x4 x[PX];
  st.load(wsw, tid);
  load_data(0, raw);
  st.store(lds, tid);
  pack_data(raw, x);
  __syncthreads();
  for (int s = 0; s < d.kt; ++s) {
    const bool more = s + 1 < d.kt;
    if (more) {
      st.load(wsw + static_cast<long>(s + 1) * slice, tid);
      load_data(s + 1, raw);
    }
    const Frag<PW>* wl = reinterpret_cast<const Frag<PW>*>(lds);
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      const Frag<PW> f = wl[t * 64 + lane];
      acc[t] = mfma_split<PX, PW, false>(x, f, acc[t]);
    }
    if (more) {
      __syncthreads();
      st.store(lds, tid);
      pack_data(raw, x);
      __syncthreads();
    }
  }
  // acc[t][j] = vexc[r = r0 + 4 (lane/16) + j][io = 16 t + lane%16]
  double ff = 0.0;
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    const int io = 16 * t + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long r = r0 + 4 * (lane >> 4) + j;
      if (r < d.N && io < d.no) {
        const long g = r * d.no + io;
        const double v = static_cast<double>(acc[t][j]);
        if (xi == nullptr) {
          o[g] = v;
        } else {
          const double x = xi[g];
          ff += v * wq[io] * (x * x - 1.0);
          df[g] = 2.0 * rho0 * x * v;
        }
      }
    }
  }
  if (xi != nullptr) {
    red[tid] = ff;
    __syncthreads();
    for (int s2 = kThreads / 2; s2 > 0; s2 >>= 1) {
      if (tid < s2) red[tid] += red[tid + s2];
      __syncthreads();
    }
    if (tid == 0) partials[blockIdx.x] = red[0];
  }
}

// The K index of angl2proj's operand slot j (0..7) of lane l in K-step s:
// the slots of lane group l/16 interleave with the other groups' in pairs,
// so that the four lanes of one row load 64 contiguous bytes per 16-byte
// load (the sum over K is the same for any such permutation, applied to the
// weights and the data alike).
inline int a2p_k_of(int s, int l, int j) { return s * kK + 8 * (j >> 1) + 2 * (l >> 4) + (j & 1); }

// Pre-swizzle a dense FP64 matrix into PW-part MFMA fragments over `tiles`
// output tiles.  As the A operand (angl2proj's weights; rows = m, cols = k):
// slot j of lane l of tile t, k-step s holds M[16t + l%16][a2p_k_of(s, l, j)];
// as the B operand (rows = k, cols = n): M[32s + 8(l/16) + j][16t + l%16].
// Out-of-range entries are 0.  The
// array is followed by kThreads zero vectors: the staging of the last slice
// reads fewer than kThreads past its end.
template <int PW>
std::vector<u32x4> swizzle(const std::vector<double>& M, int rows, int cols, bool as_a, int tiles) {
  const int ksteps = as_a ? (cols + kK - 1) / kK : (rows + kK - 1) / kK;
  const size_t nfrag = static_cast<size_t>(ksteps) * tiles * 64;
  std::vector<u32x4> out(nfrag * PW + kThreads, u32x4{0, 0, 0, 0});
  for (int s = 0; s < ksteps; ++s)
    for (int t = 0; t < tiles; ++t)
      for (int l = 0; l < 64; ++l) {
        uint32_t parts[PW][8] = {};
        for (int j = 0; j < 8; ++j) {
          const int kk = as_a ? a2p_k_of(s, l, j) : s * kK + 8 * (l >> 4) + j, mn = 16 * t + (l & 15);
          const int rr = as_a ? mn : kk, cc = as_a ? kk : mn;
          const double v = (rr < rows && cc < cols) ? M[static_cast<size_t>(rr) * cols + cc] : 0.0;
          uint32_t pr[PW];
          split_bf16<PW>(v, pr);
          for (int q = 0; q < PW; ++q) parts[q][j] = pr[q];
        }
        const size_t f = (static_cast<size_t>(s) * tiles + t) * 64 + l;
        for (int q = 0; q < PW; ++q)
          out[f * PW + q] = u32x4{parts[q][0] | (parts[q][1] << 16), parts[q][2] | (parts[q][3] << 16),
                                  parts[q][4] | (parts[q][5] << 16), parts[q][6] | (parts[q][7] << 16)};
      }
  return out;
}

template <typename T>
T* upload(const std::vector<T>& h) {
  T* d = nullptr;
  MDFT_HIP_CHECK(hipMalloc(&d, std::max<size_t>(1, h.size()) * sizeof(T)));
  if (!h.empty()) MDFT_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

// compile-time dispatch of the even tile buckets 2 .. MAXT
template <int MAXT, int T = 2, typename F>
void dispatch_tiles(int tiles, F&& f) {
  if constexpr (T > MAXT) {
    throw_if(true, "projection GEMM: no kernel for this tile count");
  } else {
    if (tiles == T)
      f(std::integral_constant<int, T>{});
    else
      dispatch_tiles<MAXT, T + 2>(tiles, f);
  }
}

// the compiled angl2proj (data parts, weight parts) pairs
template <typename F>
void dispatch_split(int px, int pw, F&& f) {
  if (px == 2 && pw == 2)
    f(std::integral_constant<int, 2>{}, std::integral_constant<int, 2>{});
  else if (px == 3 && pw == 3)
    f(std::integral_constant<int, 3>{}, std::integral_constant<int, 3>{});
  else
    throw_if(true, "projection GEMM: unsupported BF16 split");
}

}  // namespace

void projection_matrices(const AngularGrid& g, std::vector<double>& wa, std::vector<double>& wp) {
  const int nt = g.ntheta, nph = g.nphi, nps = g.npsi, np = g.np, no = g.no;
  wa.assign(static_cast<size_t>(2 * np) * no, 0.0);
  wp.assign(static_cast<size_t>(2 * np) * no, 0.0);
  const double inv = 1.0 / static_cast<double>(nph * nps);
  for (int ip = 0; ip < np; ++ip) {
    const int m = g.m[ip], mup = g.mup[ip], mu2 = g.mu2[ip];
    for (int th = 0; th < nt; ++th) {
      const double dv = g.wigner_d[th * np + ip];
      const double af = g.fm[m] * g.wtheta[th] * dv * inv;
      const double bf = dv * g.fm[m] * (mu2 == 0 ? 1.0 : 2.0);
      for (int j = 0; j < nph; ++j)
        for (int k = 0; k < nps; ++k) {
          // angles reduced exactly: mup j mod nphi and mu2 k mod npsi
          const int tj = ((mup * j) % nph + nph) % nph, tk = (mu2 * k) % nps;
          const double a = 2.0 * M_PI * (static_cast<double>(tj) / nph + static_cast<double>(tk) / nps);
          const int io = (th * nph + j) * nps + k;
          wa[static_cast<size_t>(2 * ip) * no + io] = af * std::cos(a);
          wa[static_cast<size_t>(2 * ip + 1) * no + io] = af * std::sin(a);
          wp[static_cast<size_t>(2 * ip) * no + io] = bf * std::cos(a);
          wp[static_cast<size_t>(2 * ip + 1) * no + io] = bf * std::sin(a);
        }
    }
  }
}

struct ProjGemm::Impl {
  int px, pw;
  bool p2a_mfma;  // proj2angl on the matrix cores (BF16x3) or the FP32 VALU kernel
  bool compact = false;
  int no, np;  // np: stored projections (compact: the compact count)
  int a_mt, a_kt, p_mt, p_kt;      // tile buckets and k-steps
  bool a2p_fits, p2a_fits;         // grid within the MFMA register tiles
  std::unique_ptr<ProjTables> fb;  // factorised FP32 projections
  u32x4* wa = nullptr;
  u32x4* wp = nullptr;
  double* wq = nullptr;
  bool a2p_gemm() const { return a2p_fits; }
  bool p2a_gemm() const { return p2a_mfma && p2a_fits; }
  ~Impl() {
    if (wa) (void)hipFree(wa);
    if (wp) (void)hipFree(wp);
    if (wq) (void)hipFree(wq);
  }
};

ProjGemm::ProjGemm(const AngularGrid& g, int data_parts, int weight_parts, bool p2a_mfma, bool compact)
    : impl_(std::make_unique<Impl>()) {
  dispatch_split(data_parts, weight_parts, [](auto, auto) {});  // rejects pairs without kernels
  throw_if(p2a_mfma && (data_parts != 2 || weight_parts != 2), "proj2angl on MFMA is compiled for BF16x3 only");
  Impl& I = *impl_;
  I.px = data_parts;
  I.pw = weight_parts;
  I.p2a_mfma = p2a_mfma;
  I.no = g.no;
  I.compact = compact;
  I.np = compact ? compact_projections(g) : g.np;
  const int nps = g.npsi, np = I.np, no = g.no;
  throw_if(nps % 2 == 0, "npsi must be odd");
  I.a_mt = tile_bucket((2 * np + 15) / 16);
  I.a_kt = (no + kK - 1) / kK;
  I.p_mt = tile_bucket((no + 15) / 16);
  I.p_kt = (2 * np + kK - 1) / kK;
  I.a2p_fits = I.a_mt <= kMaxA2pTiles && no >= 2;
  I.p2a_fits = I.p_mt <= kMaxP2aTiles;
  if (!I.a2p_gemm() || !I.p2a_gemm()) I.fb = std::make_unique<ProjTables>(g);
  throw_if(compact && !compact_possible(g, p2a_mfma), "projections: no compact channel order for this grid and form");
  std::vector<double> wa, wp;
  projection_matrices(g, wa, wp);
  if (compact) {
    // the angl2proj rows (2 ip + re/im) of the stored projections, in order
    const std::vector<int> code = compact_codes(g);
    std::vector<double> wc;
    wc.reserve(static_cast<size_t>(2 * np) * no);
    for (int ip = 0; ip < g.np; ++ip)
      if (!(code[ip] & kCompactImplied))
        wc.insert(wc.end(), wa.begin() + static_cast<size_t>(2 * ip) * no, wa.begin() + static_cast<size_t>(2 * ip + 2) * no);
    throw_if(wc.size() != static_cast<size_t>(2 * np) * no, "internal error: compact angl2proj rows");
    wa.swap(wc);
    // proj2angl: the same rows (a stored channel holds the fold of its pair,
    // which proj2angl weighs with the stored projection's own row)
    std::vector<double> pc;
    pc.reserve(static_cast<size_t>(2 * np) * no);
    for (int ip = 0; ip < g.np; ++ip)
      if (!(code[ip] & kCompactImplied))
        pc.insert(pc.end(), wp.begin() + static_cast<size_t>(2 * ip) * no, wp.begin() + static_cast<size_t>(2 * ip + 2) * no);
    wp.swap(pc);
  }
  if (I.a2p_gemm())
    dispatch_split(I.px, I.pw, [&](auto, auto pwc) {
      I.wa = upload(swizzle<decltype(pwc)::value>(wa, 2 * np, no, true, I.a_mt));
    });
  if (I.p2a_gemm()) {
    I.wp = upload(swizzle<2>(wp, 2 * np, no, false, I.p_mt));
    I.wq = upload(g.w);
  }
}

ProjGemm::~ProjGemm() = default;

template <typename TI>
void ProjGemm::angl2proj_t(long N, const TI* o, cplx<float>* p, const double* rho0, hipStream_t s) {
  if (N == 0) return;
  const Impl& I = *impl_;
  if (!I.a2p_gemm()) {
    angl2proj_factorized<TI, float, float>(*I.fb, N, o, p, rho0, s);
    return;
  }
  const GemmDims d{N, I.no, I.np, I.a_kt};
  const double r0 = rho0 ? *rho0 : 0.0;
  const bool even = I.no % 2 == 0;
  long nblk = 0;
  auto launch = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(nblk)), dim3(kThreads), 0, s, o, p, I.wa, d, r0);
  };
  dispatch_split(I.px, I.pw, [&](auto pxc, auto pwc) {
    constexpr int PX = decltype(pxc)::value, PW = decltype(pwc)::value;
    dispatch_tiles<kMaxA2pTiles>(I.a_mt, [&](auto mt) {
      constexpr int MT = decltype(mt)::value;
      const long rows = kWaves * 16 * a2p_groups<MT>();
      nblk = (N + rows - 1) / rows;
      throw_if(nblk > 0x7fffffffL, "grid too large");
      if (rho0)
        even ? launch(a2p_gemm_kernel<TI, PX, PW, MT, true, true>)
             : launch(a2p_gemm_kernel<TI, PX, PW, MT, true, false>);
      else
        even ? launch(a2p_gemm_kernel<TI, PX, PW, MT, false, true>)
             : launch(a2p_gemm_kernel<TI, PX, PW, MT, false, false>);
    });
  });
  MDFT_HIP_CHECK(hipGetLastError());
}

void ProjGemm::angl2proj(long N, const double* o, cplx<float>* p, const double* rho0, hipStream_t s) {
  angl2proj_t<double>(N, o, p, rho0, s);
}

void ProjGemm::angl2proj(long N, const float* o, cplx<float>* p, const double* rho0, hipStream_t s) {
  throw_if(impl_->px != 3, "angl2proj: float fields take the BF16x6 split (data exact to FP32)");
  angl2proj_t<float>(N, o, p, rho0, s);
}

int ProjGemm::p2a_blocks(long N) const {
  if (!impl_->p2a_gemm()) return proj2angl_blocks(*impl_->fb, N);
  const long rows = kWaves * kP2aRowsPerWave;
  return static_cast<int>((N + rows - 1) / rows);
}

bool ProjGemm::has_column_sums() const { return !impl_->p2a_gemm() && proj2angl_has_column_sums(*impl_->fb); }

bool ProjGemm::compact() const { return impl_->compact; }

bool ProjGemm::compact_possible(const AngularGrid& g, bool p2a_mfma) {
  // the GEMM angl2proj (register tiles), and the GEMM or the compile-time
  // FP32 proj2angl
  const int nc = compact_projections(g);
  if (nc == g.np) return false;
  if (tile_bucket((2 * nc + 15) / 16) > kMaxA2pTiles || g.no < 2) return false;
  if (p2a_mfma) return tile_bucket((g.no + 15) / 16) <= kMaxP2aTiles;
  return proj2angl_compiled(g.mmax, g.mrso);
}

int ProjGemm::proj2angl(long N, const cplx<float>* p, double* o, const double* xi, double* df, double rho0,
                        double* partials, hipStream_t s, double* colpart) {
  if (N == 0) return 0;
  const Impl& I = *impl_;
  throw_if(colpart != nullptr && !has_column_sums(), "proj2angl: no column sums in this form");
  if (!I.p2a_gemm()) {
    proj2angl_factorized<float, double, float>(*I.fb, N, p, o, xi, df, rho0, partials, s, colpart, I.compact);
    return proj2angl_blocks(*I.fb, N);
  }
  const GemmDims d{N, I.no, I.np, I.p_kt};
  const int nblk = p2a_blocks(N);
  dispatch_tiles<kMaxP2aTiles>(I.p_mt, [&](auto mt) {
    hipLaunchKernelGGL((p2a_gemm_kernel<2, 2, decltype(mt)::value>), dim3(nblk), dim3(kThreads), 0, s, p, o, xi, df,
                       I.wq, I.wp, d, rho0, partials);
  });
  MDFT_HIP_CHECK(hipGetLastError());
  return nblk;
}

int ProjGemm::proj2angl(long N, const cplx<float>* p, float* o, const float* xi, float* df, double rho0,
                        double* partials, hipStream_t s, double* colpart) {
  if (N == 0) return 0;
  const Impl& I = *impl_;
  throw_if(I.p2a_gemm(), "proj2angl: float fields take the FP32 VALU kernel");
  throw_if(colpart != nullptr && !has_column_sums(), "proj2angl: no column sums in this form");
  proj2angl_factorized<float, float, float>(*I.fb, N, p, o, xi, df, rho0, partials, s, colpart, I.compact);
  return proj2angl_blocks(*I.fb, N);
}

}  // namespace mdft
