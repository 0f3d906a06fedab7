// Convolution::execute for gfx950 (common/MDFT_Convolution.hpp:324-548).
//
// The forward and backward 3D FFTs are the Stockham passes of mdft_fft.hip.
// The per-(q, -q) pass (the reference's "convolution" team kernel, :376-543)
// runs one wavefront per pair:
//   1. R^l_{m,m1}(q) by the reference recursion (common/MDFT_Rotation.hpp:
//      103-283) in FP64, parallel over (m, m1) inside each level l, stored
//      compactly (|m|, |m1| <= l); entries the recursion reads outside a block
//      are taken as 0 (their coefficients vanish).  R is the Wigner D-matrix
//      of the q frame, R^l_{m,m1} = e^{-i m phi_q} d^l_{m,m1}(theta_q), so it
//      is split into the phases e^{-i m phi_q} and the real matrices
//      d = Re(R e^{+i m phi_q}); both rotations are then a phase and a real
//      matrix-vector product, half the work of the complex product;
//   2. the rotation to the q frame (:427-443), the c(|q|) interpolation
//      (:449-463, weights in FP64, one coalesced pass over the two table rows
//      into LDS), the OZ product (:468-496), the chop of gamma (:514-517), the
//      inverse rotation and the q / -q stores (:518-541, conjugated -q store
//      at self-conjugate mid-k points).
// The projection index is computed, not looked up: ip(m, mup, mu2) =
// base_m + (mup + m)(m/mrso + 1) + mu2.  Pairs are processed in (ix, iy, iz)
// memory order (same set and q / -q designation as the reference list), and
// the block -> pair map gives every XCD a contiguous range of pairs so that
// neighbouring pairs share L2 lines.
#include "mdft_conv.hpp"
#include "mdft_proj.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <type_traits>

namespace mdft {

namespace {

constexpr int kWave = 64;
constexpr int kXcds = 8;
constexpr int kMaxM = 6;
// compact d entries of l <= kMaxM (sum (2l+1)^2 = 455), rounded up to whole waves
constexpr int kMaxR = 512;
constexpr int kCBatch = 8;  // c(q) rows per lane in flight while staging

template <typename T>
struct OzArgs {
  int nx, ny, nz, mmax, mrso, np, np_new, nq;
  double dq;
  long npairs, chunk, N;
  long pstride;  // elements between projections: N, or nbatch N for a batch of fields
  int nbatch;    // fields in the batch (job j at offset j N within each projection)
  const double* kxyz;
  const int* pairs;
  const int* ipinfo;    // m | (mup + mmax) << 8 | mu2 << 16
  const int* ia_start;  // np + 1
  const cplx<T>* ctab;  // (nq, np_new)
  const double* rc;     // a (5x13x13) | b (5x13x13) | c (5x13) | d (5x13)
  T chop_eps;
};

// The block -> (pair, field) map of the OZ pass.  Launch index L = 8 rest +
// xcd runs on XCD xcd; rest = nbatch k + job.  The single-field block
// b0 = 8 k + xcd takes pair (b0 mod 8) chunk + b0 / 8, so each XCD owns a
// contiguous pair range and the nbatch fields of one pair run back to back on
// the same XCD (their d-table rows are read once into its L2).
template <typename A>
__device__ __forceinline__ long oz_block_pair(const A& a, int& job) {
  const long L = blockIdx.x;
  const long xcd = L % kXcds, rest = L / kXcds;
  job = static_cast<int>(rest % a.nbatch);
  const long b0 = (rest / a.nbatch) * kXcds + xcd;
  return (b0 % kXcds) * a.chunk + b0 / kXcds;
}

__device__ __forceinline__ int roff(int l) { return l * (4 * l * l - 1) / 3; }

__device__ __forceinline__ double2 rget(const double2* R, int l, int m, int m1) {
  if (m < -l || m > l || m1 < -l || m1 > l) return make_double2(0.0, 0.0);
  return R[roff(l) + (m + l) * (2 * l + 1) + (m1 + l)];
}
__device__ __forceinline__ void rset(double2* R, int l, int m, int m1, double2 v) {
  R[roff(l) + (m + l) * (2 * l + 1) + (m1 + l)] = v;
}

// prevent_underflow (common/MDFT_Math_Utils.hpp:126-137): absolute threshold per component
template <typename T>
__device__ __forceinline__ cplx<T> chop(cplx<T> v, T eps) {
  if (fabs(v.x) < eps) v.x = 0;
  if (fabs(v.y) < eps) v.y = 0;
  return v;
}

// R(q) for levels 0..mmax into R (FP64, compact), one wave.
__device__ void rotation_recursion(double2* R, const double* rc, int mmax, double q0, double q1, double q2,
                                   int lane) {
  if (lane == 0) {
    double ez0, ez1, ez2, ey0, ey1, ey2;
    if (q0 == 0.0 && q1 == 0.0 && q2 == 0.0) {
      ez0 = 0.0; ez1 = 0.0; ez2 = 1.0;
    } else {
      ez0 = q0; ez1 = q1; ez2 = q2;
    }
    if (ez0 != 0.0 || ez1 != 0.0) {  // y' = z x z'
      ey0 = -ez1; ey1 = ez0; ey2 = 0.0;
    } else {  // y' = z' x x
      ey0 = 0.0; ey1 = ez2; ey2 = -ez1;
    }
    double s = sqrt(ez0 * ez0 + ez1 * ez1 + ez2 * ez2);
    ez0 /= s; ez1 /= s; ez2 /= s;
    s = sqrt(ey0 * ey0 + ey1 * ey1 + ey2 * ey2);
    ey0 /= s; ey1 /= s; ey2 /= s;
    const double ex0 = ey1 * ez2 - ey2 * ez1, ex1 = ey2 * ez0 - ey0 * ez2, ex2 = ey0 * ez1 - ey1 * ez0;
    const double h = 1.0 / sqrt(2.0);
    rset(R, 0, 0, 0, make_double2(1.0, 0.0));
    rset(R, 1, -1, -1, make_double2((ey1 + ex0) * 0.5, (ex1 - ey0) * 0.5));
    rset(R, 1, -1, 0, make_double2(ez0 * h, ez1 * h));
    rset(R, 1, -1, 1, make_double2((ey1 - ex0) * 0.5, (-ex1 - ey0) * 0.5));
    rset(R, 1, 0, -1, make_double2(ex2 * h, -ey2 * h));
    rset(R, 1, 0, 0, make_double2(ez2, 0.0));
    rset(R, 1, 0, 1, make_double2(-ex2 * h, -ey2 * h));
    rset(R, 1, 1, -1, make_double2((ey1 - ex0) * 0.5, (ex1 + ey0) * 0.5));
    rset(R, 1, 1, 0, make_double2(-ez0 * h, ez1 * h));
    rset(R, 1, 1, 1, make_double2((ey1 + ex0) * 0.5, (ey0 - ex1) * 0.5));
  }
  __syncthreads();
  const double* ca = rc;
  const double* cb = rc + 5 * 13 * 13;
  const double* cc = rc + 2 * 5 * 13 * 13;
  const double* cdd = rc + 2 * 5 * 13 * 13 + 5 * 13;
  const int L6 = 6;
  for (int l = 2; l <= mmax; ++l) {
    const int k = l - 1, lo = l - 2;
    const int nlow = (l + 1) * (l + 1);
    for (int it = lane; it < nlow + l * l; it += kWave) {
      int m, m1;
      if (it < nlow) {
        m = -l + it / (l + 1);
        m1 = it % (l + 1);
      } else {
        const int t = it - nlow;
        m = 1 + t / l;
        m1 = 1 + t % l;
      }
      double re, im;
      if (m1 < l) {
        if (m == -l || m == l) {
          const double bv = cb[(lo * 13 + L6 + (m == -l ? -m : m)) * 13 + L6 + m1];
          const double2 r1 = rget(R, 1, m == -l ? -1 : 1, 0);
          const double2 s1 = rget(R, k, m == -l ? m + 1 : m - 1, m1);
          re = bv * (r1.x * s1.x - r1.y * s1.y);
          im = bv * (r1.x * s1.y + r1.y * s1.x);
        } else {
          const double av = ca[(lo * 13 + L6 + m) * 13 + L6 + m1];
          const double bp = cb[(lo * 13 + L6 + m) * 13 + L6 + m1];
          const double bn = cb[(lo * 13 + L6 - m) * 13 + L6 + m1];
          const double2 r0 = rget(R, 1, 0, 0), rp = rget(R, 1, 1, 0), rn = rget(R, 1, -1, 0);
          const double2 s0 = rget(R, k, m, m1), sp = rget(R, k, m + 1, m1), sn = rget(R, k, m - 1, m1);
          re = av * (r0.x * s0.x) + bp * (rp.x * sn.x - rp.y * sn.y) + bn * (rn.x * sp.x - rn.y * sp.y);
          im = av * (r0.x * s0.y) + bp * (rp.x * sn.y + rp.y * sn.x) + bn * (rn.x * sp.y + rn.y * sp.x);
        }
      } else {
        if (m == -l || m == l) {
          const double dv = cdd[lo * 13 + L6 + (m == -l ? -m : m)];
          const double2 r1 = rget(R, 1, m == -l ? -1 : 1, 1);
          const double2 s1 = rget(R, k, m == -l ? m + 1 : m - 1, m1 - 1);
          re = dv * (r1.x * s1.x - r1.y * s1.y);
          im = dv * (r1.x * s1.y + r1.y * s1.x);
        } else {
          const double cv = cc[lo * 13 + L6 + m];
          const double dp = cdd[lo * 13 + L6 + m];
          const double dn = cdd[lo * 13 + L6 - m];
          const double2 r0 = rget(R, 1, 0, 1), rp = rget(R, 1, 1, 1), rn = rget(R, 1, -1, 1);
          const double2 s0 = rget(R, k, m, m1 - 1), sp = rget(R, k, m + 1, m1 - 1), sn = rget(R, k, m - 1, m1 - 1);
          re = cv * (r0.x * s0.x - r0.y * s0.y) + dp * (rp.x * sn.x - rp.y * sn.y) + dn * (rn.x * sp.x - rn.y * sp.y);
          im = cv * (r0.x * s0.y + r0.y * s0.x) + dp * (rp.x * sn.y + rp.y * sn.x) + dn * (rn.x * sp.y + rn.y * sp.x);
        }
      }
      rset(R, l, m, m1, make_double2(re, im));
      const double sg = ((m + m1) & 1) ? -1.0 : 1.0;
      rset(R, l, -m, -m1, make_double2(sg * re, -sg * im));
    }
    __syncthreads();
  }
}

// e^{i mup phi_q} as (cos, sin) in FP64: e^{i phi} = (qx + i qy)/|q_xy|,
// phi = 0 on the z axis (the designated q of a pair never points along -z:
// its kz >= 0).
__device__ __forceinline__ double2 phase_pow(double q0, double q1, int mup) {
  const double rxy = sqrt(q0 * q0 + q1 * q1);
  const double c1 = rxy > 0.0 ? q0 / rxy : 1.0, s1 = rxy > 0.0 ? q1 / rxy : 0.0;
  double c = 1.0, s = 0.0;
  const int am = mup < 0 ? -mup : mup;
  for (int k = 0; k < am; ++k) {
    const double t = c * c1 - s * s1;
    s = s * c1 + c * s1;
    c = t;
  }
  return make_double2(c, mup < 0 ? -s : s);
}

// Set-up pass (once per Convolution and precision): the real matrices
// d^l_{mup,chi}(theta_q) = Re(R^l_{mup,chi}(q) e^{+i mup phi_q}) of every pair,
// R from the FP64 recursion, rounded to T.  Row `pair` of dtab holds the nR
// compact entries, or with padw > 0 the padded layout of the compile-time OZ
// kernels: rows (l, mup) of padw entries, chi + l in the first 2l+1 and zeros
// after; all (mmax+1)^2 rows, or with half only the (mmax+1)(mmax+2)/2 rows
// of mup >= 0 (the rows of -mup follow from
// d^l_{-mup,-chi} = (-1)^{mup-chi} d^l_{mup,chi}, which the values of the
// FP64 recursion satisfy exactly).  With sel, row b is the table of pair
// sel[4 b] (the first pair of star b), for the star kernel: the pairs of a
// star share theta_q, and d depends on theta_q alone.
template <typename T>
__global__ __launch_bounds__(kWave) void oz_dtable_kernel(T* __restrict__ dtab, OzArgs<T> a, int padw, int half,
                                                          const int* __restrict__ sel, long nrows) {
  extern __shared__ unsigned char lds_raw[];
  double2* Rd = reinterpret_cast<double2*>(lds_raw);
  const int lane = threadIdx.x;
  const long pair = blockIdx.x;
  if (pair >= nrows) return;
  const int pk = a.pairs[sel ? sel[4 * pair] : pair];
  const int ix = pk & 1023, iy = (pk >> 10) & 1023, iz = (pk >> 20) & 1023;
  const double q0 = a.kxyz[ix], q1 = a.kxyz[a.nx + iy], q2 = a.kxyz[a.nx + a.ny + iz];
  rotation_recursion(Rd, a.rc, a.mmax, q0, q1, q2, lane);
  if (padw == 0) {
    const int nR = roff(a.mmax + 1);
    for (int it = lane; it < nR; it += kWave) {
      int l = 0;
      while (roff(l + 1) <= it) ++l;
      const int mup = (it - roff(l)) / (2 * l + 1) - l;
      const double2 e = phase_pow(q0, q1, mup);
      const double2 v = Rd[it];
      dtab[pair * nR + it] = static_cast<T>(v.x * e.x - v.y * e.y);
    }
  } else {
    const int nrow = half ? (a.mmax + 1) * (a.mmax + 2) / 2 : (a.mmax + 1) * (a.mmax + 1);
    for (int pos = lane; pos < nrow * padw; pos += kWave) {
      const int row = pos / padw, col = pos - row * padw;
      int l = 0, mup;
      if (half) {
        while ((l + 1) * (l + 2) / 2 <= row) ++l;
        mup = row - l * (l + 1) / 2;
      } else {
        while ((l + 1) * (l + 1) <= row) ++l;
        mup = row - l * l - l;
      }
      double v = 0.0;
      if (col < 2 * l + 1) {
        const double2 e = phase_pow(q0, q1, mup);
        const double2 r = Rd[roff(l) + (mup + l) * (2 * l + 1) + col];
        v = r.x * e.x - r.y * e.y;
      }
      dtab[pair * nrow * padw + pos] = static_cast<T>(v);
    }
  }
}

// The per-pair LDS layouts of the OZ pass for angular grids with compile-time
// (mmax, mrso) (oz_pairs_ct_kernel), where every LDS access is a lane base
// plus a compile-time offset and every loop is unrolled (W = 2 mmax + 1):
//   G, Gm:  rows (m, mu2) of W entries, entry mup + m: the phased gathers,
//           later gamma with entry chi + m (zero padding past 2m+1);
//   Dq, Dmq: the rotated vectors in projection order;
//   dP:     one leading zero row, then rows (m, mup) of W entries d^m_{mup,chi}
//           at chi + m (the padded dtab rows of the pair).
// Rotation: Dq(m, chi, mu2) = sum_j G[m, mu2][j] dP[m, row j][chi + m], the d
// column read at stride W; the padded G entries are zero, so the rows j >= 2m+1
// (the next m's rows) add nothing.  The OZ product runs over the n blocks
// n >= |chi| with the term offset of block n a compile-time constant past the
// lane's base.  Inverse rotation: out(m, mup, mu2) = sum_j gamma[j] dP[m, mup][j]
// (and dP[m, mup][2m - j] for -q; the leading zero row keeps those reads in
// the array, and gamma's zero padding cancels them).
template <int MMAX, int MRSO>
struct OzShape {
  static constexpr int W = 2 * MMAX + 1;
  static constexpr int km(int m) { return m / MRSO + 1; }
  static constexpr int mbase(int m) {
    int a = 0;
    for (int k = 0; k < m; ++k) a += (2 * k + 1) * km(k);
    return a;
  }
  static constexpr int grow(int m) {  // G rows before m
    int a = 0;
    for (int k = 0; k < m; ++k) a += km(k);
    return a;
  }
  static constexpr int tofs(int n) {  // OZ terms of the n blocks before n
    int a = 0;
    for (int k = 0; k < n; ++k) a += 2 * (k / MRSO) + 1;
    return a;
  }
  static constexpr int NP = mbase(MMAX + 1), NG = grow(MMAX + 1), ND = (MMAX + 1) * (MMAX + 1);
  static constexpr int NSLOT = (NP + kWave - 1) / kWave;
};

// runtime values of the OzShape offsets (m, n <= MMAX: unrolled sums)
template <int MMAX, int MRSO>
__device__ __forceinline__ int grow_rt(int m) {
  int a = 0;
#pragma unroll
  for (int k = 0; k < MMAX; ++k) a += k < m ? k / MRSO + 1 : 0;
  return a;
}
template <int MMAX, int MRSO>
__device__ __forceinline__ int tofs_rt(int n) {
  int a = 0;
#pragma unroll
  for (int k = 0; k < MMAX; ++k) a += k < n ? 2 * (k / MRSO) + 1 : 0;
  return a;
}


// The compile-time OZ kernel with PB pairs per block and one thread per
// Per-pair set-up values of the FP32 compile-time OZ kernel, computed once
// per Convolution by oz_meta_kernel with the expressions of the per-pair
// kernels: the packed (ix, iy, iz), the c(q) row iq and the interpolation
// weights alpha and 1 - alpha rounded to T; and, in a table of its own, the
// phases e^{-i mup phi_q} rounded to T for mup = 0..mmax (those of -mup are
// their conjugates: phase_pow(-mup) is the conjugate of phase_pow(mup), and
// rounding commutes with negation).
template <typename T>
struct alignas(16) OzPairMeta {
  int pk, iq;
  T al, bl;
};

template <typename T>
__global__ __launch_bounds__(256) void oz_meta_kernel(OzPairMeta<T>* __restrict__ meta, cplx<T>* __restrict__ phase,
                                                      OzArgs<T> a) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.npairs) return;
  const int pk = a.pairs[i];
  const int ix = pk & 1023, iy = (pk >> 10) & 1023, iz = (pk >> 20) & 1023;
  const double q0 = a.kxyz[ix], q1 = a.kxyz[a.nx + iy], q2 = a.kxyz[a.nx + a.ny + iz];
  const double qn = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
  const double eff = qn / a.dq;
  int iq = static_cast<int>(eff);
  const double alpha = eff - static_cast<double>(iq);
  if (iq > a.nq - 2) iq = a.nq - 2;  // unreachable for |q| <= qmax (nq has 1.5 rows of margin)
  OzPairMeta<T> m;
  m.pk = pk;
  m.iq = iq;
  m.al = static_cast<T>(alpha);
  m.bl = static_cast<T>(1.0 - alpha);
  meta[i] = m;
  for (int mup = 0; mup <= a.mmax; ++mup) {
    const double2 e = phase_pow(q0, q1, mup);
    phase[i * (a.mmax + 1) + mup] = mkc<T>(static_cast<T>(e.x), static_cast<T>(-e.y));
  }
}

// OZ terms of all projections of a compile-time grid (np_new): output
// (m, chi, mu2) sums the (2 (n / mrso) + 1) terms of every n >= |chi|.
template <int MMAX, int MRSO>
constexpr int oz_terms() {
  using S = OzShape<MMAX, MRSO>;
  int t = 0;
  for (int m = 0; m <= MMAX; ++m)
    for (int chi = -m; chi <= m; ++chi) t += (m / MRSO + 1) * (S::tofs(MMAX + 1) - S::tofs(chi < 0 ? -chi : chi));
  return t;
}

// pairs per OZ block: two when their outputs fit 256 lanes (FP32); one in
// FP64, whose pair holds twice the LDS
template <typename T, int NP>
constexpr int oz_pairs_per_block() {
  return std::is_same<T, float>::value && 2 * NP <= 256 ? 2 : 1;
}

// Where the half d table's entries go in LDS: entry e = ip + k NP of a
// pair's record (row h = (l, mu >= 0) with h = l (l + 1) / 2 + mu, column c)
// is read by thread ip into register k and written to dP[(l^2 + l + mu) W + c]
// (bits 0-11) and to its mirror dP[(l^2 + l - mu) W + 2l - c] with the sign
// (-1)^{mu + c - l} (bits 12-23, kDmapNeg), or, for a padding column c > 2l,
// a zero at dP[(l^2 + l - mu) W + c] (kDmapZero).  Built once per
// Convolution as one 16-byte record per thread ip: the codes of k = 0..2
// (at most three entries per thread for mmax <= 6), then the projection's
// m | (mup + mmax) << 4 | mu2 << 8 | ia_start << 12 (its first OZ term).
constexpr int kDmapNeg = 1 << 24, kDmapZero = 1 << 25, kDmapValid = 1 << 26;
constexpr int kDmapPerThread = 3;

// the LDS places code of half-table entry e (layout above)
int oz_dtable_place(int mmax, int e) {
  const int W = 2 * mmax + 1;
  const int h = e / W, c = e % W;
  int l = 0;
  while ((l + 1) * (l + 2) / 2 <= h) ++l;
  const int mu = h - l * (l + 1) / 2;
  const int base = (l * l + l) * W;
  int v = (base + mu * W + c) | kDmapValid;
  if (c <= 2 * l) {
    v |= (base - mu * W + 2 * l - c) << 12;
    if ((mu + c + l) & 1) v |= kDmapNeg;
  } else {
    v |= (base - mu * W + c) << 12;
    v |= kDmapZero;
  }
  return v;
}

std::vector<int> oz_thread_records(const AngularGrid& g, const std::vector<int>& ia_start) {
  const int mmax = g.mmax, np = g.np;
  const int W = 2 * mmax + 1, NE = (mmax + 1) * (mmax + 2) / 2 * W, nu = (NE + np - 1) / np;
  throw_if(nu > kDmapPerThread || 2 * mmax >= 16, "internal error: OZ thread records");
  std::vector<int> code(static_cast<size_t>(np) * 4, 0);
  for (int ip = 0; ip < np; ++ip) {
    throw_if(ia_start[ip] >= (1 << 19), "internal error: OZ term index of the thread records");
    code[static_cast<size_t>(ip) * 4 + 3] =
        g.m[ip] | ((g.mup[ip] + mmax) << 4) | (g.mu2[ip] << 8) | (ia_start[ip] << 12);
  }
  for (int k = 0; k < nu; ++k)
    for (int ip = 0; ip < np; ++ip) {
      const int e = ip + k * np;
      if (e < NE) code[static_cast<size_t>(ip) * 4 + k] = oz_dtable_place(mmax, e);
    }
  return code;
}

// the LDS places codes of all NE half-table entries (the star kernel writes
// the star's one table with one entry per thread)
std::vector<int> oz_dtable_places(int mmax) {
  const int NE = (mmax + 1) * (mmax + 2) / 2 * (2 * mmax + 1);
  std::vector<int> code(NE);
  for (int e = 0; e < NE; ++e) code[e] = oz_dtable_place(mmax, e);
  return code;
}

// LDS of the star kernel for (T, MMAX, MRSO): the lerped c(q) row, the
// star's d table (a leading zero row, then (mmax + 1)^2 rows), then four pair
// regions (G, Gm, Dq, Dmq), each rounded up to 16 bytes
template <typename T, int MMAX, int MRSO>
struct StarLayout {
  using S = OzShape<MMAX, MRSO>;
  static constexpr int NC = oz_terms<MMAX, MRSO>();
  static constexpr int threads = (4 * S::NP + kWave - 1) / kWave * kWave;
  static constexpr size_t r16(size_t b) { return (b + 15) / 16 * 16; }
  static constexpr size_t cbytes = r16(static_cast<size_t>(NC) * sizeof(cplx<T>));
  static constexpr size_t dbytes = r16(static_cast<size_t>(S::W) * (S::ND + 1) * sizeof(T));
  static constexpr size_t region = r16((2 * static_cast<size_t>(S::NG) * S::W + 2 * S::NP) * sizeof(cplx<T>));
  static constexpr size_t dynamic = cbytes + dbytes + 4 * region;
  static constexpr bool fits = threads <= 1024 && dynamic + 4 * S::W * sizeof(cplx<T>) <= 64 * 1024;
};

// element idx of a table at a 32-bit byte offset from a block-uniform base
// (the load then takes a scalar base and a 32-bit vector offset)
template <typename V>
__device__ __forceinline__ V ld_off(const V* base, unsigned idx) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + idx * static_cast<unsigned>(sizeof(V)));
}

// The OZ pass for compile-time grids in FP32: PB (q, -q) pairs per block, NP
// threads per pair, one (pair, projection) output per thread (176 of 192
// lanes at mmax = 5), each pair with its own LDS region (`region` bytes).
// Blocks take groups of PB consecutive pairs in memory order, XCD-mapped like
// oz_block_pair (a.chunk counts groups).  Every input of the block is
// requested at once, right after the pair's set-up record: the q and -q
// gathers of every projection, the pair's half d table (rows mup >= 0), the
// two c(q) rows (np_new entries each, a compile-time count) and the phases.
// The mirrored d rows (d^l_{-mup,-chi} = (-1)^{mup-chi} d^l_{mup,chi}), the
// phases and the lerped c row then go to LDS.  FP32 grids up to 128
// projections take two pairs per block; FP64 takes one (its pair's LDS region
// is twice as large, and two pairs per block measured 6% slower at 64^3).
template <typename T, int MMAX, int MRSO, int PB>
__global__ __launch_bounds__(256) void oz_pairs_ct_kernel(cplx<T>* __restrict__ p, const T* __restrict__ dtab,
                                                          const OzPairMeta<T>* __restrict__ meta,
                                                          const cplx<T>* __restrict__ phase,
                                                          const int* __restrict__ dmap, OzArgs<T> a, int region) {
  using S = OzShape<MMAX, MRSO>;
  constexpr int W = S::W, NP = S::NP, NC = oz_terms<MMAX, MRSO>();
  constexpr int NH = (MMAX + 1) * (MMAX + 2) / 2, NE = NH * W;
  constexpr int NU = (NE + NP - 1) / NP, NCU = (NC + NP - 1) / NP;
  static_assert(NU <= kDmapPerThread, "d entries per thread of the thread records");
  extern __shared__ unsigned char lds_raw[];
  __shared__ cplx<T> phs[PB][W];  // e^{-i mup phi_q} of each pair
  const int tid = threadIdx.x;
  const int pp = tid / NP;  // pair of this thread within the block
  const int ip = tid - pp * NP;
  const int ppc = pp < PB ? pp : 0;
  int job;
  const long grp = oz_block_pair(a, job);
  const long pair = grp * PB + pp;
  const bool active = pp < PB && pair < a.npairs;  // uniform over the threads of one pair
  cplx<T>* G = reinterpret_cast<cplx<T>*>(lds_raw + static_cast<long>(ppc) * region);
  cplx<T>* Gm = G + S::NG * W;
  cplx<T>* Dq = Gm + S::NG * W;
  cplx<T>* Dmq = Dq + NP;
  cplx<T>* ceff = Dmq + NP;
  T* dP0 = reinterpret_cast<T*>(ceff + NC);  // leading zero row, then ND rows
  T* dP = dP0 + W;
  cplx<T>* ph = phs[ppc];
  p += static_cast<long>(job) * a.N;

  int info = 0, ias = 0, pk = 0;
  T g0x = 0, g0y = 0, g1x = 0, g1y = 0;
  long sq = 0, smq = 0;
  bool singular = false;
  if (active) {
    // inputs that need only the pair index: the half d table and the phases
    const T* dgrp = dtab + grp * (PB * NE);  // block-uniform
    T dreg[NU];
#pragma unroll
    for (int k = 0; k < NU; ++k) {
      const int e = ip + k * NP;
      dreg[k] = ld_off(dgrp, static_cast<unsigned>(pp * NE + (k + 1 < NU ? e : min(e, NE - 1))));
    }
    // this thread's LDS places of its d entries and its projection's info
    const int4 rec = reinterpret_cast<const int4*>(dmap)[ip];
    const int dcode[kDmapPerThread] = {rec.x, rec.y, rec.z};
    cplx<T> phv = mkc<T>(0, 0);
    const int amup = ip < MMAX ? MMAX - ip : ip - MMAX;
    if (ip < W) phv = phase[pair * (MMAX + 1) + amup];
    const OzPairMeta<T> mt = meta[pair];
    pk = mt.pk;
    const int ix = pk & 1023, iy = (pk >> 10) & 1023, iz = (pk >> 20) & 1023;
    const int ixm = ix == 0 ? 0 : a.nx - ix, iym = iy == 0 ? 0 : a.ny - iy, izm = iz == 0 ? 0 : a.nz - iz;
    sq = (static_cast<long>(ix) * a.ny + iy) * a.nz + iz;
    smq = (static_cast<long>(ixm) * a.ny + iym) * a.nz + izm;
    singular = (ixm == ix && iym == iy && izm == iz) && (ix == a.nx / 2 || iy == a.ny / 2 || iz == a.nz / 2);
    // gathers (:420-425), the two c rows and this thread's metadata
    const cplx<T> u = p[static_cast<long>(ip) * a.pstride + sq], v = p[static_cast<long>(ip) * a.pstride + smq];
    // the two c rows: with an even term count, two adjacent terms per lane
    // and load (16-byte accesses: the rows start at multiples of 16 bytes, and
    // so does ceff in LDS, after 2 (NG W + NP) complex values)
    constexpr bool C2 = std::is_same<T, float>::value && NC % 2 == 0;
    constexpr int NC2 = NC / 2, NCU2 = (NC2 + NP - 1) / NP;
    cplx<T> lo[C2 ? 1 : NCU], hi[C2 ? 1 : NCU];
    float4 lo4[C2 ? NCU2 : 1], hi4[C2 ? NCU2 : 1];
    if constexpr (C2) {
      const float4* c4 = reinterpret_cast<const float4*>(a.ctab);
      const unsigned crow2 = static_cast<unsigned>(mt.iq) * NC2;
#pragma unroll
      for (int k = 0; k < NCU2; ++k) {
        const int j = ip + k * NP;
        const unsigned at = crow2 + static_cast<unsigned>(k + 1 < NCU2 ? j : min(j, NC2 - 1));
        lo4[k] = ld_off(c4, at);
        hi4[k] = ld_off(c4, at + NC2);
      }
    } else {
      const unsigned crow = static_cast<unsigned>(mt.iq) * NC;
#pragma unroll
      for (int k = 0; k < NCU; ++k) {
        const int ia = ip + k * NP;
        const unsigned at = crow + static_cast<unsigned>(k + 1 < NCU ? ia : min(ia, NC - 1));
        lo[k] = ld_off(a.ctab, at);
        hi[k] = ld_off(a.ctab, at + NC);
      }
    }
    info = rec.w;
    g0x = u.x;
    g0y = u.y;
    g1x = v.x;
    g1y = v.y;
    // d rows (l, mu >= 0) and their mirrors (l, -mu), at the LDS places
    // the thread's map entries give; padding stays zero
#pragma unroll
    for (int k = 0; k < NU; ++k) {
      const int code = dcode[k];
      if (code & kDmapValid) {
        const T dv = dreg[k];
        dP[code & 0xfff] = dv;
        dP[(code >> 12) & 0xfff] = (code & kDmapZero) ? T(0) : ((code & kDmapNeg) ? -dv : dv);
      }
    }
    if (ip < W) {
      dP0[ip] = T(0);
      ph[ip] = ip < MMAX ? cconj<T>(phv) : phv;
    }
    // zero G and Gm (the padding past 2m+1 must stay zero)
    for (int i = ip; i < 2 * S::NG * W; i += NP) G[i] = mkc<T>(0, 0);
    // c(|q|) rows iq, iq+1 lerped into LDS
    const T al = mt.al, bl = mt.bl;
    if constexpr (C2) {
#pragma unroll
      for (int k = 0; k < NCU2; ++k) {
        const int j = ip + k * NP;
        if (j < NC2) {
          float4 v;
          v.x = al * hi4[k].x + bl * lo4[k].x;
          v.y = al * hi4[k].y + bl * lo4[k].y;
          v.z = al * hi4[k].z + bl * lo4[k].z;
          v.w = al * hi4[k].w + bl * lo4[k].w;
          reinterpret_cast<float4*>(ceff)[j] = v;
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < NCU; ++k) {
        const int ia = ip + k * NP;
        if (ia < NC) ceff[ia] = mkc<T>(al * hi[k].x + bl * lo[k].x, al * hi[k].y + bl * lo[k].y);
      }
    }
  }
  __syncthreads();

  const int m = info & 0xf, mupm = (info >> 4) & 0xf, mu2 = (info >> 8) & 0xf;
  ias = info >> 12;
  const int chi = mupm - MMAX;  // the same field read as chi in the q frame
  const int row = grow_rt<MMAX, MRSO>(m) + mu2;
  // phased gathers g' = g e^{-i mup phi} into G[(m, mu2)][mup + m]
  if (active) {
    const cplx<T> e = ph[mupm];
    const int at = row * W + mupm - MMAX + m;
    G[at] = cmul<T>(mkc<T>(g0x, g0y), e);
    Gm[at] = cmul<T>(mkc<T>(g1x, g1y), e);
  }
  __syncthreads();

  // rotation to the q frame (:427-443)
  if (active) {
    const cplx<T>* g0 = G + row * W;
    const cplx<T>* g1 = Gm + row * W;
    const T* d0 = dP + m * m * W + (chi + m);
    const T* d1 = dP + m * m * W + (m - chi);
    T ar = 0, ai = 0, br = 0, bi = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const cplx<T> x = g0[j], y = g1[j];
      const T u = d0[j * W], v = d1[j * W];
      ar = fma(x.x, u, ar);
      ai = fma(x.y, u, ai);
      br = fma(y.x, v, br);
      bi = fma(y.y, v, bi);
    }
    Dq[ip] = mkc<T>(ar, ai);
    Dmq[ip] = (m & 1) ? mkc<T>(-br, -bi) : mkc<T>(br, bi);
  }
  __syncthreads();

  // OZ product (:468-496) and chop (:514-517)
  if (active) {
    const int achi = chi < 0 ? -chi : chi;
    const cplx<T>* ce = ceff + (ias - tofs_rt<MMAX, MRSO>(achi));
    cplx<T> s0 = mkc<T>(0, 0), s1 = mkc<T>(0, 0);
#pragma unroll
    for (int n = 0; n <= MMAX; ++n) {
      if (n >= achi) {
        const int h = n / MRSO, kn = n / MRSO + 1;
        const cplx<T>* cn = ce + S::tofs(n);
        const int srcb = S::mbase(n) + (chi + n) * kn;
#pragma unroll
        for (int k = 0; k <= 2 * h; ++k) {
          const int nu2 = k - h;
          const cplx<T> c = cn[k];
          if (nu2 < 0) {
            const int src = srcb - nu2;
            s0 = cfma<T>(s0, c, Dq[src]);
            s1 = cfma<T>(s1, c, Dmq[src]);
          } else {
            const int src = srcb + nu2;
            s0 = cfma_conjb<T>(s0, c, Dmq[src]);
            s1 = cfma_conjb<T>(s1, c, Dq[src]);
          }
        }
      }
    }
    const int at = row * W + chi + m;
    G[at] = chop<T>(s0, a.chop_eps);
    Gm[at] = chop<T>(s1, a.chop_eps);
  }
  __syncthreads();

  // inverse rotation and stores (:518-541)
  if (active) {
    const int mup = mupm - MMAX;
    const cplx<T>* g0 = G + row * W;
    const cplx<T>* g1 = Gm + row * W;
    const T* dr = dP + (m * m + mup + m) * W;
    const T* drm = dr + 2 * m;  // d_{mup,-chi} = drm[-j]; reads past the row start meet gamma's zeros
    T ar = 0, ai = 0, br = 0, bi = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const cplx<T> x = g0[j], y = g1[j];
      const T u = dr[j], v = drm[-j];
      ar = fma(x.x, u, ar);
      ai = fma(x.y, u, ai);
      br = fma(y.x, v, br);
      bi = fma(y.y, v, bi);
    }
    const cplx<T> e = ph[mupm];  // e^{-i mup phi}; the output takes its conjugate
    const cplx<T> o0 = cmul<T>(mkc<T>(ar, ai), cconj<T>(e));
    cplx<T> o1 = cmul<T>(mkc<T>(br, bi), cconj<T>(e));
    if (m & 1) o1 = mkc<T>(-o1.x, -o1.y);
    p[static_cast<long>(ip) * a.pstride + sq] = o0;
    p[static_cast<long>(ip) * a.pstride + smq] = singular ? cconj<T>(o1) : o1;
  }
}

// The OZ pass over stars, for compile-time grids.  A star is the set of
// (q, -q) pairs whose q differ only in the signs of their components: the
// pairs of (+-kx, +-ky, +-kz), at most four.  The k grid holds -k as the exact
// negation of k, so every pair of a star has the same |q| bit for bit, hence
// the same c(q) row iq, the same interpolation weights and the same lerped
// row.  One block takes one star: the lerped row goes to LDS once, from loads
// spread over all the block's threads, and serves every pair of the star
// (four pairs per c(q) row instead of one).  The designated q of every pair
// of a star has the same kz >= 0 and |q|, hence the same theta_q, and the real
// rotation matrices d^l(theta_q) depend on theta_q alone: the star has one
// half d table (built from its first pair), loaded with one entry per thread
// and mirrored into one LDS table that all its pairs read.  Each pair keeps
// its own LDS region for G, Gm, Dq, Dmq and the per-pair arithmetic of
// oz_pairs_ct_kernel (gathers, phases, rotation, OZ product, chop, inverse
// rotation, stores).  Stars are ordered by (|ix|, |iy|, |iz|): the
// pairs of consecutive stars lie on the same z lines at neighbouring +-iz, and
// a line's cache lines serve the blocks of consecutive stars.  Blocks past the
// last star (the XCD order pads the grid) request nothing.
template <typename T, int MMAX, int MRSO, int THREADS, bool CP>
__global__ __launch_bounds__(THREADS) void oz_star_ct_kernel(cplx<T>* __restrict__ p, const T* __restrict__ dtab,
                                                             const OzPairMeta<T>* __restrict__ meta,
                                                             const cplx<T>* __restrict__ phase,
                                                             const int* __restrict__ dmap,
                                                             const int* __restrict__ dent,
                                                             const int4* __restrict__ stars, long nstars,
                                                             OzArgs<T> a, const int* __restrict__ ccode) {
  using S = OzShape<MMAX, MRSO>;
  using Lay = StarLayout<T, MMAX, MRSO>;
  constexpr int W = S::W, NP = S::NP, NC = oz_terms<MMAX, MRSO>();
  constexpr int NH = (MMAX + 1) * (MMAX + 2) / 2, NE = NH * W;
  constexpr int NDU = (NE + THREADS - 1) / THREADS, NCU = (NC + THREADS - 1) / THREADS;
  constexpr int PS = 4;  // pairs of a star
  static_assert(PS * NP <= THREADS && Lay::threads == THREADS, "one thread per (pair, projection)");
  extern __shared__ unsigned char lds_raw[];
  __shared__ cplx<T> phs[PS][W];  // e^{-i mup phi_q} of each pair
  const int tid = threadIdx.x;
  const int pp = tid / NP;  // pair of this thread within the star
  const int ip = tid - pp * NP;
  const int ppc = pp < PS ? pp : 0;
  int job;
  const long star = oz_block_pair(a, job);
  const bool sval = star < nstars;  // block-uniform
  int4 st = make_int4(-1, -1, -1, -1);
  if (sval) st = stars[star];
  const int pair = ppc == 0 ? st.x : ppc == 1 ? st.y : ppc == 2 ? st.z : st.w;
  const bool active = sval && pp < PS && pair >= 0;  // uniform over the threads of one pair
  cplx<T>* ceff = reinterpret_cast<cplx<T>*>(lds_raw);
  T* dP0 = reinterpret_cast<T*>(lds_raw + Lay::cbytes);  // leading zero row, then ND rows
  T* dP = dP0 + W;
  cplx<T>* G = reinterpret_cast<cplx<T>*>(lds_raw + Lay::cbytes + Lay::dbytes + ppc * Lay::region);
  cplx<T>* Gm = G + S::NG * W;
  cplx<T>* Dq = Gm + S::NG * W;
  cplx<T>* Dmq = Dq + NP;
  cplx<T>* ph = phs[ppc];
  p += static_cast<long>(job) * a.N;

  // the star's c(q) rows iq and iq + 1, over all threads of the block (the
  // first pair of a star is always present)
  constexpr bool C2 = std::is_same<T, float>::value && NC % 2 == 0;
  constexpr int NC2 = NC / 2, NCU2 = (NC2 + THREADS - 1) / THREADS;
  cplx<T> lo[C2 ? 1 : NCU], hi[C2 ? 1 : NCU];
  float4 lo4[C2 ? NCU2 : 1], hi4[C2 ? NCU2 : 1];
  T al = 0, bl = 0;
  T dreg[NDU];
  int dcode[NDU];
  if (sval) {
    // the star's half d table, one entry per thread
    const T* dstar = dtab + star * NE;  // block-uniform
#pragma unroll
    for (int k = 0; k < NDU; ++k) {
      const int e = min(tid + k * THREADS, NE - 1);
      dreg[k] = ld_off(dstar, static_cast<unsigned>(e));
      dcode[k] = tid + k * THREADS < NE ? dent[e] : 0;
    }
    const OzPairMeta<T> m0 = meta[st.x];  // block-uniform
    al = m0.al;
    bl = m0.bl;
    if constexpr (C2) {
      const float4* c4 = reinterpret_cast<const float4*>(a.ctab);
      const unsigned crow2 = static_cast<unsigned>(m0.iq) * NC2;
#pragma unroll
      for (int k = 0; k < NCU2; ++k) {
        const int j = tid + k * THREADS;
        const unsigned at = crow2 + static_cast<unsigned>(min(j, NC2 - 1));
        lo4[k] = ld_off(c4, at);
        hi4[k] = ld_off(c4, at + NC2);
      }
    } else {
      const unsigned crow = static_cast<unsigned>(m0.iq) * NC;
#pragma unroll
      for (int k = 0; k < NCU; ++k) {
        const int ia = tid + k * THREADS;
        const unsigned at = crow + static_cast<unsigned>(min(ia, NC - 1));
        lo[k] = ld_off(a.ctab, at);
        hi[k] = ld_off(a.ctab, at + NC);
      }
    }
  }

  int info = 0, ias = 0, pk = 0;
  T g0x = 0, g0y = 0, g1x = 0, g1y = 0;
  long sq = 0, smq = 0;
  bool singular = false;
  // the projection's channel in p; in the compact order an implied
  // projection reads its partner's channel at the mirrored point, conjugated
  // and times (-1)^mup, and hands its outputs to the partner, which stores
  // g_+(q) + (-1)^mup conj(g_-(-q)) (and the same at -q)
  long chan = ip;
  bool implied = false;
  if (active) {
    // this pair's phase and its projection's info
    const int rec = reinterpret_cast<const int4*>(dmap)[ip].w;
    cplx<T> phv = mkc<T>(0, 0);
    const int amup = ip < MMAX ? MMAX - ip : ip - MMAX;
    if (ip < W) phv = phase[static_cast<long>(pair) * (MMAX + 1) + amup];
    pk = meta[pair].pk;
    const int ix = pk & 1023, iy = (pk >> 10) & 1023, iz = (pk >> 20) & 1023;
    const int ixm = ix == 0 ? 0 : a.nx - ix, iym = iy == 0 ? 0 : a.ny - iy, izm = iz == 0 ? 0 : a.nz - iz;
    sq = (static_cast<long>(ix) * a.ny + iy) * a.nz + iz;
    smq = (static_cast<long>(ixm) * a.ny + iym) * a.nz + izm;
    singular = (ixm == ix && iym == iy && izm == iz) && (ix == a.nx / 2 || iy == a.ny / 2 || iz == a.nz / 2);
    // gathers (:420-425)
    bool odd = false;
    if constexpr (CP) {
      const int cc = ccode[ip];
      chan = cc & 0xff;
      implied = (cc & kCompactImplied) != 0;
      odd = (cc & kCompactOdd) != 0;
    }
    const long c0 = chan * a.pstride;
    const cplx<T> u = p[c0 + (implied ? smq : sq)], v = p[c0 + (implied ? sq : smq)];
    info = rec;
    g0x = u.x;
    g0y = u.y;
    g1x = v.x;
    g1y = v.y;
    if (implied) {
      g0x = odd ? -g0x : g0x;
      g0y = odd ? g0y : -g0y;
      g1x = odd ? -g1x : g1x;
      g1y = odd ? g1y : -g1y;
    }
    if (ip < W) ph[ip] = ip < MMAX ? cconj<T>(phv) : phv;
    // zero G and Gm (the padding past 2m+1 must stay zero)
    for (int i = ip; i < 2 * S::NG * W; i += NP) G[i] = mkc<T>(0, 0);
  }
  if (sval) {
    // d rows (l, mu >= 0) and their mirrors (l, -mu), at the LDS places the
    // entry codes give; padding stays zero
#pragma unroll
    for (int k = 0; k < NDU; ++k) {
      const int code = dcode[k];
      if (code & kDmapValid) {
        const T dv = dreg[k];
        dP[code & 0xfff] = dv;
        dP[(code >> 12) & 0xfff] = (code & kDmapZero) ? T(0) : ((code & kDmapNeg) ? -dv : dv);
      }
    }
    if (tid < W) dP0[tid] = T(0);
    // c(|q|) rows iq, iq+1 lerped into LDS (the expression of oz_pairs_ct_kernel)
    if constexpr (C2) {
#pragma unroll
      for (int k = 0; k < NCU2; ++k) {
        const int j = tid + k * THREADS;
        if (j < NC2) {
          float4 v;
          v.x = al * hi4[k].x + bl * lo4[k].x;
          v.y = al * hi4[k].y + bl * lo4[k].y;
          v.z = al * hi4[k].z + bl * lo4[k].z;
          v.w = al * hi4[k].w + bl * lo4[k].w;
          reinterpret_cast<float4*>(ceff)[j] = v;
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < NCU; ++k) {
        const int ia = tid + k * THREADS;
        if (ia < NC) ceff[ia] = mkc<T>(al * hi[k].x + bl * lo[k].x, al * hi[k].y + bl * lo[k].y);
      }
    }
  }
  __syncthreads();

  const int m = info & 0xf, mupm = (info >> 4) & 0xf, mu2 = (info >> 8) & 0xf;
  ias = info >> 12;
  const int chi = mupm - MMAX;  // the same field read as chi in the q frame
  const int row = grow_rt<MMAX, MRSO>(m) + mu2;
  // phased gathers g' = g e^{-i mup phi} into G[(m, mu2)][mup + m]
  if (active) {
    const cplx<T> e = ph[mupm];
    const int at = row * W + mupm - MMAX + m;
    G[at] = cmul<T>(mkc<T>(g0x, g0y), e);
    Gm[at] = cmul<T>(mkc<T>(g1x, g1y), e);
  }
  __syncthreads();

  // rotation to the q frame (:427-443)
  if (active) {
    const cplx<T>* g0 = G + row * W;
    const cplx<T>* g1 = Gm + row * W;
    const T* d0 = dP + m * m * W + (chi + m);
    const T* d1 = dP + m * m * W + (m - chi);
    T ar = 0, ai = 0, br = 0, bi = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const cplx<T> x = g0[j], y = g1[j];
      const T u = d0[j * W], v = d1[j * W];
      ar = fma(x.x, u, ar);
      ai = fma(x.y, u, ai);
      br = fma(y.x, v, br);
      bi = fma(y.y, v, bi);
    }
    Dq[ip] = mkc<T>(ar, ai);
    Dmq[ip] = (m & 1) ? mkc<T>(-br, -bi) : mkc<T>(br, bi);
  }
  __syncthreads();

  // OZ product (:468-496) and chop (:514-517)
  if (active) {
    const int achi = chi < 0 ? -chi : chi;
    const cplx<T>* ce = ceff + (ias - tofs_rt<MMAX, MRSO>(achi));
    cplx<T> s0 = mkc<T>(0, 0), s1 = mkc<T>(0, 0);
#pragma unroll
    for (int n = 0; n <= MMAX; ++n) {
      if (n >= achi) {
        const int h = n / MRSO, kn = n / MRSO + 1;
        const cplx<T>* cn = ce + S::tofs(n);
        const int srcb = S::mbase(n) + (chi + n) * kn;
#pragma unroll
        for (int k = 0; k <= 2 * h; ++k) {
          const int nu2 = k - h;
          const cplx<T> c = cn[k];
          if (nu2 < 0) {
            const int src = srcb - nu2;
            s0 = cfma<T>(s0, c, Dq[src]);
            s1 = cfma<T>(s1, c, Dmq[src]);
          } else {
            const int src = srcb + nu2;
            s0 = cfma_conjb<T>(s0, c, Dmq[src]);
            s1 = cfma_conjb<T>(s1, c, Dq[src]);
          }
        }
      }
    }
    const int at = row * W + chi + m;
    G[at] = chop<T>(s0, a.chop_eps);
    Gm[at] = chop<T>(s1, a.chop_eps);
  }
  __syncthreads();

  // inverse rotation and stores (:518-541)
  cplx<T> f0 = mkc<T>(0, 0), f1 = mkc<T>(0, 0);
  if (active) {
    const int mup = mupm - MMAX;
    const cplx<T>* g0 = G + row * W;
    const cplx<T>* g1 = Gm + row * W;
    const T* dr = dP + (m * m + mup + m) * W;
    const T* drm = dr + 2 * m;  // d_{mup,-chi} = drm[-j]; reads past the row start meet gamma's zeros
    T ar = 0, ai = 0, br = 0, bi = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const cplx<T> x = g0[j], y = g1[j];
      const T u = dr[j], v = drm[-j];
      ar = fma(x.x, u, ar);
      ai = fma(x.y, u, ai);
      br = fma(y.x, v, br);
      bi = fma(y.y, v, bi);
    }
    const cplx<T> e = ph[mupm];  // e^{-i mup phi}; the output takes its conjugate
    const cplx<T> o0 = cmul<T>(mkc<T>(ar, ai), cconj<T>(e));
    cplx<T> o1 = cmul<T>(mkc<T>(br, bi), cconj<T>(e));
    if (m & 1) o1 = mkc<T>(-o1.x, -o1.y);
    // the values the full order leaves at q and -q (one address, the second
    // store, when the point is its own mirror: q = 0 and the singular points)
    f1 = singular ? cconj<T>(o1) : o1;
    f0 = sq == smq ? f1 : o0;
    if (!CP) {
      p[chan * a.pstride + sq] = o0;
      p[chan * a.pstride + smq] = singular ? cconj<T>(o1) : o1;
    } else if (implied) {
      Dq[ip] = f0;  // free since the OZ product
      Dmq[ip] = f1;
    }
  }
  if constexpr (CP) {
    __syncthreads();
    if (active && !implied) {
      const int mup = mupm - MMAX;
      if (mu2 == 0 && mup > 0) {
        // c = g_+ + (-1)^mup conj(g_- at the mirrored point)
        const int ipr = ip - 2 * mup * (m / MRSO + 1);  // the implied (m, -mup, 0)
        const cplx<T> h0 = Dq[ipr], h1 = Dmq[ipr];
        const T sg = (mup & 1) ? T(-1) : T(1);
        f0 = mkc<T>(f0.x + sg * h1.x, f0.y - sg * h1.y);
        f1 = mkc<T>(f1.x + sg * h0.x, f1.y - sg * h0.y);
      }
      p[chan * a.pstride + sq] = f0;
      p[chan * a.pstride + smq] = f1;
    }
  }
}

// One pair per 64-lane block (per-evaluation pass).
template <typename T, bool STAGE_C>
__global__ __launch_bounds__(kWave) void oz_pair_kernel(cplx<T>* __restrict__ p, const T* __restrict__ dtab,
                                                        OzArgs<T> a) {
  extern __shared__ unsigned char lds_raw[];
  const int mmax = a.mmax, np = a.np;
  const int nR = roff(mmax + 1);
  cplx<T>* gq = reinterpret_cast<cplx<T>*>(lds_raw);
  cplx<T>* gmq = gq + np;
  cplx<T>* dq = gmq + np;
  cplx<T>* dmq = dq + np;
  T* dm = reinterpret_cast<T*>(dmq + np);  // nR real d entries
  cplx<T>* ceff = reinterpret_cast<cplx<T>*>(lds_raw + 4 * np * sizeof(cplx<T>) + ((nR * sizeof(T) + 15) / 16) * 16);
  __shared__ cplx<T> ph[2 * kMaxM + 1];  // e^{-i mup phi_q}, mup = -mmax..mmax
  __shared__ int mbase[kMaxM + 2];       // first ip of each m
  const int lane = threadIdx.x;

  int job;
  const long pair = oz_block_pair(a, job);
  if (pair >= a.npairs) return;  // uniform across the block
  p += static_cast<long>(job) * a.N;
  const int pk = a.pairs[pair];
  const int ix = pk & 1023, iy = (pk >> 10) & 1023, iz = (pk >> 20) & 1023;
  const int ixm = ix == 0 ? 0 : a.nx - ix, iym = iy == 0 ? 0 : a.ny - iy, izm = iz == 0 ? 0 : a.nz - iz;
  const long sq = (static_cast<long>(ix) * a.ny + iy) * a.nz + iz;
  const long smq = (static_cast<long>(ixm) * a.ny + iym) * a.nz + izm;
  const double q0 = a.kxyz[ix], q1 = a.kxyz[a.nx + iy], q2 = a.kxyz[a.nx + a.ny + iz];

  // gathers of g(q), g(-q) (:420-425) and this lane's projection metadata,
  // issued first and together
  // (unconditional loads at clamped indices; plain scalars: arrays of HIP's
  // vector classes are kept in scratch memory)
  T g0x[4], g0y[4], g1x[4], g1y[4];
  int info[4], ias[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int ip = min(lane + t * kWave, np - 1);
    const cplx<T> u = p[static_cast<long>(ip) * a.pstride + sq], v = p[static_cast<long>(ip) * a.pstride + smq];
    g0x[t] = u.x;
    g0y[t] = u.y;
    g1x[t] = v.x;
    g1y[t] = v.y;
    info[t] = a.ipinfo[ip];
    ias[t] = a.ia_start[ip];
  }
  // this pair's d matrices (coalesced, all loads in flight at once) and phases
  {
    const T* dsrc = dtab + pair * nR;
    T dreg[kMaxR / kWave];
#pragma unroll
    for (int u = 0; u < kMaxR / kWave; ++u) dreg[u] = dsrc[min(lane + u * kWave, nR - 1)];
#pragma unroll
    for (int u = 0; u < kMaxR / kWave; ++u) {
      const int it = lane + u * kWave;
      if (it < nR) dm[it] = dreg[u];
    }
  }
  if (lane <= 2 * mmax) {
    const double2 e = phase_pow(q0, q1, lane - mmax);
    ph[lane] = mkc<T>(static_cast<T>(e.x), static_cast<T>(-e.y));
  }
  if (lane == 0) {
    int acc = 0;
    for (int m = 0; m <= mmax; ++m) {
      mbase[m] = acc;
      acc += (2 * m + 1) * (m / a.mrso + 1);
    }
    mbase[mmax + 1] = acc;
  }
  // c(|q|): rows iq and iq+1 of the (nq, np_new) table
  const double qn = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
  const double eff = qn / a.dq;
  int iq = static_cast<int>(eff);
  const double alpha = eff - static_cast<double>(iq);
  if (iq > a.nq - 2) iq = a.nq - 2;  // unreachable for |q| <= qmax (nq has 1.5 rows of margin)
  const T al = static_cast<T>(alpha), bl = static_cast<T>(1.0 - alpha);
  const cplx<T>* c0 = a.ctab + static_cast<long>(iq) * a.np_new;
  const cplx<T>* c1 = c0 + a.np_new;
  if (STAGE_C) {
    // batches of kCBatch rows per lane with all their loads in flight
    for (int ib = lane; ib < a.np_new; ib += kCBatch * kWave) {
      // plain scalars: arrays of HIP's vector classes are kept in scratch memory
      T lx[kCBatch], ly[kCBatch], hx[kCBatch], hy[kCBatch];
#pragma unroll
      for (int u = 0; u < kCBatch; ++u) {
        const int ia = min(ib + u * kWave, a.np_new - 1);
        const cplx<T> lo = c0[ia], hi = c1[ia];
        lx[u] = lo.x;
        ly[u] = lo.y;
        hx[u] = hi.x;
        hy[u] = hi.y;
      }
#pragma unroll
      for (int u = 0; u < kCBatch; ++u) {
        const int ia = ib + u * kWave;
        if (ia < a.np_new) ceff[ia] = mkc<T>(al * hx[u] + bl * lx[u], al * hy[u] + bl * ly[u]);
      }
    }
  }
  __syncthreads();

  // store the phased gathers g'[ip] = g[ip] e^{-i mup phi}
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int ip = lane + t * kWave;
    if (ip < np) {
      const cplx<T> e = ph[(info[t] >> 8) & 0xff];
      gq[ip] = cmul<T>(mkc<T>(g0x[t], g0y[t]), e);
      gmq[ip] = cmul<T>(mkc<T>(g1x[t], g1y[t]), e);
    }
  }
  __syncthreads();

  // rotation to the q frame: d_q[m,chi,mu2] = sum_mup g'_q d_{mup,chi};
  // d_-q[m,chi,mu2] = (-1)^m sum_mup g'_-q d_{mup,-chi}
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int ip = lane + t * kWave;
    if (ip >= np) continue;
    const int m = info[t] & 0xff, chi = ((info[t] >> 8) & 0xff) - mmax, mu2 = info[t] >> 16;
    const int km = m / a.mrso + 1, w = 2 * m + 1;
    const cplx<T>* g0p = gq + mbase[m] + mu2;
    const cplx<T>* g1p = gmq + mbase[m] + mu2;
    const T* d0 = dm + roff(m) + (chi + m);
    const T* d1 = dm + roff(m) + (-chi + m);
    T ar = 0, ai = 0, br = 0, bi = 0;
    for (int j = 0; j < w; ++j) {
      const cplx<T> x = g0p[j * km], y = g1p[j * km];
      const T u = d0[j * w], v = d1[j * w];
      ar = fma(x.x, u, ar);
      ai = fma(x.y, u, ai);
      br = fma(y.x, v, br);
      bi = fma(y.y, v, bi);
    }
    dq[ip] = mkc<T>(ar, ai);
    dmq[ip] = (m & 1) ? mkc<T>(-br, -bi) : mkc<T>(br, bi);
  }
  __syncthreads();

  // OZ product: for output (m, chi, mu2), terms (n, nu2), n >= |chi|
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int ip = lane + t * kWave;
    if (ip >= np) continue;
    const int chi = ((info[t] >> 8) & 0xff) - mmax;
    const int achi = chi < 0 ? -chi : chi;
    int ia = ias[t];
    cplx<T> s0 = mkc<T>(0, 0), s1 = mkc<T>(0, 0);
    for (int n = achi; n <= mmax; ++n) {
      const int kn = n / a.mrso + 1;
      const int base = mbase[n] + (chi + n) * kn;
      const int h = n / a.mrso;
      for (int nu2 = -h; nu2 < 0; ++nu2, ++ia) {
        cplx<T> ce;
        if (STAGE_C) {
          ce = ceff[ia];
        } else {
          const cplx<T> lo = c0[ia], hi = c1[ia];
          ce = mkc<T>(al * hi.x + bl * lo.x, al * hi.y + bl * lo.y);
        }
        const int src = base - nu2;
        s0 = cfma<T>(s0, ce, dq[src]);
        s1 = cfma<T>(s1, ce, dmq[src]);
      }
      for (int nu2 = 0; nu2 <= h; ++nu2, ++ia) {
        cplx<T> ce;
        if (STAGE_C) {
          ce = ceff[ia];
        } else {
          const cplx<T> lo = c0[ia], hi = c1[ia];
          ce = mkc<T>(al * hi.x + bl * lo.x, al * hi.y + bl * lo.y);
        }
        const int src = base + nu2;
        s0 = cfma_conjb<T>(s0, ce, dmq[src]);
        s1 = cfma_conjb<T>(s1, ce, dq[src]);
      }
    }
    gq[ip] = chop<T>(s0, a.chop_eps);
    gmq[ip] = chop<T>(s1, a.chop_eps);
  }
  __syncthreads();

  // inverse rotation and stores: out[m,mup,mu2] = e^{+i mup phi} sum_chi gamma d_{mup,chi}
  const bool self = (ixm == ix && iym == iy && izm == iz);
  const bool singular = self && (ix == a.nx / 2 || iy == a.ny / 2 || iz == a.nz / 2);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int ip = lane + t * kWave;
    if (ip >= np) continue;
    const int m = info[t] & 0xff, mup = ((info[t] >> 8) & 0xff) - mmax, mu2 = info[t] >> 16;
    const int km = m / a.mrso + 1, w = 2 * m + 1;
    const cplx<T>* g0p = gq + mbase[m] + mu2;
    const cplx<T>* g1p = gmq + mbase[m] + mu2;
    const T* drow = dm + roff(m) + (mup + m) * w;
    T ar = 0, ai = 0, br = 0, bi = 0;
    for (int j = 0; j < w; ++j) {
      const cplx<T> x = g0p[j * km], y = g1p[j * km];
      const T u = drow[j], v = drow[w - 1 - j];
      ar = fma(x.x, u, ar);
      ai = fma(x.y, u, ai);
      br = fma(y.x, v, br);
      bi = fma(y.y, v, bi);
    }
    const cplx<T> e = ph[mup + mmax];  // e^{-i mup phi}; the output takes its conjugate
    cplx<T> o0 = cmul<T>(mkc<T>(ar, ai), cconj<T>(e));
    cplx<T> o1 = cmul<T>(mkc<T>(br, bi), cconj<T>(e));
    if (m & 1) o1 = mkc<T>(-o1.x, -o1.y);
    p[static_cast<long>(ip) * a.pstride + sq] = o0;
    p[static_cast<long>(ip) * a.pstride + smq] = singular ? cconj<T>(o1) : o1;
  }
}

template <typename T>
T* upload(const std::vector<T>& h) {
  T* d = nullptr;
  MDFT_HIP_CHECK(hipMalloc(&d, std::max<size_t>(1, h.size()) * sizeof(T)));
  if (!h.empty()) MDFT_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

// the angular grids with a compile-time OZ kernel: mmax 1..6 with mrso 1 or 2
template <int MMAX = 1, int MRSO = 1, typename F>
bool dispatch_oz(int mmax, int mrso, F&& f) {
  if constexpr (MMAX > kMaxM) {
    return false;
  } else if constexpr (MRSO > 2) {
    return dispatch_oz<MMAX + 1, 1>(mmax, mrso, f);
  } else {
    if (mmax == MMAX && mrso == MRSO) {
      f(std::integral_constant<int, MMAX>{}, std::integral_constant<int, MRSO>{});
      return true;
    }
    return dispatch_oz<MMAX, MRSO + 1>(mmax, mrso, f);
  }
}

}  // namespace

std::vector<int> qpairs_memory_order(int nx, int ny, int nz) {
  auto neg = [](int i, int n) { return i == 0 ? 0 : n - i; };
  std::vector<int> out;
  for (int ix = 0; ix < nx; ++ix)
    for (int iy = 0; iy < ny; ++iy)
      for (int iz = 0; iz <= nz / 2; ++iz) {
        if (neg(iz, nz) == iz) {
          // both members lie in this z plane: q is the one the reference's
          // (iy, then ix) sweep meets first
          const int iym = neg(iy, ny), ixm = neg(ix, nx);
          if (!(iy < iym || (iy == iym && ix <= ixm))) continue;
        }
        out.insert(out.end(), {ix, iy, iz});
      }
  return out;
}

std::vector<int> oz_stars(const std::vector<int>& packed, int nx, int ny, int nz) {
  auto fold = [](int i, int n) { return i <= n / 2 ? i : n - i; };
  const long hx = nx / 2 + 1, hy = ny / 2 + 1, hz = nz / 2 + 1;
  std::vector<int> slot(static_cast<size_t>(hx * hy * hz) * 4, -1);
  for (size_t i = 0; i < packed.size(); ++i) {
    const int pk = packed[i];
    const int ix = pk & 1023, iy = (pk >> 10) & 1023, iz = (pk >> 20) & 1023;
    const long key = (fold(ix, nx) * hy + fold(iy, ny)) * hz + fold(iz, nz);
    int* s = &slot[static_cast<size_t>(key) * 4];
    int k = 0;
    while (k < 4 && s[k] >= 0) ++k;
    throw_if(k == 4, "internal error: more than four pairs in a star");
    s[k] = static_cast<int>(i);
  }
  std::vector<int> out;
  for (size_t key = 0; key < slot.size() / 4; ++key)
    if (slot[key * 4] >= 0) out.insert(out.end(), slot.begin() + key * 4, slot.begin() + key * 4 + 4);
  return out;
}

ConvDevice::ConvDevice(int nx, int ny, int nz, double lx, double ly, double lz, int mmax, int mrso,
                       const LucTable& luc)
    : tab(make_conv_tables(nx, ny, nz, lx, ly, lz, mmax, mrso, luc)), fft(nx, ny, nz) {
  const AngularGrid& g = tab.ang;
  throw_if(mmax > kMaxM, "the OZ kernel supports mmax <= 6");
  throw_if(2 * mmax + 1 > 255, "angular grid too large for the packed index table");
  // pairs: same set and designation as the reference list
  std::vector<int> mem = qpairs_memory_order(nx, ny, nz);
  {
    std::vector<int> ref = qpairs_reference(nx, ny, nz);
    throw_if(ref.size() != mem.size(), "internal error: q-pair count differs from the reference list");
    auto key = [&](const std::vector<int>& v, size_t i) {
      return (static_cast<long>(v[3 * i]) * ny + v[3 * i + 1]) * nz + v[3 * i + 2];
    };
    std::vector<long> a(ref.size() / 3), b(mem.size() / 3);
    for (size_t i = 0; i < a.size(); ++i) a[i] = key(ref, i);
    for (size_t i = 0; i < b.size(); ++i) b[i] = key(mem, i);
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    throw_if(a != b, "internal error: q-pair designation differs from the reference list");
  }
  npairs = static_cast<long>(mem.size() / 3);
  std::vector<int> packed(npairs);
  for (long i = 0; i < npairs; ++i) packed[i] = mem[3 * i] | (mem[3 * i + 1] << 10) | (mem[3 * i + 2] << 20);
  std::vector<double> k3;
  k3.insert(k3.end(), tab.kx.begin(), tab.kx.end());
  k3.insert(k3.end(), tab.ky.begin(), tab.ky.end());
  k3.insert(k3.end(), tab.kz.begin(), tab.kz.end());
  std::vector<int> info(g.np);
  for (int ip = 0; ip < g.np; ++ip) {
    info[ip] = g.m[ip] | ((g.mup[ip] + mmax) << 8) | (g.mu2[ip] << 16);
    // the kernel computes ip(m, mup, mu2) = base_m + (mup + m)(m/mrso + 1) + mu2
    int base = 0;
    for (int m = 0; m < g.m[ip]; ++m) base += (2 * m + 1) * (m / mrso + 1);
    throw_if(base + (g.mup[ip] + g.m[ip]) * (g.m[ip] / mrso + 1) + g.mu2[ip] != ip,
             "internal error: projection index formula");
  }
  std::vector<double2> t64(static_cast<size_t>(tab.nq) * tab.np_new);
  std::vector<float2> t32(t64.size());
  for (int ia = 0; ia < tab.np_new; ++ia)
    for (int iq = 0; iq < tab.nq; ++iq) {
      const cd v = tab.ctab[static_cast<size_t>(ia) * tab.nq + iq];
      t64[static_cast<size_t>(iq) * tab.np_new + ia] = make_double2(v.real(), v.imag());
      t32[static_cast<size_t>(iq) * tab.np_new + ia] =
          make_float2(static_cast<float>(v.real()), static_cast<float>(v.imag()));
    }
  const RotationCoeffs& rc = rotation_coeffs();
  std::vector<double> coef;
  coef.insert(coef.end(), &rc.a[0][0][0], &rc.a[0][0][0] + 5 * 13 * 13);
  coef.insert(coef.end(), &rc.b[0][0][0], &rc.b[0][0][0] + 5 * 13 * 13);
  coef.insert(coef.end(), &rc.c[0][0], &rc.c[0][0] + 5 * 13);
  coef.insert(coef.end(), &rc.d[0][0], &rc.d[0][0] + 5 * 13);
  if (dispatch_oz(mmax, mrso, [&](auto mm, auto mr) {
        constexpr int MMAX = decltype(mm)::value, MRSO = decltype(mr)::value;
        star32 = StarLayout<float, MMAX, MRSO>::fits;
        star64 = StarLayout<double, MMAX, MRSO>::fits;
      }))
    padw = 2 * mmax + 1;
  {
    const std::vector<int> st = oz_stars(packed, nx, ny, nz);
    nstars = static_cast<long>(st.size() / 4);
    long in_stars = 0;
    for (int v : st) in_stars += v >= 0;
    throw_if(in_stars != npairs, "internal error: the stars do not cover the pair list");
    // the pairs of a star share one d table: their designated q must have
    // the same iz (then the same kz and |q|, hence the same theta_q)
    for (long i = 0; i < nstars; ++i)
      for (int k = 1; k < 4; ++k) {
        const int j = st[4 * i + k];
        throw_if(j >= 0 && (packed[j] >> 20) != (packed[st[4 * i]] >> 20),
                 "internal error: the pairs of a star differ in kz");
      }
    stars = upload(st);
  }
  kxyz = upload(k3);
  pairs = upload(packed);
  ia_start = upload(tab.ia_start);
  ipinfo = upload(info);
  ctab64 = upload(t64);
  ctab32 = upload(t32);
  rcoef = upload(coef);
}

ConvDevice::~ConvDevice() {
  for (void* q : {static_cast<void*>(kxyz), static_cast<void*>(pairs), static_cast<void*>(ia_start),
                  static_cast<void*>(ipinfo), static_cast<void*>(ctab64), static_cast<void*>(ctab32),
                  static_cast<void*>(rcoef), dtab64, dtab32, pmeta32, pphase32, pmeta64, pphase64,
                  static_cast<void*>(dmap), static_cast<void*>(stars), static_cast<void*>(dent),
                  static_cast<void*>(ccode)})
    if (q) (void)hipFree(q);
}

size_t ConvDevice::oz_lds_bytes(size_t elem) const {
  const int mmax = tab.mmax;
  const size_t nR = static_cast<size_t>(mmax + 1) * (4 * (mmax + 1) * (mmax + 1) - 1) / 3;
  const size_t stage_c = elem == sizeof(float) ? static_cast<size_t>(tab.np_new) * 2 * elem : 0;
  return 4 * static_cast<size_t>(tab.ang.np) * 2 * elem + ((nR * elem + 15) / 16) * 16 + stage_c;
}

template <typename T>
static OzArgs<T> oz_args(const ConvDevice& c, double chop_eps) {
  OzArgs<T> a;
  const ConvTables& tab = c.tab;
  a.chop_eps = static_cast<T>(chop_eps);
  a.nx = tab.nx; a.ny = tab.ny; a.nz = tab.nz; a.mmax = tab.mmax; a.mrso = tab.mrso;
  a.np = tab.ang.np; a.np_new = tab.np_new; a.nq = tab.nq; a.dq = tab.dq;
  a.npairs = c.npairs;
  a.chunk = (c.npairs + kXcds - 1) / kXcds;
  a.N = static_cast<long>(tab.nx) * tab.ny * tab.nz;
  a.pstride = a.N;
  a.nbatch = 1;
  a.kxyz = c.kxyz; a.pairs = c.pairs; a.ipinfo = c.ipinfo; a.ia_start = c.ia_start;
  a.ctab = std::is_same<T, double>::value ? reinterpret_cast<const cplx<T>*>(c.ctab64)
                                          : reinterpret_cast<const cplx<T>*>(c.ctab32);
  a.rc = c.rcoef;
  return a;
}

template <typename T>
const T* ConvDevice::dtable(hipStream_t s) const {
  void*& slot = std::is_same<T, double>::value ? dtab64 : dtab32;
  if (!slot) {
    const size_t nR = static_cast<size_t>(tab.mmax + 1) * (4 * (tab.mmax + 1) * (tab.mmax + 1) - 1) / 3;
    // the compile-time kernel reads half tables (rows mup >= 0)
    const int half = padw ? 1 : 0;
    const size_t rows = half ? static_cast<size_t>(tab.mmax + 1) * (tab.mmax + 2) / 2
                             : static_cast<size_t>(tab.mmax + 1) * (tab.mmax + 1);
    const size_t row = padw ? rows * padw : nR;
    // one table per star when the star kernel serves T, else one per pair
    const bool star = std::is_same<T, double>::value ? star64 : star32;
    const long nrows = star ? nstars : npairs;
    MDFT_HIP_CHECK(hipMalloc(&slot, std::max<size_t>(1, nrows * row * sizeof(T))));
    OzArgs<T> a = oz_args<T>(*this, 0.0);
    hipLaunchKernelGGL((oz_dtable_kernel<T>), dim3(static_cast<unsigned>(nrows)), dim3(kWave),
                       nR * sizeof(double2), s, static_cast<T*>(slot), a, padw, half, star ? stars : nullptr, nrows);
    MDFT_HIP_CHECK(hipGetLastError());
  }
  return static_cast<const T*>(slot);
}

template <typename T>
const void* ConvDevice::pair_meta(hipStream_t s, const void** phases) const {
  constexpr bool f64 = std::is_same<T, double>::value;
  void*& meta = f64 ? pmeta64 : pmeta32;
  void*& phase = f64 ? pphase64 : pphase32;
  if (!meta) {
    MDFT_HIP_CHECK(hipMalloc(&meta, std::max<size_t>(1, npairs * sizeof(OzPairMeta<T>))));
    MDFT_HIP_CHECK(hipMalloc(&phase, std::max<size_t>(1, npairs * (tab.mmax + 1) * sizeof(cplx<T>))));
    if (!dmap) dmap = upload(oz_thread_records(tab.ang, tab.ia_start));
    OzArgs<T> a = oz_args<T>(*this, 0.0);
    const unsigned nb = static_cast<unsigned>((npairs + 255) / 256);
    hipLaunchKernelGGL((oz_meta_kernel<T>), dim3(nb), dim3(256), 0, s, static_cast<OzPairMeta<T>*>(meta),
                       static_cast<cplx<T>*>(phase), a);
    MDFT_HIP_CHECK(hipGetLastError());
  }
  bool& checked = f64 ? stars_checked64 : stars_checked32;
  if (!checked) {
    // the star kernel lerps one c(q) row per star: every pair of a star
    // must have the same row and weights, bit for bit (once per Convolution)
    std::vector<OzPairMeta<T>> hm(static_cast<size_t>(npairs));
    std::vector<int> hs(static_cast<size_t>(nstars) * 4);
    MDFT_HIP_CHECK(hipMemcpyAsync(hm.data(), meta, hm.size() * sizeof(OzPairMeta<T>), hipMemcpyDeviceToHost, s));
    MDFT_HIP_CHECK(hipMemcpyAsync(hs.data(), stars, hs.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    MDFT_HIP_CHECK(hipStreamSynchronize(s));
    for (long i = 0; i < nstars; ++i) {
      const OzPairMeta<T>& m0 = hm[hs[4 * i]];
      for (int k = 1; k < 4; ++k) {
        const int j = hs[4 * i + k];
        if (j < 0) continue;
        const OzPairMeta<T>& mj = hm[j];
        throw_if(mj.iq != m0.iq || std::memcmp(&mj.al, &m0.al, sizeof(T)) != 0 ||
                     std::memcmp(&mj.bl, &m0.bl, sizeof(T)) != 0,
                 "internal error: the pairs of a star have different c(q) rows or weights");
      }
    }
    checked = true;
  }
  *phases = phase;
  return meta;
}

// launches kernel on s; with events, through hipExtLaunchKernel, so that the
// dispatch itself records its start and end times (no event packets of
// their own between the kernels)
template <typename K, typename... Args>
void launch_oz(K kernel, unsigned nb, unsigned threads, size_t lds, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1,
               Args... args) {
  if (ev0 || ev1)
    hipExtLaunchKernelGGL(kernel, dim3(nb), dim3(threads), static_cast<std::uint32_t>(lds), s, ev0, ev1, 0u,
                          args...);
  else
    hipLaunchKernelGGL(kernel, dim3(nb), dim3(threads), lds, s, args...);
  MDFT_HIP_CHECK(hipGetLastError());
}

template <typename T>
void ConvDevice::oz_pass(cplx<T>* p, double chop_eps, hipStream_t s, int nbatch, hipEvent_t ev0,
                         hipEvent_t ev1, bool compact) const {
  throw_if(compact && !compact_supported(std::is_same<T, double>::value),
           "OZ pass: the compact channel order needs the star kernel");
  if (npairs == 0) {
    if (ev0) MDFT_HIP_CHECK(hipEventRecord(ev0, s));
    if (ev1) MDFT_HIP_CHECK(hipEventRecord(ev1, s));
    return;
  }
  throw_if(tab.ang.np > 4 * kWave, "the OZ kernel handles at most 256 projections");
  throw_if(nbatch < 1, "OZ pass: bad batch");
  const T* dt = dtable<T>(s);
  OzArgs<T> a = oz_args<T>(*this, chop_eps);
  a.nbatch = nbatch;
  a.pstride = a.N * nbatch;
  const long nblk = a.chunk * kXcds * nbatch;
  throw_if(nblk > 0x7fffffffL, "too many q pairs");
  if (padw) {
    const cplx<T>* ph = nullptr;
    const OzPairMeta<T>* mt = static_cast<const OzPairMeta<T>*>(pair_meta<T>(s, reinterpret_cast<const void**>(&ph)));
    dispatch_oz(tab.mmax, tab.mrso, [&](auto mm, auto mr) {
      constexpr int MMAX = decltype(mm)::value, MRSO = decltype(mr)::value;
      using Sh = OzShape<MMAX, MRSO>;
      constexpr int PB = oz_pairs_per_block<T, Sh::NP>();
      constexpr int NC = oz_terms<MMAX, MRSO>();
      throw_if(tab.np_new != NC, "internal error: OZ term count");
      throw_if(static_cast<double>(tab.nq) * NC * sizeof(cplx<T>) >= 4294967296.0, "c(q) table too large");
      const size_t pair_lds = (2 * static_cast<size_t>(Sh::NG) * Sh::W + 2 * Sh::NP + NC) * sizeof(cplx<T>) +
                              static_cast<size_t>(Sh::W) * (Sh::ND + 1) * sizeof(T);
      const size_t region = (pair_lds + 15) / 16 * 16;
      throw_if(PB * region > 64 * 1024, "angular grid too large for the OZ kernel");
      using Lay = StarLayout<T, MMAX, MRSO>;
      if constexpr (Lay::fits) {
        OzArgs<T> ag = a;
        ag.chunk = (nstars + kXcds - 1) / kXcds;
        const long nb = ag.chunk * kXcds * nbatch;
        throw_if(nb > 0x7fffffffL, "too many q pairs");
        if (!dent) dent = upload(oz_dtable_places(MMAX));
        if (compact) {
          if (!ccode) ccode = upload(compact_codes(tab.ang));
          launch_oz(oz_star_ct_kernel<T, MMAX, MRSO, Lay::threads, true>, static_cast<unsigned>(nb),
                    static_cast<unsigned>(Lay::threads), Lay::dynamic, s, ev0, ev1, p, dt, mt, ph, dmap, dent,
                    reinterpret_cast<const int4*>(stars), nstars, ag, static_cast<const int*>(ccode));
          return;
        }
        launch_oz(oz_star_ct_kernel<T, MMAX, MRSO, Lay::threads, false>, static_cast<unsigned>(nb),
                  static_cast<unsigned>(Lay::threads), Lay::dynamic, s, ev0, ev1, p, dt, mt, ph, dmap, dent,
                  reinterpret_cast<const int4*>(stars), nstars, ag, static_cast<const int*>(nullptr));
        return;
      }
      OzArgs<T> ag = a;
      const long ngroups = (npairs + PB - 1) / PB;
      ag.chunk = (ngroups + kXcds - 1) / kXcds;
      const long nb = ag.chunk * kXcds * nbatch;
      throw_if(nb > 0x7fffffffL, "too many q pairs");
      const int threads = (PB * Sh::NP + kWave - 1) / kWave * kWave;
      launch_oz(oz_pairs_ct_kernel<T, MMAX, MRSO, PB>, static_cast<unsigned>(nb), static_cast<unsigned>(threads),
                PB * region, s, ev0, ev1, p, dt, mt, ph, dmap, ag, static_cast<int>(region));
    });
    return;
  }
  const size_t lds = oz_lds_bytes(sizeof(T));
  throw_if(lds > 64 * 1024, "angular grid too large for the OZ kernel");
  constexpr bool stage = std::is_same<T, float>::value;
  launch_oz(oz_pair_kernel<T, stage>, static_cast<unsigned>(nblk), kWave, lds, s, ev0, ev1, p, dt, a);
}

template <typename T>
void ConvDevice::execute(cplx<T>* p, double chop_eps, hipStream_t s, StageTimer* timer, int nbatch,
                         bool compact) const {
  // the compact order transforms the stored projections only
  const long batch = static_cast<long>(compact ? compact_projections(tab.ang) : tab.ang.np) * nbatch;
  if (!fft.tuned<T>()) fft.autotune<T>(batch, s);  // once per element type, before any timing mark
  if (timer) timer->mark(s, kStageConvStart);
  fft.execute<T>(p, batch, -1, s);
  if (timer) timer->mark(s, kStageFftFwd);
  if (timer && timer->oz_only())
    oz_pass<T>(p, chop_eps, s, nbatch, timer->event(kStageFftFwd), timer->event(kStageOz), compact);
  else
    oz_pass<T>(p, chop_eps, s, nbatch, nullptr, nullptr, compact);
  if (timer) timer->mark(s, kStageOz);
  fft.execute<T>(p, batch, +1, s);
  if (timer) timer->mark(s, kStageFftBwd);
}

template void ConvDevice::oz_pass<double>(cplx<double>*, double, hipStream_t, int, hipEvent_t, hipEvent_t,
                                          bool) const;
template void ConvDevice::oz_pass<float>(cplx<float>*, double, hipStream_t, int, hipEvent_t, hipEvent_t, bool) const;
template void ConvDevice::execute<double>(cplx<double>*, double, hipStream_t, StageTimer*, int, bool) const;
template void ConvDevice::execute<float>(cplx<float>*, double, hipStream_t, StageTimer*, int, bool) const;

bool ConvDevice::compact_supported(bool fp64) const {
  if (npairs == 0 || tab.ang.np > 4 * kWave || !padw) return false;
  bool ok = false;
  dispatch_oz(tab.mmax, tab.mrso, [&](auto mm, auto mr) {
    constexpr int MM = decltype(mm)::value, MR = decltype(mr)::value;
    ok = fp64 ? StarLayout<double, MM, MR>::fits : StarLayout<float, MM, MR>::fits;
  });
  return ok;
}

}  // namespace mdft
