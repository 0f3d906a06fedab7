// Convolution::execute for gfx950 (common/MDFT_Convolution.hpp:324-548).
//
// The forward and backward 3D FFTs are the Stockham passes of mdft_fft.hip.
// The per-(q, -q) pass (the reference's "convolution" team kernel, :376-543)
// runs one wavefront per pair:
//   1. R^l_{m,m1}(q) by the reference recursion (common/MDFT_Rotation.hpp:
//      103-283) in FP64, parallel over (m, m1) inside each level l, stored
//      compactly (only |m|, |m1| <= l); entries the recursion reads outside a
//      block are taken as 0 (their coefficients vanish);
//   2. the rotation to the q frame (:427-443), the c(|q|) interpolation
//      (:449-463, weights in FP64), the OZ product (:468-496) with the
//      precomputed (ip, n, nu2) -> source-projection map, the chop of gamma
//      (:514-517) and the inverse rotation and the q / -q stores (:518-541,
//      conjugated -q store at self-conjugate mid-k points).
// Pairs are processed in (ix, iy, iz) memory order (same set and q / -q
// designation as the reference list) and the block -> pair map gives every XCD
// a contiguous range of pairs, so neighbouring pairs share L2 lines.
#include "mdft_conv.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <type_traits>

namespace mdft {

namespace {

constexpr int kWave = 64;
constexpr int kXcds = 8;

template <typename T>
struct OzArgs {
  int nx, ny, nz, mmax, mrso, np, np_new, nq, nmu2;
  double dq;
  long npairs, chunk, N;
  const double* kxyz;
  const int* pairs;
  const int* ia_start;
  const int* oz_src;
  const int* ipinfo;
  const int* pmap;
  const cplx<T>* ctab;  // (nq, np_new)
  const double* rc;     // a (5x13x13) | b (5x13x13) | c (5x13) | d (5x13)
  T chop_eps;
};


__device__ __forceinline__ int roff(int l) { return l * (4 * l * l - 1) / 3; }

__device__ __forceinline__ double2 rget(const double2* R, int l, int m, int m1) {
  if (m < -l || m > l || m1 < -l || m1 > l) return make_double2(0.0, 0.0);
  return R[roff(l) + (m + l) * (2 * l + 1) + (m1 + l)];
}
__device__ __forceinline__ void rset(double2* R, int l, int m, int m1, double2 v) {
  R[roff(l) + (m + l) * (2 * l + 1) + (m1 + l)] = v;
}

template <typename T>
__device__ __forceinline__ cplx<T> to_t(double2 v) {
  return mkc<T>(static_cast<T>(v.x), static_cast<T>(v.y));
}

// prevent_underflow (common/MDFT_Math_Utils.hpp:126-137): absolute threshold per component
template <typename T>
__device__ __forceinline__ cplx<T> chop(cplx<T> v, T eps) {
  if (fabs(v.x) < eps) v.x = 0;
  if (fabs(v.y) < eps) v.y = 0;
  return v;
}

// One pair per 64-lane block.
template <typename T>
__global__ __launch_bounds__(kWave) void oz_pair_kernel(cplx<T>* __restrict__ p, OzArgs<T> a) {
  extern __shared__ unsigned char lds_raw[];
  const int nR = roff(a.mmax + 1);
  double2* R = reinterpret_cast<double2*>(lds_raw);
  cplx<T>* gq = reinterpret_cast<cplx<T>*>(R + nR);
  cplx<T>* gmq = gq + a.np;
  cplx<T>* dq = gmq + a.np;
  cplx<T>* dmq = dq + a.np;
  const int lane = threadIdx.x;

  const long b = blockIdx.x;
  const long pair = (b % kXcds) * a.chunk + b / kXcds;
  if (pair >= a.npairs) return;  // uniform across the block
  const int pk = a.pairs[pair];
  const int ix = pk & 1023, iy = (pk >> 10) & 1023, iz = (pk >> 20) & 1023;
  const int ixm = ix == 0 ? 0 : a.nx - ix, iym = iy == 0 ? 0 : a.ny - iy, izm = iz == 0 ? 0 : a.nz - iz;
  const long sq = (static_cast<long>(ix) * a.ny + iy) * a.nz + iz;
  const long smq = (static_cast<long>(ixm) * a.ny + iym) * a.nz + izm;
  const double q0 = a.kxyz[ix], q1 = a.kxyz[a.nx + iy], q2 = a.kxyz[a.nx + a.ny + iz];
  const int mmax = a.mmax;

  // ---- R(q): levels 0 and 1 (Rotation.hpp:116-174) ----
  if (lane == 0) {
    double ez0, ez1, ez2, ey0, ey1, ey2;
    if (q0 == 0.0 && q1 == 0.0 && q2 == 0.0) {
      ez0 = 0.0; ez1 = 0.0; ez2 = 1.0;
    } else {
      ez0 = q0; ez1 = q1; ez2 = q2;
    }
    if (ez0 != 0.0 || ez1 != 0.0) {
      ey0 = -ez1; ey1 = ez0; ey2 = 0.0;
    } else {
      ey0 = 0.0; ey1 = ez2; ey2 = -ez1;
    }
    double s = sqrt(ez0 * ez0 + ez1 * ez1 + ez2 * ez2);
    ez0 /= s; ez1 /= s; ez2 /= s;
    s = sqrt(ey0 * ey0 + ey1 * ey1 + ey2 * ey2);
    ey0 /= s; ey1 /= s; ey2 /= s;
    const double ex0 = ey1 * ez2 - ey2 * ez1, ex1 = ey2 * ez0 - ey0 * ez2, ex2 = ey0 * ez1 - ey1 * ez0;
    const double h = 1.0 / sqrt(2.0);
    rset(R, 0, 0, 0, make_double2(1.0, 0.0));
    if (mmax >= 1) {
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
  }
  __syncthreads();

  // ---- levels 2..mmax (Rotation.hpp:177-282) ----
  const double* ca = a.rc;
  const double* cb = a.rc + 5 * 13 * 13;
  const double* cc = a.rc + 2 * 5 * 13 * 13;
  const double* cdd = a.rc + 2 * 5 * 13 * 13 + 5 * 13;
  for (int l = 2; l <= mmax; ++l) {
    const int k = l - 1;
    const int nlow = (l + 1) * (l + 1);
    const int nitems = nlow + l * l;
    for (int it = lane; it < nitems; it += kWave) {
      int m, m1;
      if (it < nlow) {
        m = -l + it / (l + 1);
        m1 = it % (l + 1);
      } else {
        const int t = it - nlow;
        m = 1 + t / l;
        m1 = 1 + t % l;
      }
      const int L6 = 6, lo = l - 2;
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

  // ---- gather g(q), g(-q) (Convolution.hpp:420-425) ----
  for (int ip = lane; ip < a.np; ip += kWave) {
    gq[ip] = p[static_cast<long>(ip) * a.N + sq];
    gmq[ip] = p[static_cast<long>(ip) * a.N + smq];
  }
  __syncthreads();

  const int W = 2 * mmax + 1;
  // ---- rotation to the q frame (:427-443) ----
  for (int ip = lane; ip < a.np; ip += kWave) {
    const int info = a.ipinfo[ip];
    const int im = info & 0xff, ikhi = ((info >> 8) & 0xff) - mmax, imu2 = info >> 16;
    cplx<T> s0 = mkc<T>(0, 0), s1 = mkc<T>(0, 0);
    for (int imup = -im; imup <= im; ++imup) {
      const int src = a.pmap[(im * W + imup + mmax) * a.nmu2 + imu2];
      s0 = cfma<T>(s0, gq[src], to_t<T>(rget(R, im, imup, ikhi)));
      s1 = cfma<T>(s1, gmq[src], to_t<T>(rget(R, im, imup, -ikhi)));
    }
    dq[ip] = s0;
    dmq[ip] = (im & 1) ? mkc<T>(-s1.x, -s1.y) : s1;
  }
  __syncthreads();

  // ---- c(|q|) interpolation weights (:449-455) ----
  const double qn = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
  const double eff = qn / a.dq;
  int iq = static_cast<int>(eff);
  const double alpha = eff - static_cast<double>(iq);
  if (iq > a.nq - 2) iq = a.nq - 2;  // unreachable for |q| <= qmax (nq has 1.5 rows of margin)
  const T al = static_cast<T>(alpha), bl = static_cast<T>(1.0 - alpha);
  const cplx<T>* c0 = a.ctab + static_cast<long>(iq) * a.np_new;
  const cplx<T>* c1 = c0 + a.np_new;

  // ---- OZ product (:468-496); gamma overwrites g ----
  for (int ip = lane; ip < a.np; ip += kWave) {
    cplx<T> s0 = mkc<T>(0, 0), s1 = mkc<T>(0, 0);
    const int e = a.ia_start[ip + 1];
    for (int ia = a.ia_start[ip]; ia < e; ++ia) {
      const int src = a.oz_src[ia];
      const int ipm = src & 0xffff;
      const cplx<T> lo = c0[ia], hi = c1[ia];
      const cplx<T> ce = mkc<T>(al * hi.x + bl * lo.x, al * hi.y + bl * lo.y);
      if (src >> 16) {
        s0 = cfma<T>(s0, ce, dq[ipm]);
        s1 = cfma<T>(s1, ce, dmq[ipm]);
      } else {
        s0 = cfma_conjb<T>(s0, ce, dmq[ipm]);
        s1 = cfma_conjb<T>(s1, ce, dq[ipm]);
      }
    }
    gq[ip] = chop<T>(s0, a.chop_eps);
    gmq[ip] = chop<T>(s1, a.chop_eps);
  }
  __syncthreads();

  // ---- inverse rotation and stores (:518-541) ----
  const bool self = (ixm == ix && iym == iy && izm == iz);
  const bool singular = self && (ix == a.nx / 2 || iy == a.ny / 2 || iz == a.nz / 2);
  for (int ip = lane; ip < a.np; ip += kWave) {
    const int info = a.ipinfo[ip];
    const int im = info & 0xff, imup = ((info >> 8) & 0xff) - mmax, imu2 = info >> 16;
    cplx<T> s0 = mkc<T>(0, 0), s1 = mkc<T>(0, 0);
    for (int ikhi = -im; ikhi <= im; ++ikhi) {
      const int src = a.pmap[(im * W + ikhi + mmax) * a.nmu2 + imu2];
      s0 = cfma_conjb<T>(s0, gq[src], to_t<T>(rget(R, im, imup, ikhi)));
      s1 = cfma_conjb<T>(s1, gmq[src], to_t<T>(rget(R, im, imup, -ikhi)));
    }
    if (im & 1) s1 = mkc<T>(-s1.x, -s1.y);
    p[static_cast<long>(ip) * a.N + sq] = s0;
    p[static_cast<long>(ip) * a.N + smq] = singular ? cconj<T>(s1) : s1;
  }
}

template <typename T>
T* upload(const std::vector<T>& h) {
  T* d = nullptr;
  MDFT_HIP_CHECK(hipMalloc(&d, std::max<size_t>(1, h.size()) * sizeof(T)));
  if (!h.empty()) MDFT_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
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

ConvDevice::ConvDevice(int nx, int ny, int nz, double lx, double ly, double lz, int mmax, int mrso,
                       const LucTable& luc)
    : tab(make_conv_tables(nx, ny, nz, lx, ly, lz, mmax, mrso, luc)), fft(nx, ny, nz) {
  const AngularGrid& g = tab.ang;
  throw_if(g.np >= (1 << 16), "too many projections");
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
  std::vector<int> src(tab.np_new), info(g.np);
  for (int ip = 0, ia = 0; ip < g.np; ++ip) {
    const int khi = g.mup[ip];
    info[ip] = g.m[ip] | ((g.mup[ip] + mmax) << 8) | (g.mu2[ip] << 16);
    for (int n = std::abs(khi); n <= mmax; ++n)
      for (int nu2 = -n / mrso; nu2 <= n / mrso; ++nu2, ++ia) {
        const int ipm = g.P(n, khi, std::abs(nu2));
        src[ia] = ipm | (nu2 < 0 ? (1 << 16) : 0);
      }
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
  kxyz = upload(k3);
  pairs = upload(packed);
  ia_start = upload(tab.ia_start);
  oz_src = upload(src);
  ipinfo = upload(info);
  pmap = upload(g.pmap);
  ctab64 = upload(t64);
  ctab32 = upload(t32);
  rcoef = upload(coef);
}

ConvDevice::~ConvDevice() {
  for (void* q : {static_cast<void*>(kxyz), static_cast<void*>(pairs), static_cast<void*>(ia_start),
                  static_cast<void*>(oz_src), static_cast<void*>(ipinfo), static_cast<void*>(pmap),
                  static_cast<void*>(ctab64), static_cast<void*>(ctab32), static_cast<void*>(rcoef)})
    if (q) (void)hipFree(q);
}

template <typename T>
void ConvDevice::oz_pass(cplx<T>* p, double chop_eps, hipStream_t s) const {
  if (npairs == 0) return;
  OzArgs<T> a;
  a.chop_eps = static_cast<T>(chop_eps);
  const AngularGrid& g = tab.ang;
  a.nx = tab.nx; a.ny = tab.ny; a.nz = tab.nz; a.mmax = tab.mmax; a.mrso = tab.mrso;
  a.np = g.np; a.np_new = tab.np_new; a.nq = tab.nq; a.nmu2 = g.nmu2; a.dq = tab.dq;
  a.npairs = npairs;
  a.chunk = (npairs + kXcds - 1) / kXcds;
  a.N = static_cast<long>(tab.nx) * tab.ny * tab.nz;
  a.kxyz = kxyz; a.pairs = pairs; a.ia_start = ia_start; a.oz_src = oz_src; a.ipinfo = ipinfo; a.pmap = pmap;
  a.ctab = std::is_same<T, double>::value ? reinterpret_cast<const cplx<T>*>(ctab64)
                                          : reinterpret_cast<const cplx<T>*>(ctab32);
  a.rc = rcoef;
  const long nblk = a.chunk * kXcds;
  throw_if(nblk > 0x7fffffffL, "too many q pairs");
  const int nR = (tab.mmax + 1) * (4 * (tab.mmax + 1) * (tab.mmax + 1) - 1) / 3;
  const size_t lds = nR * sizeof(double2) + 4 * static_cast<size_t>(g.np) * sizeof(cplx<T>);
  throw_if(lds > 64 * 1024, "angular grid too large for the OZ kernel");
  hipLaunchKernelGGL((oz_pair_kernel<T>), dim3(static_cast<unsigned>(nblk)), dim3(kWave), lds, s, p, a);
  MDFT_HIP_CHECK(hipGetLastError());
}

template <typename T>
void ConvDevice::execute(cplx<T>* p, double chop_eps, hipStream_t s, StageTimer* timer) const {
  const long batch = tab.ang.np;
  if (timer) timer->mark(s, kStageConvStart);
  fft.execute<T>(p, batch, -1, s);
  if (timer) timer->mark(s, kStageFftFwd);
  oz_pass<T>(p, chop_eps, s);
  if (timer) timer->mark(s, kStageOz);
  fft.execute<T>(p, batch, +1, s);
  if (timer) timer->mark(s, kStageFftBwd);
}

template void ConvDevice::oz_pass<double>(cplx<double>*, double, hipStream_t) const;
template void ConvDevice::oz_pass<float>(cplx<float>*, double, hipStream_t) const;
template void ConvDevice::execute<double>(cplx<double>*, double, hipStream_t, StageTimer*) const;
template void ConvDevice::execute<float>(cplx<float>*, double, hipStream_t, StageTimer*) const;

}  // namespace mdft
