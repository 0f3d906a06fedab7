"""ctypes face of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module, and only as the checker (or the timed CPU baseline).  The
product (``mini_mdft_amd``) never imports it.

What pins each oracle function (tests/test_oracle_*.py):

* ``gauss_legendre``      -- numpy ``leggauss`` (nodes reversed, weights / 2), Sum w = 1.
* ``wigner_small_d``      -- closed forms d^1, d^2 (PDG table), d^m_00 = P_m(cos),
                             d^m_{a0} via scipy ``lpmv``, symmetries, GL orthogonality.
* ``angular_tables``      -- the grid formulas (sizes, Sum w = 8 pi^2 / mrso, io order).
* ``rotation_matrix``     -- unitarity of every l block, R(z) = I, agreement with
                             e^{i m phi_q} d^l_{m m1}(theta_q) built from the oracle's
                             own Wigner-d (itself pinned above).
* ``qpairs``              -- coverage and uniqueness (unit_test/Test_Convolution.cpp:54-77).
* ``fft1d`` / ``fft3d``   -- numpy.fft (sign, normalisation), brute-force DFT.
* ``angl2proj``           -- numpy rfft2 restatement and a dense explicit-phase matrix;
                             constant field -> only m = 0.
* ``proj2angl``           -- numpy irfft2 (non-Hermitian input included), dense matrix,
                             idempotence angl2proj(proj2angl(angl2proj(o))) = angl2proj(o).
* ``conv``                -- m = 0 closed form (N * ifftn(f(|k|) fftn(p)) for real p),
                             linearity, c = 0 -> 0, rotation covariance (90 degrees about
                             z), Luc reader round trip, table reorder against the reference's
                             loop order.
* ``delta_rho`` / ``delta_f`` -- numpy evaluation of the formulas
                             (unit_test/Test_Solvent.cpp:129-211).
* ``energy_cproj``        -- composition of the pinned stages.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmdft_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build() -> str:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    import subprocess

    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_conv_create.restype = ctypes.c_void_p
        L.orc_conv_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_int] * 3 + \
            [_ip] * 5 + [_dp, ctypes.c_int, ctypes.c_double]
        L.orc_conv_destroy.argtypes = [ctypes.c_void_p]
        L.orc_conv_info.argtypes = [ctypes.c_void_p, _ip, _dp]
        L.orc_conv_table.argtypes = [ctypes.c_void_p, _dp]
        L.orc_conv_execute.argtypes = [ctypes.c_void_p, _dp]
        L.orc_energy_cproj.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.orc_delta_rho.argtypes = [ctypes.c_long, _dp, _dp, ctypes.c_double]
        L.orc_delta_f.argtypes = [ctypes.c_long, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double,
                                  ctypes.c_double, _dp]
        L.orc_wigner_small_d.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, _dp]
        L.orc_kproj.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.orc_fft1d.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        L.orc_fft3d.argtypes = [_dp] + [ctypes.c_int] * 5
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def gauss_legendre(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    _check(lib().orc_gauss_legendre(n, _d(x), _d(w)))
    return x, w


def wigner_small_d(m: int, mu: int, mup: int, theta: float) -> float:
    out = ctypes.c_double()
    _check(lib().orc_wigner_small_d(m, mu, mup, ctypes.c_double(theta), ctypes.byref(out)))
    return out.value


def angular_sizes(mmax: int, mrso: int) -> dict:
    s = np.zeros(6, dtype=np.int32)
    _check(lib().orc_angular_sizes(mmax, mrso, _i(s)))
    return dict(zip(("ntheta", "nphi", "npsi", "no", "np", "nmu2"), (int(v) for v in s)))


def angular_tables(mmax: int, mrso: int) -> dict:
    s = angular_sizes(mmax, mrso)
    nt, no, np_, nmu2 = s["ntheta"], s["no"], s["np"], s["nmu2"]
    theta, wtheta, w = np.zeros(nt), np.zeros(nt), np.zeros(no)
    pmap = np.zeros((mmax + 1, 2 * mmax + 1, nmu2), dtype=np.int32)
    m, mup, mu, mu2 = (np.zeros(np_, dtype=np.int32) for _ in range(4))
    D = np.zeros((nt, np_))
    _check(lib().orc_angular_tables(mmax, mrso, _d(theta), _d(wtheta), _d(w), _i(pmap), _i(m), _i(mup),
                                    _i(mu), _i(mu2), _d(D)))
    s.update(theta=theta, wtheta=wtheta, w=w, pmap=pmap, m=m, mup=mup, mu=mu, mu2=mu2, D=D,
             fm=np.sqrt(2.0 * np.arange(mmax + 1) + 1.0))
    return s


def kproj(n: int, L: float) -> np.ndarray:
    k = np.zeros(n)
    _check(lib().orc_kproj(n, ctypes.c_double(L), _d(k)))
    return k


def rotation_matrix(mmax: int, q) -> np.ndarray:
    R = np.zeros((mmax + 1, 2 * mmax + 1, 2 * mmax + 1), dtype=np.complex128)
    qq = np.ascontiguousarray(q, dtype=np.float64)
    _check(lib().orc_rotation_matrix(mmax, _d(qq), R.ctypes.data_as(_dp)))
    return R


def qpairs(nx: int, ny: int, nz: int) -> np.ndarray:
    L = lib().orc_qpair_count(nx, ny, nz)
    out = np.zeros((L, 3), dtype=np.int32)
    _check(lib().orc_qpairs(nx, ny, nz, _i(out)))
    return out


def fft1d(x: np.ndarray, sign: int) -> np.ndarray:
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    _check(lib().orc_fft1d(y.size, y.ctypes.data_as(_dp), sign))
    return y


def fft3d(x: np.ndarray, sign: int) -> np.ndarray:
    """Unnormalised in-place-style 3D transform over the last three axes."""
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    nx, ny, nz = y.shape[-3:]
    batch = int(np.prod(y.shape[:-3])) if y.ndim > 3 else 1
    _check(lib().orc_fft3d(y.ctypes.data_as(_dp), batch, nx, ny, nz, sign))
    return y


def angl2proj(o: np.ndarray, n, mmax: int, mrso: int) -> np.ndarray:
    """o: (N, ntheta, nphi, npsi) float64 -> p: (np, nx, ny, nz) complex128."""
    nx, ny, nz = n
    s = angular_sizes(mmax, mrso)
    o = np.ascontiguousarray(o, dtype=np.float64)
    p = np.zeros((s["np"], nx, ny, nz), dtype=np.complex128)
    _check(lib().orc_angl2proj(nx, ny, nz, mmax, mrso, _d(o), p.ctypes.data_as(_dp)))
    return p


def proj2angl(p: np.ndarray, mmax: int, mrso: int) -> np.ndarray:
    """p: (np, nx, ny, nz) complex128 -> o: (N, ntheta, nphi, npsi) float64."""
    nx, ny, nz = p.shape[1:]
    s = angular_sizes(mmax, mrso)
    p = np.ascontiguousarray(p, dtype=np.complex128)
    o = np.zeros((nx * ny * nz, s["ntheta"], s["nphi"], s["npsi"]))
    _check(lib().orc_proj2angl(nx, ny, nz, mmax, mrso, p.ctypes.data_as(_dp), _d(o)))
    return o


def delta_rho(xi: np.ndarray, rho0: float) -> np.ndarray:
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    out = np.empty_like(xi)
    _check(lib().orc_delta_rho(xi.size, _d(xi), _d(out), ctypes.c_double(rho0)))
    return out


def delta_f(xi: np.ndarray, vexc: np.ndarray, w: np.ndarray, rho0: float, prefactor: float):
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    vexc = np.ascontiguousarray(vexc, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    no = w.size
    df = np.empty_like(xi)
    ff = ctypes.c_double()
    _check(lib().orc_delta_f(xi.size // no, no, _d(xi), _d(vexc), _d(w), _d(df), ctypes.c_double(rho0),
                             ctypes.c_double(prefactor), ctypes.byref(ff)))
    return df, ff.value


# ---------------------------------------------------------------------------
# Luc file reader: restates common/IO/MDFT_ReadCLuc.hpp:40-259.
_CPLX = re.compile(r"\(\s*([^,()\s]+)\s*(?:,\s*([^,()\s]+)\s*)?\)|([^\s()]+)")


def read_luc(path: str, np_luc: int, qmaxwanted: float):
    """Return (tuples (np_luc, 5), c (np_luc, nq) complex, dq, nq) as LucData
    would: dq from rows 3 and 4 (deltaAbscissa, :189-259), nq = int(qmax/dq +
    0.5) + 2 (:65), tuples from lines 12-16 (:142-159), nq rows after the 17
    header lines (:164-186), throwing if a row's q exceeds qmax + 2 dq (:184)."""
    if not os.path.exists(path):
        raise RuntimeError(f"File: {path}does not exist.")
    with open(path) as f:
        lines = f.read().split("\n")
    if len(lines) < 21:
        raise RuntimeError("Unexpected EOF while skipping lines in file: " + path)
    rows = lines[17:]
    absc = [float(r.split()[0]) for r in rows[:4]]
    dq = absc[3] - absc[2]
    nq = int(qmaxwanted / dq + 0.5) + 2
    toks = " ".join(lines[11:16]).split()
    tuples = np.zeros((np_luc, 5), dtype=np.int32)
    pos = 0
    for j in range(5):
        pos += 1  # label
        tuples[:, j] = [int(t) for t in toks[pos:pos + np_luc]]
        pos += np_luc
    c = np.zeros((np_luc, nq), dtype=np.complex128)
    for iq in range(nq):
        if iq >= len(rows) or not rows[iq].strip():
            raise RuntimeError("Luc file has fewer q rows than needed")
        parts = rows[iq].split(None, 1)
        q = float(parts[0])
        vals = []
        for mobj in _CPLX.finditer(parts[1] if len(parts) > 1 else ""):
            if mobj.group(3) is not None:
                vals.append(complex(float(mobj.group(3)), 0.0))
            else:
                vals.append(complex(float(mobj.group(1)), float(mobj.group(2) or 0.0)))
            if len(vals) == np_luc:
                break
        c[:, iq] = vals
        if q > qmaxwanted + 2.0 * dq:
            raise RuntimeError("q > qmaxwanted in file.")
    return tuples, c, dq, nq


class Conv:
    """Oracle Convolution (common/MDFT_Convolution.hpp:106-548) built from
    the Luc tuples/values; ``execute`` works in place on (np, nx, ny, nz)."""

    def __init__(self, n, length, mmax, mrso, tuples, c, dq):
        nx, ny, nz = (int(v) for v in n)
        lx, ly, lz = (float(v) for v in length)
        tuples = np.ascontiguousarray(tuples, dtype=np.int32)
        cols = [np.ascontiguousarray(tuples[:, j]) for j in range(5)]
        c = np.ascontiguousarray(c, dtype=np.complex128)
        self.n = (nx, ny, nz)
        self.mmax, self.mrso = mmax, mrso
        h = lib().orc_conv_create(nx, ny, nz, lx, ly, lz, mmax, mrso, tuples.shape[0], *[_i(x) for x in cols],
                                  c.ctypes.data_as(_dp), c.shape[1], ctypes.c_double(dq))
        if not h:
            raise RuntimeError(lib().orc_last_error().decode())
        self._h = ctypes.c_void_p(h)
        info = np.zeros(4, dtype=np.int32)
        dqo = ctypes.c_double()
        lib().orc_conv_info(self._h, _i(info), ctypes.byref(dqo))
        self.nq, self.np_new, self.npairs, self.np = (int(v) for v in info)
        self.dq = dqo.value

    def __del__(self):
        try:
            lib().orc_conv_destroy(self._h)
        except Exception:
            pass

    def table(self) -> np.ndarray:
        out = np.zeros((self.np_new, self.nq), dtype=np.complex128)
        lib().orc_conv_table(self._h, out.ctypes.data_as(_dp))
        return out

    def execute(self, p: np.ndarray) -> np.ndarray:
        y = np.ascontiguousarray(p, dtype=np.complex128).copy()
        _check(lib().orc_conv_execute(self._h, y.ctypes.data_as(_dp)))
        return y

    def energy_cproj(self, xi: np.ndarray, rho0: float, prefactor: float, stages: bool = False):
        """src/mini-mdft.hpp:137-193 on the oracle: returns (ff, df[, stage seconds])."""
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        df = np.empty_like(xi)
        ff = ctypes.c_double()
        st = np.zeros(5)
        _check(lib().orc_energy_cproj(self._h, _d(xi), ctypes.c_double(rho0), ctypes.c_double(prefactor),
                                      _d(df), ctypes.byref(ff), _d(st)))
        return (ff.value, df, st) if stages else (ff.value, df)


# ===========================================================================
# Full functional, external potential and L-BFGS (the "next" rows N1 / N4).
# The reference has no such code (src/mini-mdft.hpp:129-133 makes one
# evaluation; Solvent::m_vext is unused, common/MDFT_Solvent.hpp:160), so these
# restate the definitions of this build; their pins are a finite-difference
# gradient, the c = 0 limit xi^2 = exp(-V / kT), self-adjointness of the
# excess operator for the symmetric model c(q), monotone decrease and an
# analytic single-site potential (tests/test_functional.py).

#: Constants::qfact (common/MDFT_Constants.hpp:36-47): qfact q q / r in kJ/mol
QFACT = 1.602176565e-19 ** 2 * 1.0e-3 * 6.02214129e23 / (4.0 * np.pi * 8.854187817e-12 * 1.0e-10)


def rotation_matrices(mmax: int, mrso: int) -> np.ndarray:
    """(no, 3, 3) rotation of every orientation (common/MDFT_Grid.hpp:295-318):
    theta from the GL nodes, phi_j = j dphi and psi_k = k dpsi with dphi = 2 pi
    / nphi and dpsi = 2 pi / mrso / npsi (:141-142, uniform_mesh of
    common/MDFT_Math_Utils.hpp:76-83)."""
    s = angular_sizes(mmax, mrso)
    t = angular_tables(mmax, mrso)
    nt, nph, nps = s["ntheta"], s["nphi"], s["npsi"]
    dphi, dpsi = 2.0 * np.pi / nph, 2.0 * np.pi / mrso / nps
    th = np.repeat(t["theta"], nph * nps)
    ph = np.tile(np.repeat(np.arange(nph) * dphi, nps), nt)
    ps = np.tile(np.arange(nps) * dpsi, nt * nph)
    ct, st, cp, sp, cs, ss = np.cos(th), np.sin(th), np.cos(ph), np.sin(ph), np.cos(ps), np.sin(ps)
    return np.stack([np.stack([ct * cp * cs - sp * ss, -ct * cp * ss - sp * cs, st * cp], -1),
                     np.stack([ct * sp * cs + cp * ss, -ct * sp * ss + cp * cs, st * sp], -1),
                     np.stack([-st * cs, st * ss, ct], -1)], -2)


def vext(n, length, mmax: int, mrso: int, solute: np.ndarray, solvent: np.ndarray, vmax: float) -> np.ndarray:
    """(N, no) external potential in kJ/mol: for every grid node x_r = (i dx, j dy,
    k dz) and orientation, the sum over solvent sites (position x_r + R s) and
    solute sites of Lennard-Jones (Lorentz-Berthelot mixing, only between sites
    with sigma > 0) and qfact q q / d, with d the minimum-image distance, capped
    at vmax (and d floored at 1e-6 A).  solute, solvent: (n, 6) rows of
    (charge, sigma, epsilon, x, y, z)."""
    nx, ny, nz = n
    L = np.asarray(length, dtype=np.float64)
    R = rotation_matrices(mmax, mrso)
    grid = np.indices(n).reshape(3, -1).T * (L / np.asarray(n))
    V = np.zeros((nx * ny * nz, R.shape[0]))
    for qk, sk, ek, *sxyz in solvent:
        pos = grid[:, None, :] + np.einsum("oij,j->oi", R, np.asarray(sxyz))[None]
        for qj, sj, ej, *xj in solute:
            d = pos - np.asarray(xj)
            d -= L * np.round(d / L)
            r = np.maximum(np.sqrt((d * d).sum(-1)), 1e-6)
            if sk > 0 and sj > 0:
                sig, eps = 0.5 * (sk + sj), np.sqrt(ek * ej)
                V += 4.0 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
            V += QFACT * qk * qj / r
    return np.minimum(V, vmax)


def functional(conv: "Conv", xi: np.ndarray, V: np.ndarray, w: np.ndarray, rho0: float, prefactor: float,
               kT: float, dv: float):
    """F = F_id + F_ext + F_exc and its gradient with respect to xi (FP64):
      F_id  = kT rho0 dv sum w (xi^2 ln xi^2 - xi^2 + 1)
      F_ext = rho0 dv sum w xi^2 V
      F_exc = (dv / N) ff, ff the c-proj value of energy_cproj (the unnormalised
              FFTs make the reference's ff N times the discrete integral)
      g     = dv w (2 rho0 xi (kT ln xi^2 + V) + (2 prefactor / N) df)
    with df = 2 rho0 xi vexc the reference's gradient array (exact when the
    excess operator is self-adjoint).  Returns ((F, F_id, F_ext, F_exc), g)."""
    N = xi.shape[0]
    ff, df = conv.energy_cproj(xi, rho0, prefactor)
    x2 = np.maximum(xi * xi, 1e-300)
    fid = kT * rho0 * dv * np.sum(w * (x2 * np.log(x2) - x2 + 1.0))
    fext = rho0 * dv * np.sum(w * x2 * V)
    fexc = dv / N * ff
    g = dv * w * (2.0 * rho0 * xi * (kT * np.log(x2) + V) + (2.0 * prefactor / N) * df)
    return (fid + fext + fexc, fid, fext, fexc), g


def lbfgs(fg, x0: np.ndarray, h0: np.ndarray, m: int = 5, max_iter: int = 200, gtol: float = 1e-5,
          c1: float = 1e-4, max_backtrack: int = 30, eps_f: float = 1e-6):
    """The L-BFGS of this build (FP64): two-loop recursion with the diagonal
    initial inverse Hessian gamma h0, gamma = s'y / y'(h0 y) of the newest
    pair; a pair is kept when s'y > 1e-12 |s| |y|, and a full memory whose
    new pair fails that test restarts; a non-descent direction restarts from
    -h0 g.  Line search: backtracking from t = 1, or from t = min(1, 1 /
    max |h0 g|) along -h0 g (no stored pair), so that no xi moves by more
    than 1 (at most max_backtrack trials; after a rejected trial t the next is
    the minimiser of
    the quadratic through F(0), F'(0) = g.d and F(t), kept within [0.1 t, 0.5 t]
    (Nocedal and Wright, Numerical Optimization, 2nd ed., section 3.5), or t/2
    where F(t) is not finite) until the trial point x + t d satisfies Armijo,
    F(t) <= F + c1 t g.d, or the approximate Wolfe test of Hager and Zhang
    (SIAM J. Optim. 16 (2005) 170-192, their (23) with delta = c1):
    g(t).d <= (2 c1 - 1) g.d and F(t) <= F + eps_f S, S the scale of F.  The
    second test is Armijo read off the slopes, which stay accurate where F's
    rounding hides the decrease.  fg(x) -> (F, g) or (F, g, S); S = |F| when
    not given.  Converged when max |h0 g| < gtol.  Returns (x, F, g,
    iterations, evaluations, F history, status) with status 0 = converged,
    1 = max_iter iterations made, 2 = no acceptable point (x is the last
    accepted point), 4 = F or g not finite at x0."""
    def call(x):
        r = fg(x)
        return (r[0], r[1], abs(r[0])) if len(r) == 2 else r

    x = x0.copy()
    F, g, scale = call(x)
    S, Y = [], []
    hist, nev = [F], 1
    if not (np.isfinite(F) and np.all(np.isfinite(g))):
        return x, F, g, 0, nev, hist, 4
    for it in range(max_iter):
        if np.abs(h0 * g).max() < gtol:
            return x, F, g, it, nev, hist, 0
        q = g.copy()
        alpha = []
        for s, y in reversed(list(zip(S, Y))):
            a = (s @ q) / (y @ s)
            alpha.append(a)
            q -= a * y
        gamma = (S[-1] @ Y[-1]) / (Y[-1] @ (h0 * Y[-1])) if S else 1.0
        r = gamma * h0 * q
        for (s, y), a in zip(zip(S, Y), reversed(alpha)):
            b = (y @ r) / (y @ s)
            r += (a - b) * s
        d = -r
        gd = g @ d
        if not gd < 0:
            S.clear()
            Y.clear()
            d = -h0 * g
            gd = g @ d
        t = 1.0
        if not S:  # along -h0 g: no xi moves by more than 1
            gmax = np.abs(h0 * g).max()
            if gmax > 0:
                t = min(1.0, 1.0 / gmax)
        for _ in range(max_backtrack):
            xn = x + t * d
            Fn, gn, sn = call(xn)
            nev += 1
            if np.isfinite(Fn) and (Fn <= F + c1 * t * gd or
                                    (Fn <= F + eps_f * scale and gn @ d <= (2 * c1 - 1) * gd)):
                break
            # quadratic through F(0), F'(0) = g.d and F(t), kept in [0.1 t, 0.5 t];
            # halving where F(t) is not finite
            den = 2.0 * (Fn - F - gd * t)
            if np.isfinite(Fn) and den > 0:
                t = min(max(-gd * t * t / den, 0.1 * t), 0.5 * t)
            else:
                t = 0.5 * t
        else:
            return x, F, g, it, nev, hist, 2
        s, y = xn - x, gn - g
        if s @ y > 1e-12 * np.sqrt((s @ s) * (y @ y)):
            S.append(s)
            Y.append(y)
            if len(S) > m:
                S.pop(0)
                Y.pop(0)
        elif len(S) == m:
            # a full memory whose new pair fails the curvature test restarts
            # (the GPU update pass forms no Gram entries for the oldest pair
            # of a full memory, which a kept new pair drops)
            S.clear()
            Y.clear()
        x, F, g, scale = xn, Fn, gn, sn
        hist.append(F)
    if np.abs(h0 * g).max() < gtol:
        return x, F, g, max_iter, nev, hist, 0
    return x, F, g, max_iter, nev, hist, 1
