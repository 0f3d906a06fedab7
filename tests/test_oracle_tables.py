"""Pins for the oracle's init-time tables (CPU only)."""
import math

import numpy as np
import pytest
from scipy.special import eval_legendre, lpmv


def test_gauss_legendre_vs_numpy(oracle):
    # common/MDFT_Math_Utils.hpp:24-73: nodes descending, weights sum to 1
    for n in range(1, 13):
        x, w = oracle.gauss_legendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.abs(x - xr[::-1]).max() < 1e-14
        assert np.abs(w - wr[::-1] / 2).max() < 1e-14
        assert abs(w.sum() - 1.0) < 100 * np.finfo(float).eps  # Test_Math_Utils.cpp:32-44


def _d_closed(m, a, b, t):
    """Closed forms of d^m_{a,b}(t) (PDG table of d-functions)."""
    c, s = math.cos(t), math.sin(t)
    table = {
        (1, 1, 1): (1 + c) / 2, (1, 1, 0): -s / math.sqrt(2), (1, 1, -1): (1 - c) / 2, (1, 0, 0): c,
        (2, 2, 2): ((1 + c) / 2) ** 2, (2, 2, 1): -(1 + c) / 2 * s, (2, 2, 0): math.sqrt(6) / 4 * s * s,
        (2, 2, -1): -(1 - c) / 2 * s, (2, 2, -2): ((1 - c) / 2) ** 2, (2, 1, 1): (1 + c) / 2 * (2 * c - 1),
        (2, 1, 0): -math.sqrt(1.5) * s * c, (2, 1, -1): (1 - c) / 2 * (2 * c + 1), (2, 0, 0): 1.5 * c * c - 0.5,
    }
    if (m, a, b) in table:
        return table[(m, a, b)]
    if (m, -b, -a) in table:  # d_{ab} = d_{-b,-a}
        return table[(m, -b, -a)]
    if (m, b, a) in table:  # d_{ab} = (-1)^{a-b} d_{ba}
        return (-1) ** (a - b) * table[(m, b, a)]
    if (m, -a, -b) in table:  # d_{ab} = (-1)^{a-b} d_{-a,-b}
        return (-1) ** (a - b) * table[(m, -a, -b)]
    raise KeyError


def test_wigner_closed_forms(oracle):
    # common/MDFT_Wigner.hpp:62-130 computes d^m_{mu,mup} (PDG convention)
    for t in (0.0, 0.37, 1.2, 2.9, math.pi):
        for m in (1, 2):
            for a in range(-m, m + 1):
                for b in range(-m, m + 1):
                    assert abs(oracle.wigner_small_d(m, a, b, t) - _d_closed(m, a, b, t)) < 1e-14


def test_wigner_legendre_columns(oracle):
    # d^m_{00} = P_m(cos t); d^m_{a0} = sqrt((m-a)!/(m+a)!) P_m^a(cos t), P_m^a with the
    # Condon-Shortley phase that scipy's lpmv includes
    for t in (0.3, 1.1, 2.5):
        for m in range(0, 9):
            assert abs(oracle.wigner_small_d(m, 0, 0, t) - eval_legendre(m, math.cos(t))) < 1e-13
            for a in range(1, m + 1):
                ref = math.sqrt(math.factorial(m - a) / math.factorial(m + a)) * lpmv(a, m, math.cos(t))
                assert abs(oracle.wigner_small_d(m, a, 0, t) - ref) < 1e-13


def test_wigner_quick_return_and_symmetry(oracle):
    assert oracle.wigner_small_d(2, 3, 0, 0.4) == 0.0  # |mu| > m
    assert oracle.wigner_small_d(0, 0, 0, 0.4) == 1.0
    for m in range(1, 8):
        for a in range(-m, m + 1):
            for b in range(-m, m + 1):
                t = 0.77
                d = oracle.wigner_small_d(m, a, b, t)
                assert abs(d - oracle.wigner_small_d(m, -b, -a, t)) < 1e-13
                assert abs(d - (-1) ** (a - b) * oracle.wigner_small_d(m, b, a, t)) < 1e-13
                assert abs(d - (-1) ** (a - b) * oracle.wigner_small_d(m, a, b, -t)) < 1e-13


def test_wigner_gl_orthogonality(oracle):
    # sum_i w_i d^m_{ab}(t_i) d^n_{ab}(t_i) = delta_mn / (2m+1): the GL rule with mmax+1 nodes
    # is exact for these degree-(m+n) polynomials in cos t.
    mmax = 6
    T = oracle.angular_tables(mmax, 1)
    for a in range(-mmax, mmax + 1):
        for b in range(-mmax, mmax + 1):
            for m in range(max(abs(a), abs(b)), mmax + 1):
                for n in range(max(abs(a), abs(b)), mmax + 1):
                    s = sum(T["wtheta"][i] * oracle.wigner_small_d(m, a, b, T["theta"][i])
                            * oracle.wigner_small_d(n, a, b, T["theta"][i]) for i in range(mmax + 1))
                    assert abs(s - (1.0 / (2 * m + 1) if m == n else 0.0)) < 1e-13


@pytest.mark.parametrize("mmax,mrso", [(1, 1), (3, 2), (5, 2), (10, 10), (4, 3)])
def test_angular_grid(oracle, mmax, mrso):
    T = oracle.angular_tables(mmax, mrso)
    # common/MDFT_Grid.hpp:135-152
    assert T["ntheta"] == mmax + 1 and T["nphi"] == 2 * mmax + 1 and T["npsi"] == 2 * (mmax // mrso) + 1
    assert T["no"] == T["ntheta"] * T["nphi"] * T["npsi"]
    assert T["np"] == sum(1 for m in range(mmax + 1) for _ in range(-m, m + 1) for _ in range(0, m // mrso + 1))
    # Sum w = 8 pi^2 / mrso and the io order theta-phi-psi (:225-263)
    assert abs(T["w"].sum() - 8 * math.pi ** 2 / mrso) < 1e-12
    w = T["w"].reshape(T["ntheta"], T["nphi"], T["npsi"])
    ref = T["wtheta"][:, None, None] / T["nphi"] / (T["npsi"] * mrso) * 8 * math.pi ** 2
    assert np.abs(w - np.broadcast_to(ref, w.shape)).max() < 1e-15
    # index map (...Transform.hpp:76-91) and D(itheta, ip) = d^m_{mup,mu}(theta) (:143-144)
    ip = 0
    for m in range(mmax + 1):
        for mup in range(-m, m + 1):
            for mu in range(0, m + 1, mrso):
                assert T["pmap"][m, mup + mmax, mu // mrso] == ip
                assert (T["m"][ip], T["mup"][ip], T["mu"][ip], T["mu2"][ip]) == (m, mup, mu, mu // mrso)
                for it in range(T["ntheta"]):
                    assert T["D"][it, ip] == oracle.wigner_small_d(m, mup, mu, T["theta"][it])
                ip += 1


def test_angular_grid_errors(oracle):
    with pytest.raises(RuntimeError):
        oracle.angular_sizes(0, 2)
    with pytest.raises(RuntimeError):
        oracle.angular_sizes(3, 0)


@pytest.mark.parametrize("n", [1, 2, 5, 10, 64])
def test_kproj(oracle, n):
    L = 1.5
    k = oracle.kproj(n, L)
    # numpy's fftfreq except that the Nyquist entry of an even n is +n/2 (Grid.hpp:87-103)
    ref = np.fft.fftfreq(n, d=L / n) * 2 * math.pi
    if n % 2 == 0:
        ref[n // 2] = -ref[n // 2]
    assert np.abs(k - ref).max() < 1e-12
    # Hermitian symmetry (unit_test/Test_Grid.cpp:122-172)
    for i in range(n):
        j = 0 if i == 0 else n - i
        assert abs(k[j] - (k[i] if i == n // 2 and n % 2 == 0 else -k[i])) < 1e-12


def _wigner_D_closed(oracle, mmax, q):
    q = np.asarray(q, dtype=float)
    nq = np.linalg.norm(q)
    th = 0.0 if nq == 0 else math.acos(q[2] / nq)
    ph = math.atan2(q[1], q[0]) if (q[0] != 0 or q[1] != 0) else 0.0
    R = np.zeros((mmax + 1, 2 * mmax + 1, 2 * mmax + 1), dtype=complex)
    for l in range(mmax + 1):
        for m in range(-l, l + 1):
            for m1 in range(-l, l + 1):
                R[l, m + mmax, m1 + mmax] = np.exp(-1j * m * ph) * oracle.wigner_small_d(l, m, m1, th)
    return R


KX0 = 0.209439510239320  # unit_test/Test_Rotation.cpp:20-23


@pytest.mark.parametrize("q", [(0.0, 0.0, 0.0), (1.0, 2.0, 3.0), (KX0, KX0, KX0), (0.0, 0.0, 2.0),
                               (-0.3, 0.1, -0.7), (0.0, 1.3, 0.0)])
def test_rotation_matrix_closed_form(oracle, q):
    # R^l_{m,m1}(q) of common/MDFT_Rotation.hpp:103-283 is the Wigner D-matrix
    # e^{-i m phi_q} d^l_{m,m1}(theta_q) of the q-frame (z' = q, y' = z x q).
    for mmax in range(0, 6):
        R = oracle.rotation_matrix(mmax, q)
        if mmax == 0:
            assert R[0, 0, 0] == 1.0
            continue
        Rc = _wigner_D_closed(oracle, mmax, q)
        assert np.abs(R - Rc).max() < 1e-13
        for l in range(mmax + 1):  # each l block is unitary
            B = R[l, mmax - l:mmax + l + 1, mmax - l:mmax + l + 1]
            assert np.abs(B @ B.conj().T - np.eye(2 * l + 1)).max() < 1e-13


def test_rotation_matrix_q_along_minus_z(oracle):
    # for q parallel to -z the frame is (x, -y, -z): a rotation by pi about x, i.e. phi = pi
    mmax = 4
    R = oracle.rotation_matrix(mmax, (0.0, 0.0, -1.5))
    for l in range(mmax + 1):
        for m in range(-l, l + 1):
            for m1 in range(-l, l + 1):
                ref = np.exp(-1j * m * math.pi) * oracle.wigner_small_d(l, m, m1, math.pi)
                assert abs(R[l, m + mmax, m1 + mmax] - ref) < 1e-13


@pytest.mark.parametrize("n", [(5, 5, 5), (10, 10, 10), (4, 6, 5), (1, 1, 1), (2, 3, 4)])
def test_qpairs_cover_each_point_once(oracle, n):
    # unit_test/Test_Convolution.cpp:54-77: every (q, -q) is visited once and only once
    nx, ny, nz = n
    pairs = oracle.qpairs(nx, ny, nz)
    seen = np.zeros(n, dtype=int)
    for ix, iy, iz in pairs:
        assert 0 <= iz <= nz // 2
        im = ((-ix) % nx, (-iy) % ny, (-iz) % nz)
        assert seen[ix, iy, iz] == 0 and seen[im] == 0
        seen[ix, iy, iz] = 1
        seen[im] = 1
    assert seen.all()
