"""Pins for the oracle's FFT and orientation <-> projection transforms (CPU only)."""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 16, 30, 32, 48, 64, 80, 96, 13])
def test_fft1d_vs_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    # forward: e^{-2 pi i jk/n}; backward: e^{+2 pi i jk/n}; both unnormalised
    assert np.abs(oracle.fft1d(x, -1) - np.fft.fft(x)).max() < 1e-13 * max(n, 1)
    assert np.abs(oracle.fft1d(x, +1) - np.fft.ifft(x) * n).max() < 1e-13 * max(n, 1)


def test_fft1d_brute_force(oracle):
    for n in (2, 3, 5, 6, 9):
        x = np.arange(n) + 1j * np.arange(n)[::-1]
        j = np.arange(n)
        M = np.exp(-2j * np.pi * np.outer(j, j) / n)
        assert np.abs(oracle.fft1d(x, -1) - M @ x).max() < 1e-12


@pytest.mark.parametrize("shape", [(2, 3, 4, 5), (1, 6, 6, 6), (3, 4, 4, 4), (2, 5, 1, 3)])
def test_fft3d_vs_numpy(oracle, shape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    N = np.prod(shape[-3:])
    assert np.abs(oracle.fft3d(x, -1) - np.fft.fftn(x, axes=(-3, -2, -1))).max() < 1e-12 * N
    # round trip gives N x (Normalization::none both ways)
    assert np.abs(oracle.fft3d(oracle.fft3d(x, -1), +1) - N * x).max() < 1e-12 * N


def _numpy_angl2proj(T, o, mrso):
    """unit_test/Test_OrientationProjectionTransform.cpp:142-181 with numpy's rfft2."""
    nph = T["nphi"]
    c2 = np.fft.rfft2(o, axes=(-2, -1), norm="forward")
    N = o.shape[0]
    pr = np.zeros((T["np"], N), complex)
    for ip in range(T["np"]):
        m, mup, mu2 = T["m"][ip], T["mup"][ip], T["mu2"][ip]
        kphi = mup if mup >= 0 else mup + nph
        pr[ip] = (np.conj(c2[:, :, kphi, mu2]) * T["D"][:, ip] * T["wtheta"] * T["fm"][m]).sum(1)
    return pr


def _dense_matrix(T, mrso):
    """The projection as one explicit-phase complex matrix A[theta,phi,psi,ip] =
    f_m w_theta d^m_{mup,mu}(theta) e^{i(mup phi_j + mu psi_k)} / (nphi npsi)."""
    nt, nph, nps = T["ntheta"], T["nphi"], T["npsi"]
    phi = 2 * np.pi * np.arange(nph) / nph
    psi = 2 * np.pi * np.arange(nps) / (mrso * nps)
    A = np.zeros((nt, nph, nps, T["np"]), complex)
    for ip in range(T["np"]):
        m, mup, mu = T["m"][ip], T["mup"][ip], T["mu"][ip]
        amp = T["fm"][m] * T["wtheta"] * T["D"][:, ip]
        A[:, :, :, ip] = amp[:, None, None] * np.exp(1j * (mup * phi[:, None] + mu * psi[None, :]))[None] / (nph * nps)
    return A


@pytest.mark.parametrize("mmax,mrso,n", [(3, 2, (2, 3, 4)), (5, 2, (2, 2, 3)), (4, 1, (3, 1, 2)), (10, 10, (2, 2, 2))])
def test_angl2proj(oracle, mmax, mrso, n):
    T = oracle.angular_tables(mmax, mrso)
    N = int(np.prod(n))
    rng = np.random.default_rng(3)
    o = rng.standard_normal((N, T["ntheta"], T["nphi"], T["npsi"]))
    p = oracle.angl2proj(o, n, mmax, mrso).reshape(T["np"], N)
    assert np.abs(p - _numpy_angl2proj(T, o, mrso)).max() < 1e-14
    pd = np.einsum("rabc,abcp->pr", o, _dense_matrix(T, mrso))
    assert np.abs(p - pd).max() < 1e-14
    # a constant field has only the m = 0 projection, equal to the constant
    pc = oracle.angl2proj(np.full_like(o, 2.5), n, mmax, mrso).reshape(T["np"], N)
    assert np.abs(pc[0] - 2.5).max() < 1e-14 and np.abs(pc[1:]).max() < 1e-14


@pytest.mark.parametrize("mmax,mrso", [(3, 2), (5, 2), (4, 1), (6, 3), (10, 10)])
def test_angl2proj_real_field_symmetry(oracle, mmax, mrso):
    """A real field's mu = 0 projections pair up: R^m_{-mup,-mu} = (-1)^(mup-mu)
    conj(R^m_{mup,mu}) gives p^m_{-mup,0} = (-1)^mup conj(p^m_{mup,0}) (at mmax 5,
    mrso 2, 15 of the 88 channels).  A dropped sign, a swapped mup or a wrong
    conjugation in the phi quadrature breaks it; the identity without the sign
    must fail, so the check discriminates."""
    T = oracle.angular_tables(mmax, mrso)
    n = (2, 3, 2)
    N = int(np.prod(n))
    o = np.random.default_rng(11).uniform(-1, 1, (N, T["ntheta"], T["nphi"], T["npsi"]))
    p = oracle.angl2proj(o, n, mmax, mrso).reshape(T["np"], N)
    idx = {(int(a), int(b), int(c)): i for i, (a, b, c) in enumerate(zip(T["m"], T["mup"], T["mu"]))}
    pairs = [(i, idx[(m, -mup, 0)], mup) for (m, mup, mu), i in idx.items() if mu == 0 and mup < 0]
    assert len(pairs) == mmax * (mmax + 1) // 2
    scale = np.abs(p).max()
    for i, j, mup in pairs:
        assert np.abs(p[i] - (-1) ** mup * np.conj(p[j])).max() < 1e-14 * scale
    assert max(np.abs(p[i] - np.conj(p[j])).max() for i, j, mup in pairs if mup % 2) > 1e-3 * scale


@pytest.mark.parametrize("mmax,mrso,n", [(3, 2, (2, 3, 4)), (5, 2, (2, 2, 3)), (4, 1, (3, 1, 2)), (6, 4, (2, 1, 2))])
def test_proj2angl(oracle, mmax, mrso, n):
    T = oracle.angular_tables(mmax, mrso)
    N = int(np.prod(n))
    nt, nph, nps = T["ntheta"], T["nphi"], T["npsi"]
    rng = np.random.default_rng(4)
    # random complex projections: the mu2 = 0 column need not be Hermitian, the
    # c2r convention then decides (numpy irfft2 = FFTW: Re of the DC psi bin)
    pp = rng.standard_normal((T["np"],) + n) + 1j * rng.standard_normal((T["np"],) + n)
    oo = oracle.proj2angl(pp, mmax, mrso)
    P = pp.reshape(T["np"], N)
    nk = nps // 2 + 1
    S = np.zeros((N, nt, nph, nk), complex)
    for imup in range(2 * mmax + 1):
        mup = imup - mmax
        kphi = mup if mup >= 0 else mup + nph
        for mu2 in range(nk):
            s = 0
            for m in range(max(abs(mup), mrso * mu2), mmax + 1):
                ip = T["pmap"][m, imup, mu2]
                s = s + P[ip][:, None] * T["D"][None, :, ip] * T["fm"][m]
            S[:, :, kphi, mu2] = np.conj(s)
    ref = np.fft.irfft2(S, s=(nph, nps), axes=(-2, -1)) * (nph * nps)
    assert np.abs(oo - ref).max() < 1e-12


@pytest.mark.parametrize("mmax,mrso", [(3, 2), (5, 2), (4, 1), (10, 10)])
def test_projection_round_trips(oracle, mmax, mrso):
    T = oracle.angular_tables(mmax, mrso)
    n = (2, 2, 2)
    rng = np.random.default_rng(5)
    o = rng.standard_normal((8, T["ntheta"], T["nphi"], T["npsi"]))
    p1 = oracle.angl2proj(o, n, mmax, mrso)
    p2 = oracle.angl2proj(oracle.proj2angl(p1, mmax, mrso), n, mmax, mrso)
    # projections of a band-limited field are reproduced exactly (GL exactness)
    assert np.abs(p1 - p2).max() < 1e-13 * max(1.0, np.abs(p1).max())
    # unit_test/Test_OrientationProjectionTransform.cpp:313-348: constant field identity
    oc = np.ones_like(o)
    assert np.abs(oracle.proj2angl(oracle.angl2proj(oc, n, mmax, mrso), mmax, mrso) - 1.0).max() < 1e-13
