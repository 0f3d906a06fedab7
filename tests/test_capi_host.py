"""CPU checks of libmdft_hip.so: it loads, exports every declared symbol, and its
host-side (FP64) tables and Luc reader agree with the oracle.  No GPU calls."""
import ctypes
import math
import os

import numpy as np
import pytest

from mini_mdft_amd import _lib, luc
from mini_mdft_amd import mdft as M


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _lib.declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib._SIGS), set(names) ^ set(_lib._SIGS)
    assert L.mdft_version() >= 1


def test_gauss_legendre_and_wigner_match_oracle(oracle):
    for n in range(1, 12):
        x, w = np.zeros(n), np.zeros(n)
        _lib.check(_lib.lib().mdft_gauss_legendre(n, x.ctypes.data_as(_lib._pd), w.ctypes.data_as(_lib._pd)))
        xo, wo = oracle.gauss_legendre(n)
        assert np.abs(x - xo).max() < 1e-15 and np.abs(w - wo).max() < 1e-15
    out = ctypes.c_double()
    for m in range(0, 8):
        for a in range(-m, m + 1):
            for b in range(-m, m + 1):
                for t in (0.1, 1.3, 2.9):
                    _lib.check(_lib.lib().mdft_wigner_small_d(m, a, b, ctypes.c_double(t), ctypes.byref(out)))
                    assert abs(out.value - oracle.wigner_small_d(m, a, b, t)) < 1e-14


@pytest.mark.parametrize("mmax,mrso", [(1, 1), (3, 2), (5, 2), (10, 10), (6, 4)])
def test_angular_tables_match_oracle(oracle, mmax, mrso):
    ag = M.AngularGrid(mmax, mrso)
    T = oracle.angular_tables(mmax, mrso)
    assert (ag.m_ntheta, ag.m_nphi, ag.m_npsi, ag.m_no, ag.m_np) == (T["ntheta"], T["nphi"], T["npsi"], T["no"],
                                                                      T["np"])
    assert np.abs(ag.m_thetaofntheta - T["theta"]).max() < 1e-15
    assert np.abs(ag.m_wthetaofntheta - T["wtheta"]).max() < 1e-15
    assert np.abs(ag.m_w - T["w"]).max() < 1e-14
    mp = M.OrientationProjectionMap(ag)
    assert np.array_equal(mp.p(), T["pmap"])
    for a, b in ((mp.m(), T["m"]), (mp.mup(), T["mup"]), (mp.mu(), T["mu"]), (mp.mu2(), T["mu2"])):
        assert np.array_equal(a, b)
    assert np.abs(mp.wigner_small_d() - T["D"]).max() < 1e-14


def test_angular_grid_errors():
    with pytest.raises(_lib.MDFTError, match="mmax must be greater than zero"):
        M.AngularGrid(0, 2)
    with pytest.raises(_lib.MDFTError, match="molrotsymorder"):
        M.AngularGrid(3, 0)
    with pytest.raises(_lib.MDFTError, match="negative length"):
        M.SpatialGrid((4, 4, 4), (1.0, 1.0, -1.0))
    with pytest.raises(_lib.MDFTError, match="grid nodes"):
        M.SpatialGrid((4, 0, 4), (1.0, 1.0, 1.0))


def test_spatial_grid_matches_oracle(oracle):
    g = M.SpatialGrid((10, 7, 64), (1.5, 2.0, 30.0))
    for k, n, L in ((g.m_kx, 10, 1.5), (g.m_ky, 7, 2.0), (g.m_kz, 64, 30.0)):
        assert np.array_equal(k, oracle.kproj(n, L))
    assert g.m_v == 1.5 * 2.0 * 30.0


@pytest.mark.parametrize("q", [(0.0, 0.0, 0.0), (1.0, 2.0, 3.0), (0.20943951, 0.20943951, 0.20943951),
                               (0.0, 0.0, 2.0), (0.0, 0.0, -1.0), (-0.4, 0.3, 0.0)])
def test_rotation_matrix_matches_oracle(oracle, q):
    for mmax in range(0, 7):
        R = np.zeros((mmax + 1, 2 * mmax + 1, 2 * mmax + 1), dtype=np.complex128)
        qq = np.asarray(q, dtype=np.float64)
        _lib.check(_lib.lib().mdft_rotation_matrix(mmax, qq.ctypes.data_as(_lib._pd), R.ctypes.data_as(_lib._pd)))
        Ro = oracle.rotation_matrix(mmax, q)
        assert np.abs(R - Ro).max() < 1e-14


@pytest.mark.parametrize("n", [(5, 5, 5), (10, 10, 10), (4, 6, 5), (64, 64, 64), (1, 2, 3)])
def test_qpairs_match_oracle(oracle, n):
    L = _lib.lib().mdft_qpair_count(*n)
    out = np.zeros((L, 3), dtype=np.int32)
    _lib.check(_lib.lib().mdft_qpairs(*n, out.ctypes.data_as(_lib._pi)))
    assert np.array_equal(out, oracle.qpairs(*n))
    if n == (64, 64, 64):
        assert L == 131076


def _read(path, np_luc, qmax):
    nq, dq = ctypes.c_int(), ctypes.c_double()
    _lib.check(_lib.lib().mdft_luc_read(path.encode(), np_luc, ctypes.c_double(qmax), ctypes.byref(nq),
                                        ctypes.byref(dq), None, None))
    t = np.zeros((np_luc, 5), dtype=np.int32)
    c = np.zeros((np_luc, nq.value), dtype=np.complex128)
    _lib.check(_lib.lib().mdft_luc_read(path.encode(), np_luc, ctypes.c_double(qmax), None, None,
                                        t.ctypes.data_as(_lib._pi), c.ctypes.data_as(_lib._pd)))
    return t, c, dq.value, nq.value


def test_luc_reader_matches_oracle(oracle, tmp_path):
    tab = luc.synthetic_c(3, 2, 40, seed=3)
    path = os.path.join(tmp_path, "c.luc")
    luc.write_luc(path, tab)
    qmax = 30 * luc.DQ_TIP3P
    t, c, dq, nq = _read(path, 252, qmax)
    to, co, dqo, nqo = oracle.read_luc(path, 252, qmax)
    assert np.array_equal(t, to) and np.array_equal(c, co) and dq == dqo and nq == nqo
    with pytest.raises(_lib.MDFTError, match="does not exist"):
        _read(path + "wrong", 252, qmax)
    with pytest.raises(_lib.MDFTError, match="fewer q rows"):
        _read(path, 252, 60 * luc.DQ_TIP3P)


def test_solvent_constants():
    # common/MDFT_Thermo.hpp:60 and common/MDFT_Solvent.hpp:366-368, 429-431
    assert abs(M.kT(298.15) - 2.4790) < 1e-4  # RT at 298.15 K in kJ/mol
    n0 = M.SOLVENTS["spce"]["n0"]
    assert abs(M.rho0_of(n0, 2) - n0 * 2 / (8 * math.pi ** 2)) < 1e-18
    assert abs(M.prefactor_of(298.15, 2, n0) - (-2 * M.kT(298.15) * 4 * math.pi ** 2 / 2 / n0)) < 1e-9


def test_delta_f_noalloc_rejects_before_any_device_call():
    """Sizes above the scratch limit and bad arguments fail with a clear error,
    checked on the host before the first HIP call (runs without a GPU)."""
    L = _lib.lib()
    ff = ctypes.c_double()
    big = 1048568 + 1  # MDFT_DELTA_F_NOALLOC_MAX_POINTS + 1
    one = ctypes.c_void_p(8)  # never dereferenced: the size check comes first
    rc = L.mdft_delta_f_noalloc(one, one, one, one, big, 5, 0.1, -1.0, ctypes.byref(ff), _lib.DOUBLE, None)
    assert rc != 0 and "exceeds the scratch limit of 1048568" in L.mdft_last_error().decode()
    rc = L.mdft_delta_f_noalloc(one, one, one, one, 10, 0, 0.1, -1.0, ctypes.byref(ff), _lib.DOUBLE, None)
    assert rc != 0 and "bad arguments" in L.mdft_last_error().decode()
    rc = L.mdft_delta_f_noalloc(None, one, one, one, 10, 5, 0.1, -1.0, ctypes.byref(ff), _lib.DOUBLE, None)
    assert rc != 0 and "null argument" in L.mdft_last_error().decode()
    rc = L.mdft_delta_f_noalloc(one, one, one, one, 10, 5, 0.1, -1.0, ctypes.byref(ff), 7, None)
    assert rc != 0 and "element type" in L.mdft_last_error().decode()
    with open(_lib.HEADER) as f:
        assert "#define MDFT_DELTA_F_NOALLOC_MAX_POINTS 1048568" in f.read()
