"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py from the
oracle): the oracle must keep reproducing them, the host tables of the library
must match them, and on the GPU the FP64 path must match them to the parity
bar."""
import os

import numpy as np
import pytest

from mini_mdft_amd import FP64, luc, rng
from mini_mdft_amd import _lib

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


# ---- the oracle against its own fixtures (CPU) -----------------------------
def test_oracle_tables_reproduce(oracle):
    g = _load("tables")
    x, w = oracle.gauss_legendre(6)
    assert rel(x, g["gl_x"]) <= 1e-15 and rel(w, g["gl_w"]) <= 1e-15
    t = oracle.angular_tables(5, 2)
    assert rel(t["D"], g["wigner"]) <= 1e-15 and np.array_equal(t["pmap"], g["pmap"])
    for q, R in zip(g["q"], g["R"]):
        assert rel(oracle.rotation_matrix(5, q), R) <= 1e-15


def test_oracle_transforms_reproduce(oracle):
    g = _load("proj")
    assert rel(oracle.angl2proj(g["o"], (4, 4, 4), 3, 2), g["p_of_o"]) <= 1e-13
    assert rel(oracle.proj2angl(g["p"], 3, 2), g["o_of_p"]) <= 1e-13


def _conv_inputs(tag):
    n, L = ((4, 4, 4), (5.0,) * 3) if tag == "4" else ((5, 5, 5), (6.0,) * 3)
    tab = luc.synthetic_c(3, 2, luc.nq_for(luc.qmax_of_grid(n, L), luc.DQ_TIP3P), seed=7)
    return n, L, tab


@pytest.mark.parametrize("tag", ["4", "5"])
def test_oracle_conv_reproduces(oracle, tag):
    g = _load("conv")
    n, L, tab = _conv_inputs(tag)
    C = oracle.Conv(n, L, 3, 2, tab.tuples, tab.c, tab.dq)
    assert rel(C.execute(g[f"p_in_{tag}"]), g[f"p_out_{tag}"]) <= 1e-13


def _energy_inputs():
    n, L = (8, 8, 8), (10.0,) * 3
    tab = luc.synthetic_c(3, 2, luc.nq_for(luc.qmax_of_grid(n, L), luc.DQ_TIP3P), seed=11)
    xi = rng.xi_field(12345, (512, 84), 0.1)
    return n, L, tab, xi


def test_oracle_energy_reproduces(oracle):
    g = _load("energy")
    n, L, tab, xi = _energy_inputs()
    C = oracle.Conv(n, L, 3, 2, tab.tuples, tab.c, tab.dq)
    ff, df = C.energy_cproj(xi, float(g["rho0"]), float(g["prefactor"]))
    assert abs(ff - float(g["ff"])) <= 1e-13 * abs(float(g["ff"]))
    assert rel(df, g["df"]) <= 1e-13


# ---- the library's host tables against the fixtures (CPU, C ABI) -----------
def test_capi_tables_match_golden():
    import ctypes

    g = _load("tables")
    L = _lib.lib()
    x, w = np.zeros(6), np.zeros(6)
    assert L.mdft_gauss_legendre(6, x.ctypes.data_as(_lib._pd), w.ctypes.data_as(_lib._pd)) == 0
    assert rel(x, g["gl_x"]) <= 1e-14 and rel(w, g["gl_w"]) <= 1e-14
    for q, Rg in zip(g["q"], g["R"]):
        R = np.zeros((6, 11, 11), dtype=np.complex128)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        assert L.mdft_rotation_matrix(5, qq.ctypes.data_as(_lib._pd),
                                      R.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
        # entries with |m| > l are not part of the result
        for l in range(6):
            blk = np.s_[l, 5 - l:6 + l, 5 - l:6 + l]
            assert np.abs(R[blk] - Rg[blk]).max() <= 1e-14


# ---- the GPU path against the fixtures ---------------------------------------
@pytest.mark.gpu
def test_gpu_matches_golden():
    import torch

    from mini_mdft_amd import mdft as M

    dev = torch.device("cuda:0")
    # projections, 4^3, mmax 3
    g = _load("proj")
    sg = M.SpatialGrid((4, 4, 4), (1.0, 1.0, 1.0))
    ag = M.AngularGrid(3, 2)
    opt = M.OrientationProjectionTransform(sg, ag, M.OrientationProjectionMap(ag), FP64)
    p = torch.zeros((ag.m_np, 4, 4, 4), dtype=torch.complex128, device=dev)
    opt.angl2proj(torch.from_numpy(g["o"]).to(dev), p)
    assert rel(p.cpu().numpy(), g["p_of_o"]) <= 1e-13
    o = torch.zeros(g["o_of_p"].shape, dtype=torch.float64, device=dev)
    opt.proj2angl(torch.from_numpy(g["p"]).to(dev), o)
    assert rel(o.cpu().numpy(), g["o_of_p"]) <= 1e-13
    # convolution, 4^3 and 5^3
    gc = _load("conv")
    for tag in ("4", "5"):
        n, L, tab = _conv_inputs(tag)
        sg = M.SpatialGrid(n, L)
        conv = M.Convolution.from_table(tab, sg, ag, M.OrientationProjectionMap(ag), FP64)
        t = torch.from_numpy(gc[f"p_in_{tag}"].copy()).to(dev)
        conv.execute(t)
        assert rel(t.cpu().numpy(), gc[f"p_out_{tag}"]) <= 1e-11
    # one evaluation, 8^3
    ge = _load("energy")
    n, L, tab, xi = _energy_inputs()
    sg = M.SpatialGrid(n, L)
    conv = M.Convolution.from_table(tab, sg, ag, M.OrientationProjectionMap(ag), FP64)
    xt = torch.from_numpy(xi).to(dev)
    df = torch.empty_like(xt)
    ff = M.Solver(conv, float(ge["rho0"]), float(ge["prefactor"]), FP64).energy_cproj(xt, df)
    assert abs(ff - float(ge["ff"])) <= 1e-10 * abs(float(ge["ff"]))
    assert rel(df.cpu().numpy(), ge["df"]) <= 1e-10
