"""mdft_delta_f_noalloc (the get_delta_f entry point that allocates nothing)
against mdft_delta_f through the C ABI: the same kernels and summation
order, so ff and delta_f agree bit for bit, on two calls in a row (the second
smaller, reusing the library's fixed scratch array)."""
import ctypes

import numpy as np
import pytest

from mini_mdft_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MAX_POINTS = 1048568  # MDFT_DELTA_F_NOALLOC_MAX_POINTS


def _call(name, xi, vexc, w, df, rho0, pref):
    L = _lib.lib()
    no = xi.shape[-1]
    ff = ctypes.c_double()
    dtype = _lib.DOUBLE if xi.dtype == torch.float64 else _lib.FLOAT
    _lib.check(getattr(L, name)(_lib.ptr(xi), _lib.ptr(vexc), _lib.ptr(w), _lib.ptr(df), xi.numel() // no, no,
                                ctypes.c_double(rho0), ctypes.c_double(pref), ctypes.byref(ff), dtype,
                                _lib.stream_handle()))
    return ff.value


def _fields(nxyz, no, dtype, seed):
    r = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    xi = torch.from_numpy(1.0 + 0.1 * r.standard_normal((nxyz, no))).to(dev, dtype)
    vexc = torch.from_numpy(r.standard_normal((nxyz, no))).to(dev, dtype)
    w = torch.from_numpy(r.random(no)).to(dev, torch.float64)
    return xi, vexc, w


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_equals_delta_f_on_two_calls(dtype):
    assert torch.cuda.is_available()
    rho0, pref = 0.00084, -4.7
    # first call: 40^3 points x 84 orientations; second: a smaller, ragged
    # 1003 points x 330 (the last partial covers 3 points)
    for nxyz, no, seed in ((40 ** 3, 84, 1), (1003, 330, 2)):
        xi, vexc, w = _fields(nxyz, no, dtype, seed)
        df_ref = torch.empty_like(xi)
        ff_ref = _call("mdft_delta_f", xi, vexc, w, df_ref, rho0, pref)
        df = torch.full_like(xi, float("nan"))
        ff = _call("mdft_delta_f_noalloc", xi, vexc, w, df, rho0, pref)
        torch.cuda.synchronize()
        print(f"noalloc nxyz={nxyz} no={no} {dtype}: ff={ff!r} ref={ff_ref!r}", flush=True)
        assert ff == ff_ref
        assert torch.equal(df, df_ref)
        # and both against the formula (FP64 accumulation)
        x, v = xi.double().cpu().numpy(), vexc.double().cpu().numpy()
        ff_np = float(np.sum(rho0 * pref * v * w.cpu().numpy() * (x * x - 1.0)))
        assert abs(ff - ff_np) <= 1e-10 * np.abs(rho0 * pref * v * (x * x - 1.0)).sum()


def test_noalloc_largest_size_and_zero():
    """The largest accepted size (every scratch slot in use) equals mdft_delta_f;
    nxyz = 0 gives ff = 0 without a launch."""
    rho0, pref = 0.001, -2.0
    xi, vexc, w = _fields(MAX_POINTS, 1, torch.float64, 3)
    df_ref, df = torch.empty_like(xi), torch.empty_like(xi)
    ff_ref = _call("mdft_delta_f", xi, vexc, w, df_ref, rho0, pref)
    ff = _call("mdft_delta_f_noalloc", xi, vexc, w, df, rho0, pref)
    print(f"noalloc max size: ff={ff!r} ref={ff_ref!r}", flush=True)
    assert ff == ff_ref and torch.equal(df, df_ref)
    e = torch.empty((0, 5), dtype=torch.float64, device="cuda:0")
    w5 = torch.ones(5, dtype=torch.float64, device="cuda:0")
    assert _call("mdft_delta_f_noalloc", e, e, w5, e, rho0, pref) == 0.0
