"""Python mirror of the reference's hot-path classes over the C ABI.

Same names, argument meaning and error behaviour as the reference's C++ API
(common/MDFT_Grid.hpp, common/MDFT_OrientationProjectionTransform.hpp,
common/MDFT_Convolution.hpp, common/MDFT_Solvent.hpp, src/mini-mdft.hpp), with
torch tensors on the GPU standing in for Kokkos views (LayoutRight, contiguous).
Every compute call goes to libmdft_hip.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import DOUBLE, FLOAT, FP32, FP64, MIXED, MIXED_X6, MIXED_X6F, MDFTError, check, lib, ptr, stream_handle

# common/MDFT_Constants.hpp:25-56
BOLTZ = 1.38064852e-23
NAVO = 6.02214129e23
PI = math.pi


def kT(temperature: float) -> float:
    """Thermo::m_kbt = Boltz Navo T 1e-3 kJ/mol (common/MDFT_Thermo.hpp:60)."""
    return BOLTZ * NAVO * temperature * 1e-3


def rho0_of(n0: float, mrso: int) -> float:
    """n0 / (8 pi^2 / mrso) (common/MDFT_Solvent.hpp:430-431, :484-485)."""
    return n0 / (8 * PI * PI / mrso)


def prefactor_of(temperature: float, mrso: int, n0: float) -> float:
    """-2 kT 4 pi^2 / mrso / n0 (common/MDFT_Solvent.hpp:366-368)."""
    return -2 * kT(temperature) * 4 * PI * PI / mrso / n0


#: n0 (molecules / A^3) and npluc of the reference's solvents (common/MDFT_Solvent.hpp:417-532)
SOLVENTS = {
    "spce": dict(n0=0.0332891, mrso=2, npluc=(1, 6, 75, 252, 877, 2002)),
    "spce-h": dict(n0=0.0332891, mrso=2, npluc=(1, 6, 75, 252, 877, 2002)),
    "spce-m": dict(n0=0.0332891, mrso=2, npluc=(1, 6, 75, 252, 877, 2002)),
    "tip3p": dict(n0=0.03349459, mrso=2, npluc=(1, 6, 75, 252, 877, 2002)),
    "tip3p-m": dict(n0=0.03349459, mrso=2, npluc=(1, 6, 75, 252, 877, 2002)),
}


def _torch():
    import torch

    return torch


def _dtypes(precision):
    torch = _torch()
    if precision == FP64:
        return torch.float64, torch.complex128
    if precision in (FP32, MIXED_X6F):
        return torch.float32, torch.complex64
    if precision in (MIXED, MIXED_X6):
        return torch.float64, torch.complex64
    raise MDFTError("unknown precision form")


def _require(t, dtype, numel, what):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise MDFTError(f"{what} must be a CUDA (HIP) tensor")
    if t.dtype != dtype:
        raise MDFTError(f"{what} must have dtype {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise MDFTError(f"{what} must be contiguous (LayoutRight)")
    if t.numel() != numel:
        raise MDFTError(f"{what} has {t.numel()} elements, expected {numel}")


class SpatialGrid:
    """common/MDFT_Grid.hpp:23-104."""

    def __init__(self, n_nodes, length):
        if any(float(v) <= 0 for v in length):
            raise MDFTError("The supercell cannot have negative length.")
        if any(int(v) <= 0 for v in n_nodes):
            raise MDFTError("The space is divided into grid nodes.")
        self.m_n_nodes = tuple(int(v) for v in n_nodes)
        self.m_length = tuple(float(v) for v in length)
        self.m_nx, self.m_ny, self.m_nz = self.m_n_nodes
        self.m_lx, self.m_ly, self.m_lz = self.m_length
        self.m_dl = tuple(l / n for l, n in zip(self.m_length, self.m_n_nodes))
        self.m_v = self.m_lx * self.m_ly * self.m_lz
        self.m_dv = self.m_dl[0] * self.m_dl[1] * self.m_dl[2]
        self.m_kx, self.m_ky, self.m_kz = (self._kproj(n, l) for n, l in zip(self.m_n_nodes, self.m_length))

    @staticmethod
    def _kproj(n, L):
        k = np.zeros(n)
        check(lib().mdft_kproj(n, ctypes.c_double(L), k.ctypes.data_as(_lib._pd)))
        return k

    @property
    def N(self):
        return self.m_nx * self.m_ny * self.m_nz


class AngularGrid:
    """common/MDFT_Grid.hpp:106-335 (the tables the hot path uses)."""

    def __init__(self, mmax: int, molrotsymorder: int):
        s = np.zeros(6, dtype=np.int32)
        check(lib().mdft_angular_sizes(mmax, molrotsymorder, s.ctypes.data_as(_lib._pi)))
        self.m_mmax, self.m_molrotsymorder = mmax, molrotsymorder
        self.m_ntheta, self.m_nphi, self.m_npsi, self.m_no, self.m_np, self.nmu2 = (int(v) for v in s)
        self.m_thetaofntheta = np.zeros(self.m_ntheta)
        self.m_wthetaofntheta = np.zeros(self.m_ntheta)
        self.m_w = np.zeros(self.m_no)
        self._pmap = np.zeros((mmax + 1, 2 * mmax + 1, self.nmu2), dtype=np.int32)
        self._m, self._mup, self._mu, self._mu2 = (np.zeros(self.m_np, dtype=np.int32) for _ in range(4))
        self._d = np.zeros((self.m_ntheta, self.m_np))
        P = _lib._pd
        Q = _lib._pi
        check(lib().mdft_angular_tables(mmax, molrotsymorder, self.m_thetaofntheta.ctypes.data_as(P),
                                        self.m_wthetaofntheta.ctypes.data_as(P), self.m_w.ctypes.data_as(P),
                                        self._pmap.ctypes.data_as(Q), self._m.ctypes.data_as(Q),
                                        self._mup.ctypes.data_as(Q), self._mu.ctypes.data_as(Q),
                                        self._mu2.ctypes.data_as(Q), self._d.ctypes.data_as(P)))
        self.m_quadrature_norm = 8.0 * PI * PI
        self.m_dphi = 2.0 * PI / self.m_nphi
        self.m_dpsi = 2.0 * PI / molrotsymorder / self.m_npsi


class OrientationProjectionMap:
    """common/MDFT_OrientationProjectionTransform.hpp:20-156."""

    def __init__(self, angular_grid: AngularGrid):
        self.angular_grid = angular_grid

    def p(self):
        return self.angular_grid._pmap

    def m(self):
        return self.angular_grid._m

    def mup(self):
        return self.angular_grid._mup

    def mu(self):
        return self.angular_grid._mu

    def mu2(self):
        return self.angular_grid._mu2

    def wigner_small_d(self):
        return self.angular_grid._d


class OrientationProjectionTransform:
    """common/MDFT_OrientationProjectionTransform.hpp:161-485 on the GPU."""

    def __init__(self, spatial_grid: SpatialGrid, angular_grid: AngularGrid, map: OrientationProjectionMap,
                 precision: int = FP64):
        self.sg, self.ag, self.precision = spatial_grid, angular_grid, precision
        h = ctypes.c_void_p()
        check(lib().mdft_opt_create(spatial_grid.m_nx, spatial_grid.m_ny, spatial_grid.m_nz, angular_grid.m_mmax,
                                    angular_grid.m_molrotsymorder, precision, ctypes.byref(h)))
        self._h, self._destroy = h, lib().mdft_opt_destroy
        self.real_dtype, self.complex_dtype = _dtypes(precision)

    def __del__(self):
        # the destroy entry point is bound at creation: module globals may
        # already be gone when this runs at interpreter exit
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)

    def angl2proj(self, o, p, stream=None):
        """o: (N, ntheta, nphi, npsi) real -> p: (np, nx, ny, nz) complex (:267-380)."""
        _require(o, self.real_dtype, self.sg.N * self.ag.m_no, "o")
        _require(p, self.complex_dtype, self.sg.N * self.ag.m_np, "p")
        check(lib().mdft_angl2proj(self._h, ptr(o), ptr(p), stream_handle(stream)))

    def proj2angl(self, p, o, stream=None):
        """p: (np, nx, ny, nz) complex -> o: (N, ntheta, nphi, npsi) real (:388-484)."""
        _require(o, self.real_dtype, self.sg.N * self.ag.m_no, "o")
        _require(p, self.complex_dtype, self.sg.N * self.ag.m_np, "p")
        check(lib().mdft_proj2angl(self._h, ptr(p), ptr(o), stream_handle(stream)))


class Convolution:
    """common/MDFT_Convolution.hpp:20-549 on the GPU."""

    def __init__(self, filename, spatial_grid: SpatialGrid, angular_grid: AngularGrid,
                 map: OrientationProjectionMap, np_luc: int, precision: int = FP64, _table=None):
        self.sg, self.ag, self.precision = spatial_grid, angular_grid, precision
        h = ctypes.c_void_p()
        g = spatial_grid
        if _table is None:
            check(lib().mdft_conv_create(str(filename).encode(), g.m_nx, g.m_ny, g.m_nz, g.m_lx, g.m_ly, g.m_lz,
                                         angular_grid.m_mmax, angular_grid.m_molrotsymorder, np_luc, precision,
                                         ctypes.byref(h)))
        else:
            tuples, c, dq = _table
            tuples = np.ascontiguousarray(tuples, dtype=np.int32)
            c = np.ascontiguousarray(c, dtype=np.complex128)
            check(lib().mdft_conv_create_from_table(g.m_nx, g.m_ny, g.m_nz, g.m_lx, g.m_ly, g.m_lz,
                                                    angular_grid.m_mmax, angular_grid.m_molrotsymorder,
                                                    tuples.shape[0], tuples.ctypes.data_as(_lib._pi),
                                                    c.ctypes.data_as(_lib._pd), c.shape[1], ctypes.c_double(dq),
                                                    precision, ctypes.byref(h)))
        self._h, self._destroy = h, lib().mdft_conv_destroy
        nq, npn, npairs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        dq_ = ctypes.c_double()
        check(lib().mdft_conv_info(h, ctypes.byref(nq), ctypes.byref(npn), ctypes.byref(npairs), ctypes.byref(dq_)))
        self.nq, self.np_new, self.npairs, self._dq = nq.value, npn.value, npairs.value, dq_.value
        self.complex_dtype = _dtypes(precision)[1]

    @classmethod
    def from_table(cls, table, spatial_grid, angular_grid, map, precision=FP64):
        """Construct from an in-memory Luc table (mini_mdft_amd.luc.LucTable)."""
        return cls(None, spatial_grid, angular_grid, map, table.np_luc, precision,
                   _table=(table.tuples, table.c, table.dq))

    def __del__(self):
        # the destroy entry point is bound at creation: module globals may
        # already be gone when this runs at interpreter exit
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)

    def dq(self):
        return self._dq

    def gamma_p_map(self):
        """The (q, -q) list in the reference's order (:126-176)."""
        g = self.sg
        out = np.zeros((self.npairs, 3), dtype=np.int32)
        check(lib().mdft_qpairs(g.m_nx, g.m_ny, g.m_nz, out.ctypes.data_as(_lib._pi)))
        return out

    def mnmunukhi_q(self):
        out = np.zeros((self.np_new, self.nq), dtype=np.complex128)
        check(lib().mdft_conv_table(self._h, out.ctypes.data_as(_lib._pd)))
        return out

    def fft_config(self) -> dict:
        """The FFT configuration this convolution tuned on its first execute."""
        v = [ctypes.c_int() for _ in range(6)]
        check(lib().mdft_conv_fft_config(self._h, *(ctypes.byref(x) for x in v)))
        return dict(zip(("tuned", "plane", "plane_threads", "wz", "wy", "wx"), (x.value for x in v)))

    def execute(self, deltarho_p, stream=None):
        """In place on (np, nx, ny, nz) complex (:324-548)."""
        _require(deltarho_p, self.complex_dtype, self.sg.N * self.ag.m_np, "deltarho_p")
        check(lib().mdft_conv_execute(self._h, ptr(deltarho_p), stream_handle(stream)))


def _dtype_code(t):
    torch = _torch()
    if t.dtype == torch.float64:
        return DOUBLE
    if t.dtype == torch.float32:
        return FLOAT
    raise MDFTError("fields must be float64 or float32")


def get_delta_rho(xi, delta_rho, rho0, stream=None):
    """common/MDFT_Solvent.hpp:575-596."""
    if delta_rho.numel() != xi.numel():
        raise MDFTError("size of delta_rho must be the same as the size of xi")
    _require(delta_rho, xi.dtype, xi.numel(), "delta_rho")
    _require(xi, xi.dtype, xi.numel(), "xi")
    check(lib().mdft_delta_rho(ptr(xi), ptr(delta_rho), xi.numel(), ctypes.c_double(rho0), _dtype_code(xi),
                               stream_handle(stream)))


def get_delta_f(xi, vexc, w, delta_f, rho0, prefactor, stream=None) -> float:
    """common/MDFT_Solvent.hpp:613-660; xi, vexc, delta_f (nx, ny, nz, no); w (no,) float64 on the
    device.  Returns ff (the call synchronises)."""
    torch = _torch()
    if tuple(xi.shape) != tuple(vexc.shape):
        raise MDFTError("angular grid size of xi and vexc must be the same")
    no = xi.shape[-1]
    _require(w, torch.float64, no, "w")
    for t, name in ((xi, "xi"), (vexc, "vexc"), (delta_f, "delta_f")):
        _require(t, xi.dtype, xi.numel(), name)
    ff = ctypes.c_double()
    check(lib().mdft_delta_f(ptr(xi), ptr(vexc), ptr(w), ptr(delta_f), xi.numel() // no, no, ctypes.c_double(rho0),
                             ctypes.c_double(prefactor), ctypes.byref(ff), _dtype_code(xi), stream_handle(stream)))
    return ff.value


class Solver:
    """One c-proj functional + gradient evaluation (src/mini-mdft.hpp:137-193)
    with delta_rho fused into angl2proj and get_delta_f into proj2angl."""

    def __init__(self, conv: Convolution, rho0: float, prefactor: float, precision: int = FP64, capacity: int = 1):
        h = ctypes.c_void_p()
        check(lib().mdft_solver_create_batch(conv._h, ctypes.c_double(rho0), ctypes.c_double(prefactor), precision,
                                             capacity, ctypes.byref(h)))
        self._h, self.conv, self.precision, self.capacity = h, conv, precision, capacity
        self._destroy = lib().mdft_solver_destroy
        self.rho0, self.prefactor = rho0, prefactor
        self.real_dtype = _dtypes(precision)[0]
        b = ctypes.c_size_t()
        check(lib().mdft_solver_workspace_bytes(h, ctypes.byref(b)))
        self.workspace_bytes = b.value

    def __del__(self):
        # the destroy entry point is bound at creation: module globals may
        # already be gone when this runs at interpreter exit
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)

    def energy_cproj(self, xi, df, stream=None) -> float:
        n = self.conv.sg.N * self.conv.ag.m_no
        _require(xi, self.real_dtype, n, "xi")
        _require(df, self.real_dtype, n, "df")
        ff = ctypes.c_double()
        check(lib().mdft_energy_cproj(self._h, ptr(xi), ptr(df), ctypes.byref(ff), stream_handle(stream)))
        return ff.value

    def energy_cproj_batch(self, xi, df, stream=None):
        """Evaluate the nbatch fields of xi (nbatch, N, no) in one pass of every
        stage (capacity >= nbatch); returns the nbatch ff values, df in place."""
        n = self.conv.sg.N * self.conv.ag.m_no
        nb = xi.shape[0]
        _require(xi, self.real_dtype, nb * n, "xi")
        _require(df, self.real_dtype, nb * n, "df")
        ff = (ctypes.c_double * nb)()
        check(lib().mdft_energy_cproj_batch(self._h, nb, ptr(xi), ptr(df), ff, stream_handle(stream)))
        return list(ff)

    STAGES = ("angl2proj", "fft_forward", "oz", "fft_backward", "proj2angl_delta_f")

    def set_mean_fp64(self, on: bool) -> None:
        """The r-mean (q = 0) mode of ff in FP64 (default on for MDFT_MIXED_X6;
        available for MDFT_FP64 / MDFT_FP32, where it changes ff by rounding only)."""
        check(lib().mdft_solver_set_mean_fp64(self._h, 1 if on else 0))

    def set_compact(self, on: bool) -> None:
        """The compact channel order of the fused evaluation (default on in
        every form where the grid allows): the mu = 0,
        mup < 0 projections, implied by their partners in the real delta_rho
        and folded into them after the OZ pass, are neither stored nor
        transformed; results change by rounding only."""
        check(lib().mdft_solver_set_compact(self._h, 1 if on else 0))

    @property
    def compact(self) -> bool:
        v = ctypes.c_int(0)
        check(lib().mdft_solver_compact(self._h, ctypes.byref(v)))
        return bool(v.value)

    def set_timing(self, enable: bool, oz_only: bool = False) -> None:
        """Per-stage hipEvent timers (the reference Solver's stage timers); with
        oz_only, only the OZ launch is timed, by the start and end events that
        the launch itself records."""
        check(lib().mdft_solver_set_timing(self._h, (2 if oz_only else 1) if enable else 0))

    def stage_ms(self):
        """(accumulated ms per stage as a dict, number of evaluations); waits for them."""
        ms = (ctypes.c_double * 5)()
        cnt = ctypes.c_long()
        check(lib().mdft_solver_stage_ms(self._h, ms, ctypes.byref(cnt)))
        return dict(zip(self.STAGES, list(ms))), cnt.value

    def energy_cproj_async(self, xi, df, ff_device, stream=None) -> None:
        torch = _torch()
        n = self.conv.sg.N * self.conv.ag.m_no
        _require(xi, self.real_dtype, n, "xi")
        _require(df, self.real_dtype, n, "df")
        _require(ff_device, torch.float64, 1, "ff_device")
        check(lib().mdft_energy_cproj_async(self._h, ptr(xi), ptr(df), ptr(ff_device), stream_handle(stream)))


def fft3d(data, sign: int, stream=None):
    """In-place unnormalised 3D FFT over the last three axes of a complex tensor."""
    torch = _torch()
    if data.dtype not in (torch.complex128, torch.complex64) or not data.is_cuda or not data.is_contiguous():
        raise MDFTError("fft3d needs a contiguous complex CUDA tensor")
    nx, ny, nz = data.shape[-3:]
    batch = data.numel() // (nx * ny * nz)
    dt = DOUBLE if data.dtype == torch.complex128 else FLOAT
    check(lib().mdft_fft3d(ptr(data), batch, nx, ny, nz, sign, dt, stream_handle(stream)))


def fft3d_config(data, sign: int, plane: bool, plane_threads: int = 512, wz: int = 0, wy: int = 0, wx: int = 0,
                 stream=None):
    """fft3d in a given configuration (fused z/y plane pass, or per-axis tiles
    of wz/wy/wx lines); bitwise the same result as fft3d."""
    torch = _torch()
    if data.dtype not in (torch.complex128, torch.complex64) or not data.is_cuda or not data.is_contiguous():
        raise MDFTError("fft3d needs a contiguous complex CUDA tensor")
    nx, ny, nz = data.shape[-3:]
    batch = data.numel() // (nx * ny * nz)
    dt = DOUBLE if data.dtype == torch.complex128 else FLOAT
    check(lib().mdft_fft3d_config(ptr(data), batch, nx, ny, nz, sign, dt, 1 if plane else 0, plane_threads, wz, wy,
                                  wx, stream_handle(stream)))


def fill_uniform(t, seed: int, lo: float, hi: float, offset: int = 0, stream=None):
    """t[i] = lo + (hi - lo) u(seed, offset + i), the stream of mini_mdft_amd.rng.uniform."""
    check(lib().mdft_fill_uniform(ptr(t), t.numel(), seed & 0xFFFFFFFFFFFFFFFF, offset, ctypes.c_double(lo),
                                  ctypes.c_double(hi), _dtype_code(t), stream_handle(stream)))


def rotation_matrices(angular_grid: AngularGrid) -> np.ndarray:
    """AngularGrid's rotation tables m_rotxx .. m_rotzz (common/MDFT_Grid.hpp:295-318)
    as an (no, 3, 3) array."""
    out = np.zeros((angular_grid.m_no, 3, 3))
    check(lib().mdft_rotation_matrices(angular_grid.m_mmax, angular_grid.m_molrotsymorder,
                                       out.ctypes.data_as(_lib._pd)))
    return out


def vext(spatial_grid: SpatialGrid, angular_grid: AngularGrid, solute, solvent, vmax: float, out=None, stream=None):
    """External potential (kJ/mol) of the solute sites on the grid, an (N, no)
    float64 CUDA tensor: Lennard-Jones (Lorentz-Berthelot) plus qfact q q / d
    over solvent x solute site pairs at the minimum-image distance, capped at
    vmax.  solute, solvent: (n, 6) rows (charge, sigma, epsilon, x, y, z), the
    solvent's in the molecular frame (mini_mdft_amd.solute)."""
    torch = _torch()
    g = spatial_grid
    solute = np.ascontiguousarray(solute, dtype=np.float64).reshape(-1, 6)
    solvent = np.ascontiguousarray(solvent, dtype=np.float64).reshape(-1, 6)
    if out is None:
        out = torch.empty((g.N, angular_grid.m_no), dtype=torch.float64, device="cuda")
    _require(out, torch.float64, g.N * angular_grid.m_no, "vext")
    check(lib().mdft_vext(g.m_nx, g.m_ny, g.m_nz, g.m_lx, g.m_ly, g.m_lz, angular_grid.m_mmax,
                          angular_grid.m_molrotsymorder, solute.shape[0], solute.ctypes.data_as(_lib._pd),
                          solvent.shape[0], solvent.ctypes.data_as(_lib._pd), ctypes.c_double(vmax), ptr(out),
                          stream_handle(stream)))
    return out


def xi_from_vext(V, kT: float, out=None, stream=None):
    """xi = exp(-V / (2 kT)) on the GPU (V / kT below -600 taken as -600): the
    minimum of F_id + F_ext alone, the density guess from V_ext of the original
    MDFT, as a start point for Functional.minimize.  The reference fills xi
    with 1 (common/MDFT_Solvent.hpp:287-296) because it has no V_ext."""
    torch = _torch()
    if out is None:
        out = torch.empty_like(V)
    _require(V, torch.float64, out.numel(), "xi_from_vext")
    _require(out, torch.float64, V.numel(), "xi_from_vext")
    check(lib().mdft_xi_from_vext(ptr(V), ctypes.c_longlong(V.numel()), ctypes.c_double(kT), ptr(out),
                                  stream_handle(stream)))
    return out


class Functional:
    """F = F_id + F_ext + F_exc and its gradient on the GPU, with an FP64
    L-BFGS minimiser (the reference's Solver makes one excess evaluation,
    src/mini-mdft.hpp:129-133; its iteration settings are unused,
    common/MDFT_System.hpp:188-189).  The excess term is the solver's c-proj
    evaluation in its precision form (MDFT_FP64 or a mixed form)."""

    STATUS = {0: "converged", 1: "max_iter", 2: "line search", 4: "non-finite start"}

    def __init__(self, solver: Solver, vext_tensor, kT_: float):
        torch = _torch()
        _require(vext_tensor, torch.float64, solver.conv.sg.N * solver.conv.ag.m_no, "vext")
        h = ctypes.c_void_p()
        check(lib().mdft_functional_create(solver._h, ptr(vext_tensor), ctypes.c_double(kT_), ctypes.byref(h)))
        self._h, self._destroy = h, lib().mdft_functional_destroy
        self.solver, self.vext, self.kT = solver, vext_tensor, kT_
        self.n = solver.conv.sg.N * solver.conv.ag.m_no

    def __del__(self):
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)

    def eval(self, xi, g, stream=None):
        """(F, F_id, F_ext, F_exc) in kJ/mol; g = dF/dxi written in place."""
        torch = _torch()
        _require(xi, torch.float64, self.n, "xi")
        _require(g, torch.float64, self.n, "g")
        F = (ctypes.c_double * 4)()
        check(lib().mdft_functional_eval(self._h, ptr(xi), ptr(g), F, stream_handle(stream)))
        return tuple(F)

    def minimize(self, xi, m: int = 5, max_iter: int = 200, gtol: float = 1e-5, stream=None) -> dict:
        """L-BFGS from xi, in place.  Returns F (4 terms), iterations,
        evaluations, status, residual (max |h0 g|) and the F history."""
        torch = _torch()
        _require(xi, torch.float64, self.n, "xi")
        F = (ctypes.c_double * 4)()
        it, nev, st = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        res = ctypes.c_double()
        hist = np.zeros(max_iter + 1)
        check(lib().mdft_lbfgs_minimize(self._h, ptr(xi), m, max_iter, ctypes.c_double(gtol), F, ctypes.byref(it),
                                        ctypes.byref(nev), ctypes.byref(st), ctypes.byref(res),
                                        hist.ctypes.data_as(_lib._pd), stream_handle(stream)))
        return dict(F=tuple(F), iterations=it.value, evaluations=nev.value, status=st.value,
                    residual=res.value, history=hist[:it.value + 1].copy())


class Server:
    """Batched and served minimisation (no reference counterpart): solutes of
    any shape under one HBM budget, same-shape solutes advanced together in
    engines of `capacity` slots whose ticks evaluate all slots in one batched
    pass of every stage, replayed as one HIP graph.  `table` is the c(q)
    (mini_mdft_amd.luc.LucTable) for every mmax that will be requested."""

    STATUS = {0: "converged", 1: "max_iter", 2: "line search", 3: "rejected (HBM budget)", 4: "non-finite start",
              5: "admission failed (see error)"}

    def __init__(self, table, mrso: int, n0: float, temperature: float, precision: int = MIXED_X6,
                 capacity: int = 4, m: int = 5, max_iter: int = 400, gtol: float = 1e-5,
                 budget_bytes: float = 200e9):
        tup = np.ascontiguousarray(table.tuples, dtype=np.int32)
        c = np.ascontiguousarray(table.c, dtype=np.complex128)
        h = ctypes.c_void_p()
        check(lib().mdft_server_create(tup.shape[0], tup.ctypes.data_as(_lib._pi), c.ctypes.data_as(_lib._pd),
                                       c.shape[1], ctypes.c_double(table.dq), mrso, ctypes.c_double(n0),
                                       ctypes.c_double(temperature), precision, capacity, m, max_iter,
                                       ctypes.c_double(gtol), ctypes.c_double(budget_bytes), ctypes.byref(h)))
        self._h, self._destroy = h, lib().mdft_server_destroy
        self.mrso, self._outs = mrso, {}

    def __del__(self):
        h, destroy = getattr(self, "_h", None), getattr(self, "_destroy", None)
        if h and destroy is not None:
            destroy(h)

    def use_graphs(self, on: bool) -> None:
        check(lib().mdft_server_use_graphs(self._h, 1 if on else 0))

    def start_from_vext(self, on: bool) -> None:
        """Start points of the solves (set before the first step): xi = 1 (the
        default) or xi = exp(-V / (2 kT)) from each solute's V_ext (xi_from_vext)."""
        check(lib().mdft_server_start_from_vext(self._h, 1 if on else 0))

    def submit(self, n_nodes, length, mmax: int, solute, solvent, vmax: float, out=None) -> int:
        """Queues one solute; `out` (a float64 CUDA tensor of N x no) receives
        the final xi.  Returns the request id."""
        solute = np.ascontiguousarray(solute, dtype=np.float64).reshape(-1, 6)
        solvent = np.ascontiguousarray(solvent, dtype=np.float64).reshape(-1, 6)
        rid = ctypes.c_longlong()
        check(lib().mdft_server_submit(self._h, *(int(v) for v in n_nodes), *(ctypes.c_double(v) for v in length),
                                       mmax, solute.shape[0], solute.ctypes.data_as(_lib._pd), solvent.shape[0],
                                       solvent.ctypes.data_as(_lib._pd), ctypes.c_double(vmax),
                                       ptr(out) if out is not None else None, ctypes.byref(rid)))
        if out is not None:
            self._outs[rid.value] = out  # kept alive until the solve ends
        return rid.value

    def step(self):
        """Admissions, then one tick of every engine; returns (active, pending)."""
        a, p = ctypes.c_int(), ctypes.c_int()
        check(lib().mdft_server_step(self._h, ctypes.byref(a), ctypes.byref(p)))
        return a.value, p.value

    def poll(self):
        """One finished solve as a dict, or None."""
        rid, it, nev, st, got = (ctypes.c_longlong(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(),
                                 ctypes.c_int())
        F = (ctypes.c_double * 4)()
        res, tq, tt = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().mdft_server_poll(self._h, ctypes.byref(rid), F, ctypes.byref(it), ctypes.byref(nev),
                                     ctypes.byref(st), ctypes.byref(res), ctypes.byref(tq), ctypes.byref(tt),
                                     ctypes.byref(got)))
        if not got.value:
            return None
        self._outs.pop(rid.value, None)
        err = ctypes.create_string_buffer(1024)
        check(lib().mdft_server_poll_error(self._h, err, len(err)))
        return dict(id=rid.value, F=tuple(F), iterations=it.value, evaluations=nev.value, status=st.value,
                    residual=res.value, queued_s=tq.value, total_s=tt.value, error=err.value.decode())

    def info(self):
        b, e, a, p = ctypes.c_double(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().mdft_server_info(self._h, ctypes.byref(b), ctypes.byref(e), ctypes.byref(a), ctypes.byref(p)))
        return dict(bytes_in_use=b.value, engines=e.value, active=a.value, pending=p.value)

    def run(self, max_steps: int = 100000):
        """Steps until nothing is active or pending; returns the results by id."""
        out = {}
        for _ in range(max_steps):
            a, p = self.step()
            while (r := self.poll()) is not None:
                out[r["id"]] = r
            if a == 0 and p == 0:
                break
        while (r := self.poll()) is not None:
            out[r["id"]] = r
        return out
