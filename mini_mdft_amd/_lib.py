"""ctypes binding of libmdft_hip.so (include/mdft_hip.h).

The library is loaded from ``mini_mdft_amd/lib``; there is no fallback: if it
is missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmdft_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "mdft_hip.h")

FP64, FP32, MIXED, MIXED_X6, MIXED_X6F = 0, 1, 2, 3, 4
DOUBLE, FLOAT = 0, 1

_lib = None

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_sz = ctypes.c_size_t
_ull = ctypes.c_ulonglong
_pi = ctypes.POINTER(ctypes.c_int)
_pd = ctypes.POINTER(ctypes.c_double)

_SIGS = {
    "mdft_last_error": (ctypes.c_char_p, []),
    "mdft_version": (_i, []),
    "mdft_gauss_legendre": (_i, [_i, _pd, _pd]),
    "mdft_wigner_small_d": (_i, [_i, _i, _i, _d, _pd]),
    "mdft_angular_sizes": (_i, [_i, _i, _pi]),
    "mdft_angular_tables": (_i, [_i, _i, _pd, _pd, _pd, _pi, _pi, _pi, _pi, _pi, _pd]),
    "mdft_kproj": (_i, [_i, _d, _pd]),
    "mdft_rotation_matrix": (_i, [_i, _pd, _pd]),
    "mdft_qpair_count": (_i, [_i, _i, _i]),
    "mdft_qpairs": (_i, [_i, _i, _i, _pi]),
    "mdft_luc_read": (_i, [ctypes.c_char_p, _i, _d, _pi, _pd, _pi, _pd]),
    "mdft_opt_create": (_i, [_i, _i, _i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "mdft_opt_destroy": (_i, [_vp]),
    "mdft_angl2proj": (_i, [_vp, _vp, _vp, _vp]),
    "mdft_proj2angl": (_i, [_vp, _vp, _vp, _vp]),
    "mdft_conv_create": (_i, [ctypes.c_char_p, _i, _i, _i, _d, _d, _d, _i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "mdft_conv_create_from_table": (_i, [_i, _i, _i, _d, _d, _d, _i, _i, _i, _pi, _pd, _i, _d, _i,
                                         ctypes.POINTER(_vp)]),
    "mdft_conv_info": (_i, [_vp, _pi, _pi, _pi, _pd]),
    "mdft_conv_table": (_i, [_vp, _pd]),
    "mdft_conv_execute": (_i, [_vp, _vp, _vp]),
    "mdft_conv_destroy": (_i, [_vp]),
    "mdft_fft3d": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp]),
    "mdft_fft3d_config": (_i, [_vp, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _i, _vp]),
    "mdft_conv_fft_config": (_i, [_vp, _pi, _pi, _pi, _pi, _pi, _pi]),
    "mdft_delta_rho": (_i, [_vp, _vp, _sz, _d, _i, _vp]),
    "mdft_delta_f": (_i, [_vp, _vp, _vp, _vp, _sz, _i, _d, _d, _pd, _i, _vp]),
    "mdft_delta_f_noalloc": (_i, [_vp, _vp, _vp, _vp, _sz, _i, _d, _d, _pd, _i, _vp]),
    "mdft_solver_create": (_i, [_vp, _d, _d, _i, ctypes.POINTER(_vp)]),
    "mdft_solver_workspace_bytes": (_i, [_vp, ctypes.POINTER(_sz)]),
    "mdft_solver_destroy": (_i, [_vp]),
    "mdft_energy_cproj": (_i, [_vp, _vp, _vp, _pd, _vp]),
    "mdft_energy_cproj_async": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "mdft_solver_set_timing": (_i, [_vp, _i]),
    "mdft_solver_set_mean_fp64": (_i, [_vp, _i]),
    "mdft_solver_set_compact": (_i, [_vp, _i]),
    "mdft_solver_compact": (_i, [_vp, ctypes.POINTER(_i)]),
    "mdft_solver_stage_ms": (_i, [_vp, _pd, ctypes.POINTER(ctypes.c_long)]),
    "mdft_fill_uniform": (_i, [_vp, _sz, _ull, _ull, _d, _d, _i, _vp]),
    "mdft_solver_create_batch": (_i, [_vp, _d, _d, _i, _i, ctypes.POINTER(_vp)]),
    "mdft_energy_cproj_batch": (_i, [_vp, _i, _vp, _vp, _pd, _vp]),
    "mdft_energy_cproj_batch_async": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "mdft_rotation_matrices": (_i, [_i, _i, _pd]),
    "mdft_vext": (_i, [_i, _i, _i, _d, _d, _d, _i, _i, _i, _pd, _i, _pd, _d, _vp, _vp]),
    "mdft_xi_from_vext": (_i, [_vp, ctypes.c_longlong, _d, _vp, _vp]),
    "mdft_functional_create": (_i, [_vp, _vp, _d, ctypes.POINTER(_vp)]),
    "mdft_functional_destroy": (_i, [_vp]),
    "mdft_functional_eval": (_i, [_vp, _vp, _vp, _pd, _vp]),
    "mdft_lbfgs_minimize": (_i, [_vp, _vp, _i, _i, _d, _pd, _pi, _pi, _pi, _pd, _pd, _vp]),
    "mdft_server_create": (_i, [_i, _pi, _pd, _i, _d, _i, _d, _d, _i, _i, _i, _i, _d, _d, ctypes.POINTER(_vp)]),
    "mdft_server_destroy": (_i, [_vp]),
    "mdft_server_use_graphs": (_i, [_vp, _i]),
    "mdft_server_start_from_vext": (_i, [_vp, _i]),
    "mdft_server_submit": (_i, [_vp, _i, _i, _i, _d, _d, _d, _i, _i, _pd, _i, _pd, _d, _vp,
                                ctypes.POINTER(ctypes.c_longlong)]),
    "mdft_server_step": (_i, [_vp, _pi, _pi]),
    "mdft_server_poll": (_i, [_vp, ctypes.POINTER(ctypes.c_longlong), _pd, _pi, _pi, _pi, _pd, _pd, _pd, _pi]),
    "mdft_server_info": (_i, [_vp, _pd, _pi, _pi, _pi]),
    "mdft_server_poll_error": (_i, [_vp, ctypes.c_char_p, _i]),
}


class MDFTError(RuntimeError):
    """A failure reported by the C ABI (status != 0)."""


def declared_symbols() -> list:
    """Names of every function include/mdft_hip.h declares."""
    with open(HEADER) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mdft_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MDFTError(f"{LIB_PATH} is not built: run __graft_entry__.build() (or make -C "
                            f"mini_mdft_amd/csrc)")
        # torch first: its HIP runtime (NEEDED as libamdhip64.so, soname
        # libamdhip64.so.7) is then the one the library binds to.  Loaded the
        # other way round, the process holds two HIP runtimes and the second
        # one sees no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise MDFTError(lib().mdft_last_error().decode())


def ptr(t) -> ctypes.c_void_p:
    """Device (or host) pointer of a torch tensor."""
    return ctypes.c_void_p(t.data_ptr())


def stream_handle(stream=None) -> ctypes.c_void_p:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
