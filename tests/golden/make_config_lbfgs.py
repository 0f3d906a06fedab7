#!/usr/bin/env python3
"""Writes tests/golden/lbfgs_64c_mmax5.npz: the FP64 oracle's L-BFGS on the
64^3 configuration (30 A box, mmax 5, the 20-site solute of seed 1 with full
charges, tip3p sites, V_ext capped at 100 kT, the symmetric model c(q) of seed
2024), m = 5, 50 iterations (gtol = 0) from xi = 1: the F history, the
iteration and evaluation counts, the status and summary numbers of the final
point.  The oracle needs about half an hour on 8 cores for it, so the GPU test
compares against these numbers instead of running it.

Usage (from the repository root): python tests/golden/make_config_lbfgs.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.dirname(HERE)):
    sys.path.insert(0, p)

import configs as C  # noqa: E402
import oracle as O  # noqa: E402

ITERS, M_PAIRS = 50, 5


def main():
    O.set_num_threads(os.cpu_count() or 1)
    cfg = C.CONFIG1
    n, L, mmax = cfg["n"], cfg["L"], cfg["mmax"]
    tab = C.model_table(cfg)
    conv = O.Conv(n, L, mmax, C.MRSO, tab.tuples, tab.c, tab.dq)
    kT, rho0, pref = C.kT(), C.rho0(), C.prefactor()
    t0 = time.time()
    V = O.vext(n, L, mmax, C.MRSO, C.config1_solute(), C.WATER, C.VMAX_KT * kT)
    print(f"vext {time.time() - t0:.1f} s", flush=True)
    w = O.angular_tables(mmax, C.MRSO)["w"]
    N, no = V.shape
    dv = float(np.prod(L) / N)
    h0 = np.broadcast_to(1.0 / (2 * rho0 * dv * w * kT), (N, no)).ravel()

    def fg(x):
        (F, *parts), g = O.functional(conv, x.reshape(N, -1), V, w, rho0, pref, kT, dv)
        print(f"  F = {F!r} ({time.time() - t0:.0f} s)", flush=True)
        return F, g.ravel(), sum(abs(p) for p in parts)

    x, F, g, it, nev, hist, status = O.lbfgs(fg, np.ones(N * no), h0, m=M_PAIRS, max_iter=ITERS, gtol=0.0)
    np.savez(os.path.join(HERE, "lbfgs_64c_mmax5.npz"), history=np.asarray(hist), iterations=it, evaluations=nev,
             status=status, residual=float(np.abs(h0 * g).max()), x_sum=float(x.sum()),
             x_sq_sum=float((x * x).sum()), m=M_PAIRS)
    print(f"done: {it} iterations, {nev} evaluations, status {status}, F = {F!r}", flush=True)


if __name__ == "__main__":
    main()
