#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the FP64 oracle.

The fixtures are small, seed-documented input/output vectors of the oracle:
  tables.npz   GL nodes/weights (n = 6), the Wigner table and projection map of
               mmax = 5, mrso = 2, and R(q) at mmax = 5 for the q vectors of the
               reference's rotation test (unit_test/Test_Rotation.cpp:20-23: q = 0,
               (1, 2, 3), kx0 (1, 1, 1)) plus q along -z
  proj.npz     angl2proj and proj2angl on a 4^3 grid, mmax = 3, mrso = 2: the
               input o (seed 101) and p (seed 102), and both outputs
  conv.npz     Convolution::execute on 4^3 and 5^3 grids (boxes 5 and 6 A),
               mmax = 3, synthetic c(q) nmax = 3 (seed 7): input p (seeds 201,
               202) and output
  energy.npz   one c-proj evaluation at 8^3 (box 10 A), mmax = 3, spce n0 at
               298.15 K, synthetic c(q) (seed 11), xi from the splitmix64 field
               (seed 12345, amplitude 0.1): ff and df
They let the tests check the oracle against itself over time (regression) and
the GPU path against fixed numbers without rebuilding the expected values.

Usage (from the repository root): python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from mini_mdft_amd import luc, rng  # noqa: E402

KX0 = 2 * np.pi / 10.0
QS = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 3.0], [KX0, KX0, KX0], [0.0, 0.0, -2.0]])
SPCE_N0, TEMP, MRSO = 0.0332891, 298.15, 2


def rho0_pref(n0, mrso, T):
    kT = 1.380649e-23 * 6.02214076e23 * T * 1e-3
    return n0 * mrso / (8 * np.pi ** 2), -2 * kT * 4 * np.pi ** 2 / mrso / n0


def main():
    x, w = O.gauss_legendre(6)
    t = O.angular_tables(5, MRSO)
    R = np.stack([O.rotation_matrix(5, q) for q in QS])
    np.savez_compressed(os.path.join(HERE, "tables.npz"), gl_x=x, gl_w=w, wigner=t["D"], pmap=t["pmap"],
                        theta=t["theta"], w=t["w"], q=QS, R=R)

    n = (4, 4, 4)
    s = O.angular_sizes(3, MRSO)
    o = np.random.default_rng(101).random((64, s["ntheta"], s["nphi"], s["npsi"]))
    p = np.random.default_rng(102).standard_normal((s["np"],) + n) + 1j * np.random.default_rng(103).standard_normal(
        (s["np"],) + n)
    np.savez_compressed(os.path.join(HERE, "proj.npz"), o=o, p_of_o=O.angl2proj(o, n, 3, MRSO), p=p,
                        o_of_p=O.proj2angl(p, 3, MRSO))

    out = {}
    for tag, n, L, seed in (("4", (4, 4, 4), (5.0,) * 3, 201), ("5", (5, 5, 5), (6.0,) * 3, 202)):
        tab = luc.synthetic_c(3, MRSO, luc.nq_for(luc.qmax_of_grid(n, L), luc.DQ_TIP3P), seed=7)
        C = O.Conv(n, L, 3, MRSO, tab.tuples, tab.c, tab.dq)
        g = np.random.default_rng(seed)
        pin = g.standard_normal((s["np"],) + n) + 1j * g.standard_normal((s["np"],) + n)
        out[f"p_in_{tag}"] = pin
        out[f"p_out_{tag}"] = C.execute(pin.copy())
    np.savez_compressed(os.path.join(HERE, "conv.npz"), **out)

    n, L = (8, 8, 8), (10.0,) * 3
    tab = luc.synthetic_c(3, MRSO, luc.nq_for(luc.qmax_of_grid(n, L), luc.DQ_TIP3P), seed=11)
    C = O.Conv(n, L, 3, MRSO, tab.tuples, tab.c, tab.dq)
    rho0, pref = rho0_pref(SPCE_N0, MRSO, TEMP)
    xi = rng.xi_field(12345, (512, s["no"]), 0.1)
    ff, df = C.energy_cproj(xi, rho0, pref)
    np.savez_compressed(os.path.join(HERE, "energy.npz"), ff=ff, df=df, rho0=rho0, prefactor=pref)


if __name__ == "__main__":
    main()
