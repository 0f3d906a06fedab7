"""The benchmark configurations as test inputs (shared by the configuration
tests and the fixture script tests/golden/make_config_lbfgs.py).

* config 0: 32^3 grid in a 20 A box, mmax 3, FP64; a single ion at the box
  centre (q = +1 e, sigma 2.5 A, epsilon 0.5 kJ/mol);
* config 1 / 2: 64^3 in 30 A, mmax 5, FP64 / mixed; 20 LJ + charge sites
  within 5 A of the centre (min separation 1.2 A, sigma U[2.5, 3.8] A,
  epsilon U[0.1, 1.0] kJ/mol, charges U[-0.8, 0.8] e shifted to zero net);
* configs 3 / 4: grids up to 80^3 and 96^3 with box 0.47 A x side, mmax 5.

All use the rigid 3-site water of the configurations (tip3p sites:
common/MDFT_Solvent.hpp:471-532), bulk density 0.0332891 A^-3 (the spce n0,
common/MDFT_Solvent.hpp:429), T = 298.15 K and molrotsymorder 2.  The
evaluation inputs are the synthetic c(q) (seed 2024) and xi = 1 + 0.1 U(-1, 1)
(splitmix64 seed 12345).  Minimisations use the exchange-symmetric model c(q)
(seed 2024): with the synthetic c the excess operator is indefinite and
L-BFGS runs away within four iterations at config 0 (F = 381 -> -2.7e206
kJ/mol), so no trajectory of it is a meaningful check.
"""
from __future__ import annotations

from mini_mdft_amd import luc, solute
from mini_mdft_amd import mdft as M

TEMP, N0, MRSO = 298.15, 0.0332891, 2
C_SEED, XI_SEED, XI_AMP = 2024, 12345, 0.1
VMAX_KT = 100.0
WATER = solute.SOLVENT_SITES["tip3p"]

CONFIG0 = dict(n=(32, 32, 32), L=(20.0, 20.0, 20.0), mmax=3)
CONFIG1 = dict(n=(64, 64, 64), L=(30.0, 30.0, 30.0), mmax=5)


def kT():
    return M.kT(TEMP)


def rho0():
    return M.rho0_of(N0, MRSO)


def prefactor():
    return M.prefactor_of(TEMP, MRSO, N0)


def nq(cfg):
    return luc.nq_for(luc.qmax_of_grid(cfg["n"], cfg["L"]), luc.DQ_TIP3P)


def synthetic_table(cfg):
    return luc.synthetic_c(cfg["mmax"], MRSO, nq(cfg), seed=C_SEED)


def model_table(cfg):
    return luc.model_c_symmetric(cfg["mmax"], MRSO, nq(cfg), seed=C_SEED)


def config0_solute():
    return solute.single_ion(1.0, 2.5, 0.5, CONFIG0["L"])


def config1_solute(seed=1):
    return solute.synthetic_solute(20, seed, CONFIG1["L"], radius=5.0, min_sep=1.2, charge_scale=1.0)


def large(side):
    """configs 3 / 4: a side^3 grid in a (0.47 side)^3 A box, mmax 5."""
    return dict(n=(side,) * 3, L=(0.47 * side,) * 3, mmax=5)
