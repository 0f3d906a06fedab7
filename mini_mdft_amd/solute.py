"""Solute and solvent site models for the external potential.

The solvent sites are the reference's rigid 3-site water models
(common/MDFT_Solvent.hpp:417-532: charge, LJ sigma and epsilon and the
molecular-frame position of O, H1, H2).  The synthetic solute follows the
benchmark configurations: nsite Lennard-Jones + charge sites at seed-uniform
positions within a radius of the box centre, a minimum separation, sigma, epsilon
and charge drawn uniformly, charges shifted to a zero net charge and then
multiplied by the solute-charge scale factor of Solute::read_solute
(common/MDFT_Solute.hpp:126-150).  Sites are moved to the box centre as
Solute::mv_solute_to_center does (:152-178).
"""
from __future__ import annotations

import numpy as np

from .rng import uniform

#: solvent sites (charge e, sigma A, epsilon kJ/mol, x, y, z A in the molecular frame)
SOLVENT_SITES = {
    "spce": np.array([[-0.8476, 3.166, 0.65, 0.0, 0.0, 0.0],
                      [0.4238, 0.0, 0.0, 0.816495, 0.0, 0.5773525],
                      [0.4238, 0.0, 0.0, -0.816495, 0.0, 0.5773525]]),
    "tip3p": np.array([[-0.834, 3.15061, 0.636386, 0.0, 0.0, 0.0],
                       [0.417, 0.0, 0.0, 0.756950, 0.0, 0.585882],
                       [0.417, 0.0, 0.0, -0.756950, 0.0, 0.585882]]),
}


def synthetic_solute(nsite: int, seed: int, box, radius: float = 5.0, min_sep: float = 1.2,
                     charge_scale: float = 1.0) -> np.ndarray:
    """(nsite, 6) array of (charge, sigma, epsilon, x, y, z): positions uniform
    in a ball of `radius` A (rejection sampling, pairwise separation >= min_sep),
    sigma ~ U[2.5, 3.8] A, epsilon ~ U[0.1, 1.0] kJ/mol, charge ~ U[-0.8, 0.8] e
    shifted to a net charge of 0 and multiplied by charge_scale; centred in the
    box."""
    pos = []
    k = 0
    while len(pos) < nsite:
        u = uniform(seed, 3, offset=3 * k)
        k += 1
        p = radius * (2.0 * u - 1.0)
        if np.dot(p, p) > radius * radius:
            continue
        if all(np.linalg.norm(p - q) >= min_sep for q in pos):
            pos.append(p)
        if k > 100000:
            raise RuntimeError("could not place the solute sites")
    pos = np.asarray(pos)
    u = uniform(seed + 1, 3 * nsite).reshape(nsite, 3)
    sigma = 2.5 + 1.3 * u[:, 0]
    eps = 0.1 + 0.9 * u[:, 1]
    q = -0.8 + 1.6 * u[:, 2]
    q = (q - q.mean()) * charge_scale
    centre = 0.5 * np.asarray(box, dtype=np.float64)
    pos = pos - pos.mean(axis=0) + centre
    return np.column_stack([q, sigma, eps, pos])


def single_ion(charge: float, sigma: float, epsilon: float, box) -> np.ndarray:
    """One site at the box centre (the config-0 solute)."""
    c = 0.5 * np.asarray(box, dtype=np.float64)
    return np.array([[charge, sigma, epsilon, c[0], c[1], c[2]]])
