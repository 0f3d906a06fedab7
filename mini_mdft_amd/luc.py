"""The Luc c(q) projection file: writer, synthetic generator and tuple sets.

The format is the one read by common/IO/MDFT_ReadCLuc.hpp:123-259:

* lines 1-10: free header text (skipped, :140);
* line 11: a label line (:144);
* lines 12-16: ``label v_1 ... v_np`` for m, n, mu, nu, khi (:146-159);
* line 17: free text (the second ``getline`` at :161-162; deltaAbscissa skips
  17 lines, :63, :189-222);
* then one row per q: ``q (re,im) (re,im) ...`` in ``std::complex`` stream
  format (:164-181), starting at q = 0 with a uniform step dq.

The reader that the product uses is the C++ one in the HIP library
(``mdft_luc_read``); this module writes files and generates the synthetic c(q)
that the benchmark configurations define.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .rng import uniform

#: dq of the reference's tip3p c(q) file (unit_test/Test_IO.cpp:104)
DQ_TIP3P = 6.135923151542601e-2


def luc_tuples(nmax: int, mrso: int) -> np.ndarray:
    """All (m, n, mu, nu, khi) with m, n <= nmax, mu, nu multiples of mrso with
    |mu| <= m, |nu| <= n and |khi| <= min(m, n), in descending order (the
    reference file lists (nmax, nmax, ...) first and (0, 0, 0, 0, 0) last,
    unit_test/Test_IO.cpp:133-142).  Its length is npluc[nmax]
    (common/MDFT_Solvent.hpp:488: 1, 6, 75, 252, 877, 2002 for mrso = 2)."""
    out = []
    for m in range(nmax, -1, -1):
        for n in range(nmax, -1, -1):
            for mu in range((m // mrso) * mrso, -(m // mrso) * mrso - 1, -mrso):
                for nu in range((n // mrso) * mrso, -(n // mrso) * mrso - 1, -mrso):
                    for khi in range(min(m, n), -min(m, n) - 1, -1):
                        out.append((m, n, mu, nu, khi))
    return np.asarray(out, dtype=np.int32).reshape(-1, 5)


def qmax_of_grid(n, length) -> float:
    """|(max kx, max ky, max kz)| as Convolution's constructor computes it
    (common/MDFT_Convolution.hpp:214-231): kmax along each axis is
    (n//2) * 2 pi / L."""
    ks = [(int(ni) // 2) * 2.0 * math.pi / float(li) for ni, li in zip(n, length)]
    return math.sqrt(sum(k * k for k in ks))


def nq_for(qmax: float, dq: float) -> int:
    """Rows the reader needs: int(qmax / dq + 0.5) + 2 (common/IO/MDFT_ReadCLuc.hpp:65)."""
    return int(qmax / dq + 0.5) + 2


@dataclass
class LucTable:
    tuples: np.ndarray  # (np_luc, 5) int: m, n, mu, nu, khi
    c: np.ndarray  # (np_luc, nq) complex128
    dq: float

    @property
    def np_luc(self) -> int:
        return int(self.tuples.shape[0])

    @property
    def nq(self) -> int:
        return int(self.c.shape[1])


def synthetic_c(nmax: int, mrso: int, nq: int, seed: int, dq: float = DQ_TIP3P,
                nterms: int = 3) -> LucTable:
    """Synthetic c(q): for every tuple, sum_k a_k sinc(q r0_k) exp(-q^2/8) with
    a_k ~ U[-5, 5] and r0_k ~ U[2, 6] Angstrom from the counter-based stream
    (seed, 2 * (nterms * tuple + k) + {0, 1}); sinc(0) = 1.  Real valued."""
    tup = luc_tuples(nmax, mrso)
    ntup = tup.shape[0]
    u = uniform(seed, 2 * nterms * ntup).reshape(ntup, nterms, 2)
    a = -5.0 + 10.0 * u[..., 0]
    r0 = 2.0 + 4.0 * u[..., 1]
    q = np.arange(nq, dtype=np.float64) * dq
    x = q[None, None, :] * r0[:, :, None]
    sinc = np.where(x == 0.0, 1.0, np.sin(x) / np.where(x == 0.0, 1.0, x))
    c = (a[:, :, None] * sinc).sum(axis=1) * np.exp(-q * q / 8.0)[None, :]
    return LucTable(tup, c.astype(np.complex128), dq)


def model_c_symmetric(nmax: int, mrso: int, nq: int, seed: int, dq: float = DQ_TIP3P,
                      amplitude: float = 1.0) -> LucTable:
    """A model c(q) for minimisations: exchange symmetric and stabilising.

    Only the tuples (m, m, 0, 0, 0) are nonzero: the correlation of the
    molecular axes along q, one channel per m <= nmax.  For this family the
    excess operator K = proj2angl . conv . angl2proj is self-adjoint in the
    w-weighted orientation inner product and negative semidefinite, so F_exc
    = (dv / N) prefactor rho0^2 <u, K u>_w (prefactor < 0) is a nonnegative
    quadratic form and F is bounded below; the gradient 2 prefactor w vexc
    d(Delta rho) is exact (tests/test_functional.py checks the symmetry and
    the spectrum).  Each value is -a exp(-(q r0)^2 / 4) with a = amplitude
    U[0.2, 1] / (1 + m) and r0 ~ U[2, 3.5] A (the stream of (seed, 2 tuple +
    {0, 1})): negative like the core of a physical c, decayed by the Nyquist
    wave numbers of the grids used (where even grids pair a mode with itself
    and break the symmetry).  The independent draws of synthetic_c, and
    channels with mu != 0 or khi != 0, make K indefinite."""
    tup = luc_tuples(nmax, mrso)
    u = uniform(seed, 2 * tup.shape[0]).reshape(-1, 2)
    q = np.arange(nq, dtype=np.float64) * dq
    c = np.zeros((tup.shape[0], nq), dtype=np.complex128)
    for i, (m, n, mu, nu, khi) in enumerate(tup):
        if m == n and mu == 0 and nu == 0 and khi == 0:
            a = amplitude * (0.2 + 0.8 * u[i, 0]) / (1.0 + m)
            r0 = 2.0 + 1.5 * u[i, 1]
            c[i] = -a * np.exp(-(q * r0) ** 2 / 4.0)
    return LucTable(tup, c, dq)


def write_luc(path, table: LucTable, header: str = "synthetic c(q)") -> None:
    """Write ``table`` in the Luc text format described in the module docstring."""
    tup = table.tuples
    lines = []
    for i in range(10):
        lines.append(f"# {header} (header line {i + 1})")
    lines.append("# projections")
    for j, lab in enumerate(("m", "n", "mu", "nu", "khi")):
        lines.append(lab + " " + " ".join(str(int(v)) for v in tup[:, j]))
    lines.append("# q  c(q) for each projection")
    for iq in range(table.nq):
        q = iq * table.dq
        row = [repr(float(q))]
        for v in table.c[:, iq]:
            row.append(f"({float(v.real)!r},{float(v.imag)!r})")
        lines.append(" ".join(row))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
