"""The full functional, V_ext and the L-BFGS of the oracle (CPU), and on the GPU
the library's versions of them against the oracle.

The reference evaluates only F_exc (src/mini-mdft.hpp:129-133) and leaves
Solvent::m_vext unused (common/MDFT_Solvent.hpp:160), so the oracle's
functional, vext and lbfgs restate this build's definitions; what pins them:

* rotation_matrices -- scipy's intrinsic ZYZ Euler matrices Rz(phi) Ry(theta)
  Rz(psi) at the grid's (theta, phi, psi), and orthogonality;
* vext              -- closed forms for one LJ site and for one charge seen by
  an off-centre solvent site (distance through the scipy rotation), minimum
  image, the cap, and QFACT against the Coulomb constant 1389.35 kJ/mol A / e^2;
* functional        -- a central finite difference of F along a random
  direction against g . d (a dropped or mis-scaled term of g fails it), and
  F_exc against the quadratic form (dv / N) prefactor rho0^2 <u, K u>_w built
  from the pinned transforms;
* model_c_symmetric -- <a, K b>_w = <b, K a>_w for the excess operator K, and
  the same test failing for the non-symmetric synthetic c (the test can fail);
* lbfgs             -- the c = 0 minimiser xi^2 = exp(-V / kT) (closed form),
  monotone F, and agreement with scipy's L-BFGS-B minimum.
"""
import math

import numpy as np
import pytest

from mini_mdft_amd import luc, solute
from mini_mdft_amd import mdft as M

# an odd grid: on even grids the Nyquist planes (where q and -q share an index)
# make the excess operator slightly non-self-adjoint (1.7e-5 relative at 8^3 in
# a 10 A box; below 1e-10 at 64^3 in 30 A, where c(q) has decayed)
N3, L3, MMAX, MRSO = (9, 9, 9), (10.0, 10.0, 10.0), 3, 2
NPTS = 729
TEMP = 298.15
KT = M.kT(TEMP)
N0 = M.SOLVENTS["tip3p"]["n0"]
RHO0 = M.rho0_of(N0, MRSO)
PREF = M.prefactor_of(TEMP, MRSO, N0)
DV = float(np.prod(L3) / np.prod(N3))
NQ = luc.nq_for(luc.qmax_of_grid(N3, L3), luc.DQ_TIP3P)


def _setup(oracle, amp=1.0, charge_scale=0.1, nsite=3, seed=3, vmax_kt=30.0, c_sym=True):
    s = oracle.angular_sizes(MMAX, MRSO)
    w = oracle.angular_tables(MMAX, MRSO)["w"]
    tab = luc.model_c_symmetric(MMAX, MRSO, NQ, seed=5, amplitude=amp) if c_sym else \
        luc.synthetic_c(MMAX, MRSO, NQ, seed=5)
    conv = oracle.Conv(N3, L3, MMAX, MRSO, tab.tuples, tab.c, tab.dq)
    sol = solute.synthetic_solute(nsite, seed, L3, radius=2.0, charge_scale=charge_scale)
    V = oracle.vext(N3, L3, MMAX, MRSO, sol, solute.SOLVENT_SITES["tip3p"], vmax_kt * KT)
    return s, w, conv, V


def test_qfact_is_the_coulomb_constant(oracle):
    # e^2 N_A / (4 pi eps0) in kJ/mol A = 1389.35
    assert abs(oracle.QFACT - 1389.3545) < 1e-3


def test_rotation_matrices_are_zyz_euler(oracle):
    from scipy.spatial.transform import Rotation

    s = oracle.angular_sizes(MMAX, MRSO)
    t = oracle.angular_tables(MMAX, MRSO)
    R = oracle.rotation_matrices(MMAX, MRSO)
    assert R.shape == (s["no"], 3, 3)
    io = 0
    for th in t["theta"]:
        for j in range(s["nphi"]):
            for k in range(s["npsi"]):
                ph, ps = 2 * np.pi * j / s["nphi"], 2 * np.pi * k / (s["npsi"] * MRSO)
                E = Rotation.from_euler("ZYZ", [ph, th, ps]).as_matrix()
                assert np.abs(R[io] - E).max() <= 1e-14
                io += 1
    assert np.abs(np.einsum("oij,okj->oik", R, R) - np.eye(3)).max() <= 1e-14


def test_vext_single_lj_site(oracle):
    """A neutral LJ solute at the box centre and a one-site solvent at the
    molecular origin: V = 4 eps ((sig/d)^12 - (sig/d)^6) with Lorentz-Berthelot
    mixing and the minimum-image d, the same for every orientation."""
    L = np.asarray(L3)
    sol = np.array([[0.0, 3.0, 0.4, 5.0, 5.0, 5.0]])
    solv = np.array([[0.0, 3.2, 0.6, 0.0, 0.0, 0.0]])
    V = oracle.vext(N3, L3, MMAX, MRSO, sol, solv, 1e30)
    sig, eps = 3.1, math.sqrt(0.24)
    for r in (0, 7, 100, 728, 364):
        i, j, k = np.unravel_index(r, N3)
        d = np.array([i, j, k]) * L / np.asarray(N3) - 5.0
        d -= L * np.round(d / L)
        x = np.linalg.norm(d)
        ref = 4 * eps * ((sig / x) ** 12 - (sig / x) ** 6) if x > 0 else None
        if ref is None:
            continue
        assert np.allclose(V[r], ref, rtol=1e-13, atol=0)


def test_vext_charge_off_centre_site_and_cap(oracle):
    """A charge seen by a solvent site at s = (0.3, -0.2, 0.9) in the molecular
    frame: V(r, o) = QFACT q1 q2 / |x_r + R_o s - x| (R_o through scipy), and the
    cap vmax where the site sits on the charge."""
    from scipy.spatial.transform import Rotation

    s = oracle.angular_sizes(MMAX, MRSO)
    t = oracle.angular_tables(MMAX, MRSO)
    L = np.asarray(L3)
    x0 = np.array([3.3, 6.1, 4.7])
    sol = np.array([[0.7, 0.0, 0.0, *x0]])
    sv = np.array([0.3, -0.2, 0.9])
    solv = np.array([[-0.5, 0.0, 0.0, *sv]])
    vmax = 50.0
    V = oracle.vext(N3, L3, MMAX, MRSO, sol, solv, vmax)
    rng = np.random.default_rng(0)
    for _ in range(40):
        r, io = int(rng.integers(NPTS)), int(rng.integers(s["no"]))
        it, rem = divmod(io, s["nphi"] * s["npsi"])
        j, k = divmod(rem, s["npsi"])
        E = Rotation.from_euler("ZYZ", [2 * np.pi * j / s["nphi"], t["theta"][it],
                                        2 * np.pi * k / (s["npsi"] * MRSO)]).as_matrix()
        d = np.array(np.unravel_index(r, N3)) * L / np.asarray(N3) + E @ sv - x0
        d -= L * np.round(d / L)
        ref = min(oracle.QFACT * 0.7 * -0.5 / np.linalg.norm(d), vmax)
        assert abs(V[r, io] - ref) <= 1e-12 * abs(ref)
    # a positive pair at a tiny distance is capped
    Vp = oracle.vext(N3, L3, MMAX, MRSO, np.array([[0.7, 0.0, 0.0, 0.0, 0.0, 0.0]]),
                     np.array([[0.5, 0.0, 0.0, 0.0, 0.0, 0.0]]), vmax)
    assert np.all(Vp[0] == vmax) and Vp.max() == vmax


def _K(oracle, conv, o):
    """The excess operator proj2angl . conv . angl2proj on an (N, no) field."""
    s = oracle.angular_sizes(MMAX, MRSO)
    p = oracle.angl2proj(o.reshape(-1, s["ntheta"], s["nphi"], s["npsi"]), N3, MMAX, MRSO)
    return oracle.proj2angl(conv.execute(p), MMAX, MRSO).reshape(o.shape)


@pytest.mark.parametrize("sym", [True, False])
def test_model_c_makes_the_excess_operator_self_adjoint(oracle, sym):
    s, w, conv, _ = _setup(oracle, c_sym=sym)
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal((2, NPTS, s["no"]))
    ab = np.sum(w * a * _K(oracle, conv, b))
    ba = np.sum(w * b * _K(oracle, conv, a))
    asym = abs(ab - ba) / abs(ab)
    if sym:
        assert asym <= 1e-12
    else:
        assert asym >= 1e-3  # the independent draws of synthetic_c are not symmetric


@pytest.mark.parametrize("sym", [True, False])
def test_model_c_makes_the_excess_operator_negative_semidefinite(oracle, sym):
    """The dense w-weighted matrix of K on an odd 3^3 grid (mmax 2): no
    positive eigenvalue for the model c, positive ones for synthetic_c."""
    n, L, mmax = (3, 3, 3), (4.0,) * 3, 2
    s = oracle.angular_sizes(mmax, MRSO)
    w = oracle.angular_tables(mmax, MRSO)["w"]
    nq = luc.nq_for(luc.qmax_of_grid(n, L), luc.DQ_TIP3P)
    tab = luc.model_c_symmetric(mmax, MRSO, nq, seed=5) if sym else luc.synthetic_c(mmax, MRSO, nq, seed=5)
    conv = oracle.Conv(n, L, mmax, MRSO, tab.tuples, tab.c, tab.dq)
    dim = 27 * s["no"]
    A = np.zeros((dim, dim))
    for j in range(dim):
        e = np.zeros(dim)
        e[j] = 1.0
        p = oracle.angl2proj(e.reshape(27, s["ntheta"], s["nphi"], s["npsi"]), n, mmax, MRSO)
        A[:, j] = oracle.proj2angl(conv.execute(p), mmax, MRSO).ravel()
    A *= np.tile(w, 27)[:, None]
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    if sym:
        assert ev.max() <= 1e-12 * abs(ev).max() and ev.min() < 0
    else:
        assert ev.max() > 1e-3 * abs(ev).max()


def test_functional_gradient_finite_difference(oracle):
    s, w, conv, V = _setup(oracle)
    rng = np.random.default_rng(2)
    xi = 1.0 + 0.1 * rng.standard_normal((NPTS, s["no"]))
    d = rng.standard_normal(xi.shape)
    (F, fid, fext, fexc), g = oracle.functional(conv, xi, V, w, RHO0, PREF, KT, DV)
    assert abs(F - (fid + fext + fexc)) <= 1e-12 * abs(F)
    h = 1e-5
    Fp = oracle.functional(conv, xi + h * d, V, w, RHO0, PREF, KT, DV)[0][0]
    Fm = oracle.functional(conv, xi - h * d, V, w, RHO0, PREF, KT, DV)[0][0]
    fd = (Fp - Fm) / (2 * h)
    gd = float(np.sum(g * d))
    assert abs(fd - gd) <= 1e-7 * abs(gd)


def test_functional_excess_is_the_quadratic_form(oracle):
    """ff = prefactor rho0 sum w (xi^2 - 1) vexc (common/MDFT_Solvent.hpp:643-657)
    with vexc = K(delta_rho) and delta_rho = rho0 (xi^2 - 1), so F_exc = (dv /
    N) prefactor rho0^2 <u, K u>_w for u = xi^2 - 1, built here from the pinned
    transforms; and F_exc scales as t^2 along u = t u1."""
    s, w, conv, V = _setup(oracle)
    rng = np.random.default_rng(4)
    u = np.clip(0.2 * rng.standard_normal((NPTS, s["no"])), -0.3, 0.3)
    def fexc(t):
        return oracle.functional(conv, np.sqrt(1.0 + t * u), V, w, RHO0, PREF, KT, DV)[0][3]
    f1 = fexc(1.0)
    ref = DV / NPTS * PREF * RHO0 ** 2 * np.sum(w * u * _K(oracle, conv, u))
    assert abs(f1 - ref) <= 1e-10 * abs(ref)  # the sum over w u K u cancels to 1e-3 of its terms
    assert abs(fexc(0.0)) <= 1e-14 * abs(f1)
    for t in (2.0, 3.0):
        assert abs(fexc(t) - t * t * f1) <= 1e-10 * abs(t * t * f1)


def test_lbfgs_c0_limit_is_boltzmann(oracle):
    """With c = 0 the functional is separable and its minimisers are xi^2 =
    exp(-V / kT) at every node; F depends on xi^2 only, so a node may end at
    either sign of xi."""
    s, w, conv, V = _setup(oracle, amp=0.0, vmax_kt=10.0)
    h0 = np.broadcast_to(1.0 / (2 * RHO0 * DV * w * KT), (NPTS, s["no"])).ravel()
    def fg(x):
        (F, *parts), g = oracle.functional(conv, x.reshape(NPTS, -1), V, w, RHO0, PREF, KT, DV)
        return F, g.ravel(), sum(abs(p) for p in parts)
    x, F, g, it, nev, hist, status = oracle.lbfgs(fg, np.ones(NPTS * s["no"]), h0, m=5, max_iter=300, gtol=1e-10)
    assert status == 0
    assert np.abs(np.abs(x) - np.exp(-V / (2 * KT)).ravel()).max() <= 1e-9
    # F at the minimum: kT rho0 dv sum w (1 - exp(-V/kT))
    assert abs(F - KT * RHO0 * DV * np.sum(w * (1 - np.exp(-V / KT)))) <= 1e-10 * abs(F)


def test_lbfgs_monotone_and_matches_scipy(oracle):
    from scipy.optimize import minimize

    s, w, conv, V = _setup(oracle)
    h0 = np.broadcast_to(1.0 / (2 * RHO0 * DV * w * KT), (NPTS, s["no"])).ravel()
    def fg(x):
        (F, *parts), g = oracle.functional(conv, x.reshape(NPTS, -1), V, w, RHO0, PREF, KT, DV)
        return F, g.ravel(), sum(abs(p) for p in parts)
    x, F, g, it, nev, hist, status = oracle.lbfgs(fg, np.ones(NPTS * s["no"]), h0, m=5, max_iter=300, gtol=1e-6)
    assert status == 0 and np.abs(h0 * g).max() < 1e-6
    # F decreases, up to the approximate Wolfe test's allowance of 1e-6 of its scale
    tol = 2e-6 * abs(hist[0])
    assert all(b <= a + tol for a, b in zip(hist, hist[1:]))
    # scipy's L-BFGS-B in the scaled variable y = x / sqrt(h0) reaches the same minimum
    sc = np.sqrt(h0)
    res = minimize(lambda y: (lambda F, g, _: (F, g * sc))(*fg(y * sc)), 1.0 / sc, jac=True, method="L-BFGS-B",
                   options=dict(maxiter=1000, gtol=1e-12, ftol=1e-16))
    assert abs(res.fun - F) <= 1e-10 * abs(F)
    assert np.abs(np.abs(res.x * sc) - np.abs(x)).max() <= 1e-4


def test_lbfgs_approximate_wolfe_carries_a_noisy_f(oracle):
    """With F perturbed by a deterministic noise of 1e-9 of its scale (the
    rounding of a lower-precision evaluation) and the gradient exact, the
    slope test keeps the line search decisive and the solve converges; with
    that test switched off (eps_f = -inf) Armijo alone stalls: it ends without
    converging, out of acceptable steps (status 2) or out of iterations on
    steps that F's noise accepts (status 1)."""
    s, w, conv, V = _setup(oracle)
    h0 = np.broadcast_to(1.0 / (2 * RHO0 * DV * w * KT), (NPTS, s["no"])).ravel()

    def fg(x):
        (F, *parts), g = oracle.functional(conv, x.reshape(NPTS, -1), V, w, RHO0, PREF, KT, DV)
        scale = sum(abs(p) for p in parts)
        noise = 1e-9 * scale * np.sin(1e7 * float(np.sum(x[::97])))
        return F + noise, g.ravel(), scale

    x0 = np.ones(NPTS * s["no"])
    res = oracle.lbfgs(fg, x0, h0, m=5, max_iter=400, gtol=1e-7)
    assert res[6] == 0
    res = oracle.lbfgs(fg, x0, h0, m=5, max_iter=400, gtol=1e-7, eps_f=-np.inf)
    assert res[6] in (1, 2) and np.abs(h0 * res[2]).max() > 1e-7


# ---- the library (C ABI) against the oracle -----------------------------------
def test_capi_rotation_matrices_match_oracle(oracle):
    for mmax, mrso in ((3, 2), (5, 2), (2, 1)):
        ag = M.AngularGrid(mmax, mrso)
        assert np.abs(M.rotation_matrices(ag) - oracle.rotation_matrices(mmax, mrso)).max() <= 1e-15


def _gpu_problem(oracle, n=N3, L=L3, mmax=MMAX, amp=1.0, nsite=3, charge_scale=0.1, vmax_kt=30.0, precision=None):
    import torch

    prec = M.FP64 if precision is None else precision
    sg, ag = M.SpatialGrid(n, L), M.AngularGrid(mmax, MRSO)
    nq = luc.nq_for(luc.qmax_of_grid(n, L), luc.DQ_TIP3P)
    tab = luc.model_c_symmetric(mmax, MRSO, nq, seed=5, amplitude=amp)
    conv = M.Convolution.from_table(tab, sg, ag, M.OrientationProjectionMap(ag), M.FP64 if prec == M.FP64 else M.FP32)
    solver = M.Solver(conv, RHO0, PREF, prec)
    sol = solute.synthetic_solute(nsite, 3, L, radius=2.0, charge_scale=charge_scale)
    solv = solute.SOLVENT_SITES["tip3p"]
    V = M.vext(sg, ag, sol, solv, vmax_kt * KT)
    fun = M.Functional(solver, V, KT)
    oconv = oracle.Conv(n, L, mmax, MRSO, tab.tuples, tab.c, tab.dq)
    return dict(sg=sg, ag=ag, solver=solver, V=V, fun=fun, oconv=oconv, sol=sol, solv=solv, tab=tab,
                dv=float(np.prod(L) / np.prod(n)), N=int(np.prod(n)), torch=torch)


@pytest.mark.gpu
@pytest.mark.parametrize("n,L,mmax,nsite,cs", [((9, 9, 9), (10.0,) * 3, 3, 3, 1.0),
                                               ((16, 12, 10), (14.0, 11.0, 9.5), 5, 20, 1.0),
                                               ((8, 8, 8), (10.0,) * 3, 2, 1, 0.1)])
def test_gpu_vext_matches_oracle(oracle, n, L, mmax, nsite, cs):
    sg, ag = M.SpatialGrid(n, L), M.AngularGrid(mmax, MRSO)
    sol = solute.synthetic_solute(nsite, 11, L, radius=min(L) / 3, charge_scale=cs)
    solv = solute.SOLVENT_SITES["tip3p"]
    vmax = 100 * KT
    V = M.vext(sg, ag, sol, solv, vmax).cpu().numpy()
    ref = oracle.vext(n, L, mmax, MRSO, sol, solv, vmax)
    assert np.abs(V - ref).max() <= 1e-12 * np.abs(ref).max()
    assert (V == vmax).sum() == (ref == vmax).sum()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(9, 9, 9), (8, 8, 8)])
def test_gpu_functional_eval_matches_oracle(oracle, n):
    P = _gpu_problem(oracle, n=n)
    torch = P["torch"]
    s = oracle.angular_sizes(MMAX, MRSO)
    xi = (1.0 + 0.1 * np.random.default_rng(7).standard_normal((P["N"], s["no"])))
    g = torch.empty((P["N"], s["no"]), dtype=torch.float64, device="cuda")
    F = P["fun"].eval(torch.from_numpy(xi).cuda(), g)
    Vh = P["V"].cpu().numpy()
    (Fr, *parts), gr = oracle.functional(P["oconv"], xi, Vh, P["ag"].m_w, RHO0, PREF, KT, P["dv"])
    for a, b in zip(F, (Fr, *parts)):
        assert abs(a - b) <= 1e-10 * max(abs(b), abs(Fr))
    assert np.abs(g.cpu().numpy() - gr).max() <= 1e-10 * np.abs(gr).max()


@pytest.mark.gpu
def test_gpu_lbfgs_trajectory_matches_oracle(oracle):
    """Ten iterations with gtol = 0 and the converged solve: the F history,
    the point and the counts agree with the oracle's two-loop L-BFGS."""
    P = _gpu_problem(oracle)
    torch = P["torch"]
    w = P["ag"].m_w
    Vh = P["V"].cpu().numpy()
    no = P["ag"].m_no
    h0 = np.broadcast_to(1.0 / (2 * RHO0 * P["dv"] * w * KT), (P["N"], no)).ravel()

    def fg(x):
        (F, *parts), g = oracle.functional(P["oconv"], x.reshape(P["N"], -1), Vh, w, RHO0, PREF, KT, P["dv"])
        return F, g.ravel(), sum(abs(p) for p in parts)

    x0 = np.ones(P["N"] * no)
    ref = oracle.lbfgs(fg, x0, h0, m=5, max_iter=10, gtol=0.0)
    xi = torch.ones((P["N"], no), dtype=torch.float64, device="cuda")
    r = P["fun"].minimize(xi, m=5, max_iter=10, gtol=0.0)
    assert r["iterations"] == ref[3] == 10 and r["evaluations"] == ref[4]
    assert np.abs(r["history"] - np.asarray(ref[5])).max() <= 1e-10 * abs(ref[5][0])
    assert np.abs(xi.cpu().numpy().ravel() - ref[0]).max() <= 1e-8
    # converged
    ref = oracle.lbfgs(fg, x0, h0, m=5, max_iter=300, gtol=1e-6)
    xi = torch.ones((P["N"], no), dtype=torch.float64, device="cuda")
    r = P["fun"].minimize(xi, m=5, max_iter=300, gtol=1e-6)
    assert r["status"] == ref[6] == 0 and r["residual"] < 1e-6
    assert abs(r["iterations"] - ref[3]) <= 2
    assert abs(r["F"][0] - ref[1]) <= 1e-10 * abs(ref[1])


@pytest.mark.gpu
def test_gpu_solve_from_vext_guess_matches_oracle(oracle):
    """xi = exp(-V / (2 kT)) (the minimum of F_id + F_ext alone) against
    numpy, and the FP64 solve from it against the oracle's L-BFGS from the
    same start: the same minimum as the solve from xi = 1, in fewer
    iterations."""
    P = _gpu_problem(oracle)
    torch = P["torch"]
    w = P["ag"].m_w
    Vh = P["V"].cpu().numpy()
    no = P["ag"].m_no
    xg = M.xi_from_vext(P["V"], KT)
    ref_x = np.exp(-0.5 * Vh / KT)
    assert np.abs(xg.cpu().numpy() - ref_x).max() <= 4e-16 * np.abs(ref_x).max()
    # F_id + F_ext is stationary there: its gradient 2 rho0 xi (kT ln xi^2 + V) vanishes
    assert np.abs(KT * np.log(ref_x ** 2) + Vh).max() <= 1e-12 * np.abs(Vh).max()
    h0 = np.broadcast_to(1.0 / (2 * RHO0 * P["dv"] * w * KT), (P["N"], no)).ravel()

    def fg(x):
        (F, *parts), g = oracle.functional(P["oconv"], x.reshape(P["N"], -1), Vh, w, RHO0, PREF, KT, P["dv"])
        return F, g.ravel(), sum(abs(p) for p in parts)

    ref = oracle.lbfgs(fg, ref_x.ravel().copy(), h0, m=5, max_iter=300, gtol=1e-6)
    xi = xg.clone()
    r = P["fun"].minimize(xi, m=5, max_iter=300, gtol=1e-6)
    assert r["status"] == ref[6] == 0 and r["residual"] < 1e-6
    assert abs(r["iterations"] - ref[3]) <= 2
    assert abs(r["F"][0] - ref[1]) <= 1e-10 * abs(ref[1])
    one = torch.ones((P["N"], no), dtype=torch.float64, device="cuda")
    r1 = P["fun"].minimize(one, m=5, max_iter=300, gtol=1e-6)
    assert r1["status"] == 0 and r["iterations"] < r1["iterations"]
    assert abs(r["F"][0] - r1["F"][0]) <= 1e-7 * abs(r1["F"][0])


@pytest.mark.gpu
def test_gpu_lbfgs_c0_limit(oracle):
    P = _gpu_problem(oracle, amp=0.0, vmax_kt=10.0)
    torch = P["torch"]
    xi = torch.ones((P["N"], P["ag"].m_no), dtype=torch.float64, device="cuda")
    r = P["fun"].minimize(xi, m=5, max_iter=300, gtol=1e-10)
    assert r["status"] == 0
    exact = torch.exp(-P["V"] / (2 * KT))
    assert float((xi.abs() - exact).abs().max()) <= 1e-9
    # F decreases, up to the approximate Wolfe test's allowance of 1e-6 of its scale
    tol = 2e-6 * sum(abs(p) for p in r["F"][1:])
    assert all(b <= a + tol for a, b in zip(r["history"], r["history"][1:]))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [M.MIXED_X6, M.MIXED, M.MIXED_X6F])
def test_gpu_lbfgs_mixed_vs_fp64(oracle, prec):
    """A mixed-form solve reaches the FP64 minimum: |dF| <= 0.01 kJ/mol, and
    the FP64 gradient at the mixed minimiser is small (its relative L2 against
    the gradient at the start is stated)."""
    P64 = _gpu_problem(oracle, n=(16, 16, 16), L=(12.0,) * 3, nsite=5)
    Pm = _gpu_problem(oracle, n=(16, 16, 16), L=(12.0,) * 3, nsite=5, precision=prec)
    torch = P64["torch"]
    shape = (P64["N"], P64["ag"].m_no)
    x64 = torch.ones(shape, dtype=torch.float64, device="cuda")
    r64 = P64["fun"].minimize(x64, max_iter=300, gtol=1e-5)
    xm = torch.ones(shape, dtype=torch.float64, device="cuda")
    rm = Pm["fun"].minimize(xm, max_iter=300, gtol=1e-5)
    assert r64["status"] == 0 and rm["status"] == 0
    assert abs(rm["F"][0] - r64["F"][0]) <= 0.01
    g = torch.empty(shape, dtype=torch.float64, device="cuda")
    F_at_m = P64["fun"].eval(xm, g)
    g0 = torch.empty(shape, dtype=torch.float64, device="cuda")
    P64["fun"].eval(torch.ones(shape, dtype=torch.float64, device="cuda"), g0)
    assert abs(F_at_m[0] - r64["F"][0]) <= 0.01
    assert float(g.norm() / g0.norm()) <= 1e-3


@pytest.mark.gpu
def test_gpu_functional_errors(oracle):
    import torch

    sg, ag = M.SpatialGrid((4, 4, 4), (5.0,) * 3), M.AngularGrid(2, MRSO)
    tab = luc.model_c_symmetric(2, MRSO, luc.nq_for(luc.qmax_of_grid((4, 4, 4), (5.0,) * 3), luc.DQ_TIP3P), 1)
    V = M.vext(sg, ag, solute.single_ion(0.1, 2.5, 0.5, (5.0,) * 3), solute.SOLVENT_SITES["spce"], 50.0)
    conv32 = M.Convolution.from_table(tab, sg, ag, M.OrientationProjectionMap(ag), M.FP32)
    with pytest.raises(M.MDFTError, match="FP32"):
        M.Functional(M.Solver(conv32, RHO0, PREF, M.FP32), V, KT)
    conv = M.Convolution.from_table(tab, sg, ag, M.OrientationProjectionMap(ag), M.FP64)
    fun = M.Functional(M.Solver(conv, RHO0, PREF, M.FP64), V, KT)
    xi = torch.ones((64, ag.m_no), dtype=torch.float64, device="cuda")
    for m in (0, 9):
        with pytest.raises(M.MDFTError, match="m must be"):
            fun.minimize(xi, m=m)
    with pytest.raises(M.MDFTError, match="bad sites"):
        M.vext(sg, ag, np.zeros((1, 6)), np.zeros((0, 6)), 1.0)


def test_lbfgs_nonfinite_start_ends_at_once(oracle):
    s, w, conv, V = _setup(oracle)
    h0 = np.broadcast_to(1.0 / (2 * RHO0 * DV * w * KT), (NPTS, s["no"])).ravel()
    def fg(x):
        (F, *parts), g = oracle.functional(conv, x.reshape(NPTS, -1), V, w, RHO0, PREF, KT, DV)
        return F, g.ravel(), sum(abs(p) for p in parts)
    x0 = np.ones(NPTS * s["no"])
    x0[123] = np.nan
    res = oracle.lbfgs(fg, x0, h0)
    assert res[6] == 4 and res[3] == 0 and res[4] == 1


@pytest.mark.gpu
def test_gpu_lbfgs_nonfinite_start_ends_at_once(oracle):
    """A NaN injected into xi: the solve ends after the start evaluation with
    status 4 (no line search on NaN); the server refuses such inputs up front."""
    P = _gpu_problem(oracle)
    torch = P["torch"]
    xi = torch.ones((P["N"], P["ag"].m_no), dtype=torch.float64, device="cuda")
    xi[5, 7] = float("nan")
    r = P["fun"].minimize(xi, m=5, max_iter=50, gtol=1e-6)
    assert r["status"] == 4 and r["iterations"] == 0 and r["evaluations"] == 1
    # the server refuses non-finite sites and a NaN cap at submission
    tab = luc.model_c_symmetric(MMAX, MRSO, NQ + 4, seed=5)
    srv = M.Server(tab, MRSO, N0, TEMP, M.FP64, capacity=1)
    bad = solute.synthetic_solute(2, 3, L3, radius=2.0, charge_scale=0.1)
    bad[1, 0] = float("nan")
    with pytest.raises(M.MDFTError, match="finite"):
        srv.submit(N3, L3, MMAX, bad, solute.SOLVENT_SITES["tip3p"], 30 * KT)
    with pytest.raises(M.MDFTError, match="vmax"):
        srv.submit(N3, L3, MMAX, bad[:1], solute.SOLVENT_SITES["tip3p"], float("nan"))
