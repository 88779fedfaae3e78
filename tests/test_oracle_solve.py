"""Pins the FP64 oracle's saddle solver, preconditioner, GMRES and descent step (CPU).

* Lean golden ``rs-icosa-precond``: ε and z = L′⁻¹ blockdiag(L_{2−s} ⊗ I₃, 0) L′⁻¹ r with the
  fixture's own L_{2−s} (tests/oracle.rs:45-48: 1e-10) — pins W, M, ε, the barycenter row and
  the L′ solve;
* ``fractional_symbol_ratio`` = 4 I(σ) I(2−s) ≈ 200 at s = 5/3 (solve/precond.rs:512-564);
* the augmented preconditioner reproduces the constraint rows, C_Q (P⁻¹ [r; b])_x = b, and is
  symmetric (precond.rs:596-612);
* GMRES's solution satisfies A′ x = b to its tolerance, checked against an independent dense
  solve (gmres.rs:413);
* Lean golden ``rs-octa-saddle-none``: the direction x = A′⁻¹ [g; 0] with the dense metric
  (χ = 0, w = (1, 1), (4, 8), barycenter rows) and ⟨g, x⟩ — pins A′, the constraint rows and GMRES;
* ε and L′ against an independent numpy assembly (cotan W, lumped M, barycenter row);
* Lean golden ``rs-icosa-step-stuck``: one step of the session (exact energy, the dense metric
  χ = 0, the barycenter only, an absolute start τ₀ = 1e6, at most three halvings) reproduces F₀,
  ⟨g, x⟩, ‖x‖, ℓ̄₀, τ₀, every trial's τ, outcome and F, the stuck result with four halvings and F
  after the projection (tests/oracle.rs:557-640, 1e-10; ℓ̄₀ 1e-12) — the trial energies at
  τ ≥ 1.25e5 are F(f − τ x) with τ x dominating, so they pin the direction's shape;
* one descent step lowers the energy and keeps the barycenter.
"""
import math

import numpy as np
import pytest

import oracle
from golden import fixtures as fx


def test_lean_precond_icosa():
    c, o = fx.read_case("rs-icosa-precond"), fx.read_out("rs-icosa-precond")
    sd = oracle.Saddle(c["faces"], c["positions"], *c["exponents"], chi=0.0, eps_rel=c["eps_rel"])
    assert abs(sd.eps - o["eps"]) <= 1e-12 * o["eps"]
    nv = len(c["positions"])
    r = np.zeros(3 * nv + 3)
    r[: 3 * nv] = np.array(c["r"]).reshape(-1)
    lfr = fx.matrix(o, "Lfr")
    y = sd.lprime_solve(r)
    v = np.zeros_like(y)
    for k in range(3):
        v[k: 3 * nv: 3] = lfr @ y[k: 3 * nv: 3]
    z = sd.lprime_solve(v)
    zref = np.array([o["rows"]["z"][i] for i in range(nv)]).reshape(-1)
    assert np.abs(z[: 3 * nv] - zref).max() <= 1e-10 * np.abs(zref).max()


def test_lean_saddle_octa():
    c, o = fx.read_case("rs-octa-saddle-none"), fx.read_out("rs-octa-saddle-none")
    assert c["weights"] == (1.0, 1.0, 0.0)
    p, f = np.array(c["positions"]), np.array(c["faces"])
    nv = len(p)
    g = np.array([o["rows"]["g"][i] for i in range(nv)]).reshape(-1)
    x = np.array([o["rows"]["x"][i] for i in range(nv)]).reshape(-1)
    _, d, _ = oracle.exact(f, p, *c["exponents"])
    assert np.abs(d.reshape(-1) - g).max() <= 1e-12 * np.abs(g).max()
    sd = oracle.Saddle(f, p, *c["exponents"], chi=0.0)
    b = np.zeros(sd.n)
    b[: 3 * nv] = g
    xs, rep = sd.solve(b, tol=1e-13, maxit=500)
    assert rep["converged"] == 1.0
    assert np.abs(xs[: 3 * nv] - x).max() <= 1e-10 * np.abs(x).max()
    assert abs(g @ xs[: 3 * nv] - o["gx"]) <= 1e-10 * abs(o["gx"])


def test_lprime_matches_independent_assembly():
    from conftest import load_package, lprime_axis_matrix, lprime_residual

    p, f = load_package().generate.config_icosphere_l3()
    sd = oracle.Saddle(f, p)
    L, eps = lprime_axis_matrix(f, p, oracle.vertex_areas(f, p))
    assert abs(sd.eps - eps) <= 1e-13 * eps
    r = np.random.default_rng(21).standard_normal(sd.n)
    # L′'s last pivot is −cᵀK⁻¹c ~ 1/ε: solves keep about 1e-10 relative residual (precond.rs:16-18)
    assert lprime_residual(L, r, sd.lprime_solve(r)) <= 1e-8


def test_symbol_ratio_closed_form():
    s = 5.0 / 3.0

    def I(t):
        return math.pi * math.gamma(1 - t) / (t * 4.0**t * math.gamma(1 + t))

    assert abs(oracle.fractional_symbol_ratio(s) - 4 * I(s - 1) * I(2 - s)) <= 1e-13 * 200
    assert 199.0 < oracle.fractional_symbol_ratio(s) < 201.0


@pytest.fixture(scope="module")
def small():
    from conftest import load_package

    g = load_package().generate
    p, f = g.icosphere(2)
    return g.normal_jitter(p, f, 0.02, 9), f


def test_augmented_preconditioner_structure(small):
    p, f = small
    sd = oracle.Saddle(f, p, chi=0.5)
    nv = len(p)
    rng = np.random.default_rng(2)
    r = rng.standard_normal(sd.n)
    z = sd.precond(r)
    va = oracle.vertex_areas(f, p)
    crow = va / va.sum()
    # L′ has a near-null space of size ε = 1e-10 tr W / tr M; its unrefined solves keep about
    # 1e-8 relative accuracy on it (precond.rs:16-18), hence the tolerance.
    for k in range(3):
        assert abs(crow @ z[k: 3 * nv: 3] - r[3 * nv + k]) <= 1e-7 * np.abs(r).max()
    u = rng.standard_normal(sd.n)
    assert abs(u @ sd.precond(r) - r @ sd.precond(u)) <= 1e-7 * np.abs(u) @ np.abs(sd.precond(r))


def test_gmres_against_dense_solve(small):
    p, f = small
    sd = oracle.Saddle(f, p, chi=0.5)
    n = sd.n
    A = np.stack([sd.apply(e) for e in np.eye(n)], axis=1)
    e0, d, _ = oracle.exact(f, p)
    rhs = np.zeros(n)
    rhs[:-3] = d.reshape(-1)
    x, rep = sd.solve(rhs, restart=50, tol=1e-10, maxit=400)
    assert rep["converged"] == 1.0
    xd = np.linalg.solve(A, rhs)
    assert np.linalg.norm(x - xd) <= 1e-7 * np.linalg.norm(xd)
    assert np.linalg.norm(A @ x - rhs) <= 1e-9 * np.linalg.norm(rhs)


def test_gmres_history_is_the_minimal_residual(small):
    """GmresReport::history (gmres.rs:217): after k iterations of the first cycle the Arnoldi
    estimate is min over y in K_k(A P^-1, b) of ||b - A P^-1 y|| / ||b||, computed here by least
    squares on an orthonormal Krylov basis of the dense A P^-1; cycle_residuals (gmres.rs:214) start
    at 1, each later one is the explicit residual that the previous cycle's last estimate predicts,
    and the last is the reported relative residual."""
    p, f = small
    sd = oracle.Saddle(f, p, chi=0.5)
    n = sd.n
    eye = np.eye(n)
    A = np.stack([sd.apply(e) for e in eye], axis=1)
    Pinv = np.stack([sd.precond(e) for e in eye], axis=1)
    M = A @ Pinv
    _, d, _ = oracle.exact(f, p)
    b = np.zeros(n)
    b[:-3] = d.reshape(-1)
    _, rep = sd.solve(b, restart=50, tol=1e-8, maxit=50)
    hist = rep["history"]
    assert len(hist) == rep["iterations"] >= 5
    Q = (b / np.linalg.norm(b))[:, None]
    # P^-1 has the near-null space of L' (eps = 1e-10 tr W / tr M), so the dense A P^-1 carries
    # rounding of about 1e-10 relative to its norm: compared while the residual is above 1e-6
    checked = 0
    for k, est in enumerate(hist[:12], 1):
        MQ = M @ Q
        c = np.linalg.lstsq(MQ, b, rcond=None)[0]
        want = np.linalg.norm(b - MQ @ c) / np.linalg.norm(b)
        if want < 1e-6:
            break
        assert abs(est - want) <= 1e-5 * want, (k, est, want)
        checked += 1
        v = M @ Q[:, -1]
        for _ in range(2):
            v -= Q @ (Q.T @ v)
        Q = np.hstack([Q, (v / np.linalg.norm(v))[:, None]])
    assert checked >= 5
    # restarted: 3 cycles of 4 iterations
    _, rep = sd.solve(b, restart=4, tol=1e-12, maxit=12)
    cyc, hist = rep["cycle_residuals"], rep["history"]
    assert rep["cycles"] == 3 and len(hist) == 12 and len(cyc) == 4
    assert cyc[0] == 1.0 and cyc[-1] == rep["relative_residual"]
    for c in range(1, 4):
        assert abs(cyc[c] - hist[4 * c - 1]) <= 1e-7 * cyc[c] + 1e-15, (c, cyc[c], hist[4 * c - 1])
    assert all(h1 <= h0 * (1 + 1e-12) for h0, h1 in zip(hist, hist[1:]))


def test_descent_step_lowers_energy_and_keeps_barycenter(small):
    p, f = small
    s = oracle.Session(f, p, theta=0.5)
    va = oracle.vertex_areas(f, p)
    b0 = (va[:, None] * p).sum(axis=0) / va.sum()
    rep = s.step(2)
    assert rep["accepted"] == 1.0 and rep["gmres_converged"] == 1.0
    assert rep["energy1"] < rep["energy0"]
    q = s.positions()
    vq = oracle.vertex_areas(f, q)
    b1 = (vq[:, None] * q).sum(axis=0) / vq.sum()
    assert np.abs(b1 - b0).max() <= 1e-14


def step_stuck_session(fx_case):
    c = fx_case
    return oracle.Session(c["faces"], c["positions"], *c["exponents"], chi=0.0, tol=1e-13, maxit=500,
                          energy_exact=True, start=("abs", c["start"][1]), armijo_c=c["armijo"],
                          max_halvings=c["maxhalv"])


def test_lean_step_stuck_icosa():
    c, o = fx.read_case("rs-icosa-step-stuck"), fx.read_out("rs-icosa-step-stuck")
    assert c["weights"] == (1.0, 1.0, 0.0) and c["constraints"] == [["barycenter", "all"]]
    assert c["start"][0] == "abs" and c["newton"] == ("tol", 5)
    rep = step_stuck_session(c).step(1)
    rel = lambda a, b: abs(a - b) / abs(b)
    assert rel(rep["energy0"], o["energy0"]) <= 1e-10
    assert rel(rep["slope"], o["slope"]) <= 1e-10
    assert rel(rep["direction_norm"], o["xnorm"]) <= 1e-10
    assert rel(rep["lbar"], o["lbar"]) <= 1e-12
    assert rel(rep["tau0"], o["tau0"]) <= 1e-10
    want = {"accepted": 0, "armijo-failed": 1, "non-finite": 2}
    assert len(rep["trial_list"]) == len(o["trials"]) == 4
    for (tau, f, outcome), (_, wtau, wout, wf, _) in zip(rep["trial_list"], o["trials"]):
        assert outcome == want[wout]
        assert rel(tau, wtau) <= 1e-10
        assert rel(f, wf) <= 1e-10
    assert o["result"][0] == "stuck" and rep["accepted"] == 0.0
    assert rep["halvings"] == o["result"][2] == 4 and rep["tau"] == o["result"][1] == 0.0
    assert rel(rep["energy1"], o["energy1"]) <= 1e-10
    assert rel(rep["energy_after"], o["energy-final"]) <= 1e-10
    assert o["newton"] == 0 and o["phi"] == [0.0, 0.0, 0.0]


def test_lean_step_stuck_catches_a_wrong_halving_cap():
    """The same replay with the cap off by one takes a fifth trial (τ = 6.25e4) that the fixture
    does not have: the trial records pin the cap rule `halvings > max_halvings`."""
    c = dict(fx.read_case("rs-icosa-step-stuck"))
    c["maxhalv"] = c["maxhalv"] + 1
    rep = step_stuck_session(c).step(1)
    assert len(rep["trial_list"]) == 5


def test_preconditioner_variants(small):
    """PreconditionerOptions (solve/precond.rs:466-582): with y = L′⁻¹ r and v = [L^V y; 0],
    augmented P⁻¹_s r = L′⁻¹ ((s L^V − K) y) + y and congruence P⁻¹ r = L′⁻¹ v, so
    P⁻¹_symbol − P⁻¹_unit = (s − 1) P⁻¹_congruence (s the symbol scale); the congruence form is
    symmetric and its range meets the barycenter rows (C₁ z = 0, precond.rs:586-594); one step of
    refinement in the L′ solves changes the application by round-off only."""
    p, f = small
    sym = oracle.Saddle(f, p, chi=0.5)
    unit = oracle.Saddle(f, p, chi=0.5, scale="unit")
    cong = oracle.Saddle(f, p, chi=0.5, form="congruence")
    refd = oracle.Saddle(f, p, chi=0.5, refine=True)
    rng = np.random.default_rng(4)
    r, u = rng.standard_normal(sym.n), rng.standard_normal(sym.n)
    zs, zu, zc = sym.precond(r), unit.precond(r), cong.precond(r)
    scale = oracle.lib().orc_saddle_frac_scale(sym._h)
    assert scale != 1.0 and oracle.lib().orc_saddle_frac_scale(unit._h) == 1.0
    assert np.abs((zs - zu) - (scale - 1.0) * zc).max() <= 1e-9 * np.abs(zs - zu).max()
    assert abs(u @ cong.precond(r) - r @ cong.precond(u)) <= 1e-9 * np.abs(u) @ np.abs(zc)
    nv = sym.nv
    va = oracle.vertex_areas(f, p)
    for k in range(3):
        assert abs((va / va.sum()) @ zc[k: 3 * nv: 3]) <= 1e-9 * np.abs(zc).max()
    zr = refd.precond(r)
    assert not np.array_equal(zr, zs)
    assert np.abs(zr - zs).max() <= 1e-6 * np.abs(zs).max()
