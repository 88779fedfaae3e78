"""Pins of the oracle's nonlinear constraint rows, Schur complement and constrained step (CPU):
the reference's constraint fixtures (values Φ and Jacobians J, ``RS/tests/oracle.rs:332-375``,
tolerances PHI = JACOBIAN = 1e-12 at :111-112), central differences of the area and volume, and
the reference's constrained step fixtures replayed through the oracle session
(``RS/tests/oracle.rs:557-662``, SOLVE = 1e-10 at :113)."""
import numpy as np
import pytest

import oracle
from golden import fixtures as fx

REL = lambda a, b: abs(a - b) / abs(b)


def _targets(case):
    """The nonlinear rows of a case in order (kind, target or None) and whether a barycenter row
    is present."""
    rows, bary = [], False
    for c in case["constraints"]:
        if c[0] in ("area", "volume"):
            assert c[1] == "all"
            rows.append((c[0], fx.hexf(c[2]) if len(c) > 2 else None))
        elif c[0] == "barycenter":
            assert c[1:] == ["all"]
            bary = True
    return rows, bary


@pytest.mark.parametrize("name", ["rs-icosa-cons-area-volume", "rs-mobius-cons-area", "rs-patch-cons-open"])
def test_lean_constraint_fixtures(name):
    c, o = fx.read_case(name), fx.read_out(name)
    p, f = c["positions"], c["faces"]
    nv = p.shape[0]
    rows, _ = _targets(c)
    kinds = [k for k, _ in rows]
    vals, grads = oracle.constraint_rows(f, p, kinds)
    J = fx.matrix(o, "J")
    targets = [t if t is not None else v for (_, t), v in zip(rows, vals)]
    scale = max([1.0] + [abs(t) for t in targets])
    va = oracle.vertex_areas(f, p)
    r = 0
    for idx, kind, mode, want_phi in o["crows"]:
        if kind == "barycenter":
            # Φ = 0 at the target's own positions; J = a_v / Σ a on the row's axis
            axis = sum(1 for i, k2, _, _ in o["crows"] if k2 == "barycenter" and i < idx)
            want = np.zeros(3 * nv)
            want[axis::3] = va / va.sum()
            assert mode == "positional" and want_phi == 0.0
            assert np.max(np.abs(J[idx] - want)) <= 1e-12 * np.max(np.abs(want))
            continue
        assert mode == "nonlinear" and kind == kinds[r]
        assert abs((vals[r] - targets[r]) - want_phi) <= 1e-12 * scale, (name, kind)
        assert np.max(np.abs(grads[r] - J[idx])) <= 1e-12 * max(1.0, np.max(np.abs(J[idx]))), (name, kind)
        r += 1
    assert r == len(kinds)


def test_area_and_volume_gradients_are_derivatives():
    """Central differences of the total area and the signed volume against their gradients, and
    the closed forms on a cube-like icosphere: the volume gradient sums to zero (translation) and
    Σ_v f_v·∇_v V = 3 V (homogeneity of degree 3), Σ_v f_v·∇_v A = 2 A."""
    from conftest import load_package
    g = load_package().generate
    p, f = g.icosphere(1)
    p = g.normal_jitter(p, f, 0.05, 3)
    vals, grads = oracle.constraint_rows(f, p, ["area", "volume"])
    h = 1e-6
    rng = np.random.default_rng(0)
    for i in rng.choice(3 * p.shape[0], 12, replace=False):
        q1, q2 = p.copy(), p.copy()
        q1.reshape(-1)[i] += h
        q2.reshape(-1)[i] -= h
        d = (oracle.constraint_rows(f, q1, ["area", "volume"])[0] - oracle.constraint_rows(f, q2, ["area", "volume"])[0]) / (2 * h)
        assert np.all(np.abs(d - grads[:, i]) <= 1e-7 * max(1.0, np.abs(grads[:, i]).max()))
    x = p.reshape(-1)
    assert abs(x @ grads[0] - 2 * vals[0]) <= 1e-12 * vals[0]
    assert abs(x @ grads[1] - 3 * vals[1]) <= 1e-12 * vals[1]
    assert np.abs(grads[1].reshape(-1, 3).sum(axis=0)).max() <= 1e-12 * np.abs(grads[1]).max()


def constrained_session(c):
    rows, bary = _targets(c)
    assert bary
    start = c["start"]
    s = ("maxdisp", start[1]) if start[0] == "maxdisp" else ("abs", start[1]) if start[0] == "abs" else ("reference",)
    return oracle.Session(c["faces"], c["positions"], *c["exponents"], chi=0.0, tol=1e-13, maxit=500,
                          energy_exact=True, start=s, armijo_c=c["armijo"], max_halvings=c["maxhalv"],
                          constraints=rows, newton=c["newton"])


@pytest.mark.parametrize("name", ["rs-icosa-step", "rs-mobius-step", "rs-torus-step-ref"])
def test_lean_constrained_steps(name):
    """The step with nonlinear rows: the Schur direction (energy0, slope, ‖x‖), the trials, the
    Newton projection (rounds, Φ after it) and the final energy and positions."""
    c, o = fx.read_case(name), fx.read_out(name)
    assert c["weights"] == (1.0, 1.0, 0.0)
    s = constrained_session(c)
    rep = s.step(1)
    print(name, {k: rep[k] for k in ("energy0", "slope", "direction_norm", "tau0", "newton_rounds", "min_pivot")})
    assert REL(rep["energy0"], o["energy0"]) <= 1e-10
    assert REL(rep["slope"], o["slope"]) <= 1e-10
    assert REL(rep["direction_norm"], o["xnorm"]) <= 1e-10
    assert REL(rep["lbar"], o["lbar"]) <= 1e-12
    assert REL(rep["tau0"], o["tau0"]) <= 1e-10
    want = {"accepted": 0, "armijo-failed": 1, "non-finite": 2}
    assert len(rep["trial_list"]) == len(o["trials"])
    for (tau, fv, outcome), (_, wtau, wout, wf, _) in zip(rep["trial_list"], o["trials"]):
        assert outcome == want[wout]
        assert REL(tau, wtau) <= 1e-10
        assert REL(fv, wf) <= 1e-10
    assert o["result"][0] == "accepted" and rep["accepted"] == 1.0 and rep["halvings"] == o["result"][2]
    assert REL(rep["energy1"], o["energy1"]) <= 1e-10
    assert rep["newton_rounds"] == o["newton"]
    # (the fixture's `pivot` is the smallest over the largest LU pivot of the whole saddle matrix
    # in the Lean oracle, which neither the reference's runner nor this path forms)
    assert 0.0 < rep["min_pivot"] <= 1.0
    rows, _ = _targets(c)
    vals0, _ = oracle.constraint_rows(c["faces"], c["positions"], [k for k, _ in rows])
    scale = max([1.0] + [abs(t if t is not None else v) for (_, t), v in zip(rows, vals0)])
    assert np.max(np.abs(np.array(rep["phi"]) - np.array(o["phi"]))) <= 1e-10 * scale
    assert np.max(np.abs(np.array(rep["phi_before"]) - np.array(o["phi-before"]))) <= 1e-10 * scale
    assert REL(rep["energy_after"], o["energy-final"]) <= 1e-10
    want_x = np.array([o["rows"]["x"][i] for i in range(len(o["rows"]["x"]))])
    assert np.max(np.abs(s.positions() - want_x)) <= 1e-10


def test_lean_constrained_step_catches_a_missing_newton_projection():
    """Without the Newton rounds the area and volume rows stay off by the step's drift (1e-3),
    eight orders above the fixture's Φ after the projection: the Φ records pin the rounds."""
    c, o = fx.read_case("rs-icosa-step"), fx.read_out("rs-icosa-step")
    c = dict(c)
    c["newton"] = ("fixed", 0)
    rep = constrained_session(c).step(1)
    assert rep["newton_rounds"] == 0
    assert np.max(np.abs(np.array(rep["phi"][3:]) - np.array(o["phi"][3:]))) > 1e-4


@pytest.mark.parametrize("name", ["rs-icosa-saddle", "rs-mobius-saddle", "rs-torus-saddle"])
def test_lean_saddle_with_schur_rows(name):
    """The reference's saddle fixtures with nonlinear rows (RS/tests/oracle.rs:377-419): with the
    exact metric's FP64 solves z_r = A′⁻¹ [C_N row r; 0] and y = A′⁻¹ [g; 0], the oracle's Schur
    complement and the step's own (``ddgrs_schur``, host code) give the fixture's S, the direction
    x, ⟨g, x⟩ and the Newton correction h to SOLVE = 1e-10."""
    from conftest import load_package
    pkg = load_package()
    c, o = fx.read_case(name), fx.read_out(name)
    p, f = c["positions"], c["faces"]
    nv = p.shape[0]
    rows, bary = _targets(c)
    assert bary and rows and all(t is not None for _, t in rows)
    w = c["weights"]
    sd = oracle.Saddle(f, p, *c["exponents"], chi=0.0, c_lo=1.0 if w[1] != 0.0 else 0.0)
    vals, C = oracle.constraint_rows(f, p, [k for k, _ in rows])
    phin = vals - np.array([t for _, t in rows])
    assert np.max(np.abs(phin - np.array(o["phin"]))) <= 1e-10
    _, g, _ = oracle.exact(f, p, *c["exponents"])
    g = g.reshape(-1)
    want_g = np.array([o["rows"]["g"][i] for i in range(nv)]).reshape(-1)
    assert np.abs(g - want_g).max() <= 1e-12 * np.abs(want_g).max()

    def solve(b):
        rhs = np.zeros(sd.n)
        rhs[: 3 * nv] = b
        x, rep = sd.solve(rhs, tol=1e-13, maxit=500)
        assert rep["converged"] == 1.0
        return x[: 3 * nv]

    Z = np.stack([solve(C[r]) for r in range(C.shape[0])])
    y = solve(g)
    want_x = np.array([o["rows"]["x"][i] for i in range(nv)]).reshape(-1)
    want_h = np.array([o["rows"]["h"][i] for i in range(nv)]).reshape(-1)
    want_S = fx.matrix(o, "S")
    for label, impl in (("oracle", oracle.schur), ("library", pkg.schur)):
        x, h, S, pivot = impl(C, Z, y, phin)
        print(name, label, "S", S.tolist(), "pivot", pivot)
        assert np.abs(S - want_S).max() <= 1e-10 * np.abs(want_S).max(), label
        assert np.abs(x - want_x).max() <= 1e-10 * np.abs(want_x).max(), label
        assert abs(g @ x - o["gx"]) <= 1e-10 * np.linalg.norm(g) * np.linalg.norm(x), label
        assert np.abs(h - want_h).max() <= 1e-10 * np.abs(want_h).max(), label
        # C_N x = 0 on the kept rows (the fixture's cnx is round-off)
        assert np.abs(C @ x).max() <= 1e-10 * np.abs(C).max() * np.abs(x).max()
