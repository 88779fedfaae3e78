"""The step's saddle solver and the Hs descent step on the device against the FP64 oracle (GPU).

* L′⁻¹ (FP64 LDLᵀ solves on the device): L′'s last pivot is −cᵀK⁻¹c ~ 1/ε (ε = 1e-10 tr W/tr M),
  so any FP64 solve of it has a relative residual near 1e-10 and two solves with different
  elimination or summation orders differ by up to ~1e-8 (scipy's spsolve against the oracle:
  4.8e-9 on the level-3 icosphere).  The test asks for a residual against an independent scipy
  assembly of L′ within 10× the oracle's own, and a difference from the oracle ≤ 1e-7; ε and the
  fractional scale to 1e-12.
* Lean golden ``rs-octa-saddle-none``: the device direction x = A′⁻¹ [g; 0] (χ = 0, FP32 metric
  values) within 1e-5 of the fixture's x.
* A′ x and P⁻¹ r (FP32 hierarchical products inside): relative 2-norm error ≤ 1e-5.
* GMRES: converged to 1e-6 on the device's operator, iteration count within 1 of the oracle's, and
  the solution within 1e-4 (relative 2-norm) of the oracle's.
* The step: energy before ≤ 1e-5 relative (the Barnes–Hut energy on the same tree and θ), the
  accepted trial energy ≤ 1e-5, the same number of line-search trials, τ ≤ 1e-4 relative, moved
  positions ≤ 1e-4 of the displacement, the barycenter kept to 1e-12.
* Lean golden ``rs-icosa-step-stuck`` through the device session (exact energy, χ = 0, absolute
  start 1e6, three halvings): F₀ and every trial energy ≤ 1e-5 relative (FP32 pair kernels),
  ⟨g, x⟩ and ‖x‖ ≤ 1e-4 (GMRES to 1e-6 on FP32 metric values), τ₀ and each τ exactly, every
  outcome, the stuck result after four halvings, the positions unchanged.
* C1 proper: one exact-energy step on the level-3 icosphere (1,280 faces) against the oracle's
  exact-energy step: the same trial records, F before / after ≤ 1e-5, τ ≤ 1e-4.
"""
import numpy as np
import pytest

import oracle
from conftest import lprime_axis_matrix, lprime_residual
from golden import fixtures as fx

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def meshes(pkg):
    return {
        "icosphere_l3": pkg.generate.config_icosphere_l3,
        "torus_16k": pkg.generate.config_torus_16k,
    }


@pytest.mark.parametrize("name", ["icosphere_l3", "torus_16k"])
def test_saddle_parts(pkg, ctx, name):
    p, f = meshes(pkg)[name]()
    s = pkg.Surface(ctx, f, p)
    sd = pkg.Saddle(s)
    od = oracle.Saddle(f, p)
    info = sd.info()
    assert abs(info["eps"] - od.eps) <= 1e-12 * od.eps
    assert abs(info["fractional_scale"] - od.frac_scale) <= 1e-12 * od.frac_scale
    n = od.n
    rng = np.random.default_rng(21)
    x = rng.standard_normal((2, n))
    L, _ = lprime_axis_matrix(f, p, oracle.vertex_areas(f, p))
    e_l, res, res0 = 0.0, 0.0, 0.0
    for q in range(2):
        xl, xl0 = sd.lprime_solve(x[q]), od.lprime_solve(x[q])
        e_l = max(e_l, rel(xl, xl0))
        res, res0 = max(res, lprime_residual(L, x[q], xl)), max(res0, lprime_residual(L, x[q], xl0))
    e_a = max(rel(sd.apply_batch(x[q]), od.apply(x[q])) for q in range(2))
    e_p = max(rel(sd.precondition(x[q]), od.precond(x[q])) for q in range(2))
    print(f"{name}: L'^-1 {e_l:.2e} (residual {res:.2e}, oracle {res0:.2e})  A' {e_a:.2e}  P^-1 {e_p:.2e}")
    assert res <= max(10.0 * res0, 1e-13)
    assert e_l <= 1e-7
    assert e_a <= 1e-5
    assert e_p <= 1e-5


@pytest.mark.parametrize(
    "form,scale,refine",
    [("congruence", "symbol", False), ("augmented", "unit", False), ("augmented", "symbol", True), ("congruence", "unit", True)],
)
def test_preconditioner_variants(pkg, ctx, form, scale, refine):
    """PreconditionerOptions' form, fractional scale and refinement (precond.rs:466) on the device
    against the oracle's restatement of the same variant, and GMRES with the variant reaching the
    default preconditioner's solution."""
    p, f = pkg.generate.config_icosphere_l3()
    s = pkg.Surface(ctx, f, p)
    sd = pkg.Saddle(s, options=pkg.SaddleOptions(form=form, scale=scale, refine=refine))
    od = oracle.Saddle(f, p, form=form, scale=scale, refine=refine)
    base = pkg.Saddle(s)
    info = sd.info()
    want_scale = 1.0 if scale == "unit" else base.info()["fractional_scale"]
    assert info["fractional_scale"] == want_scale
    rng = np.random.default_rng(31)
    x = rng.standard_normal((2, od.n))
    e_p = max(rel(sd.precondition(x[q]), od.precond(x[q])) for q in range(2))
    # the step's right-hand side [g; 0] with Σ_v g_v = 0, as a translation-invariant energy's
    # differential has: then the multipliers of the solution vanish (λ_a = Σ_v g_va, since B̃ kills
    # translations and Σ c_v = 1).  The congruence form's range has no multiplier component, so
    # right-preconditioned GMRES with it cannot reach any other right-hand side (precond.rs:586-594).
    b = x[0].copy()
    b[-3:] = 0.0
    g = b[:-3].reshape(-1, 3)
    g -= g.mean(axis=0)
    xb, rb = base.solve(b)
    xv, rv = sd.solve(b)
    print(f"{form}/{scale}/refine={refine}: P^-1 {e_p:.2e}; GMRES {rv.iterations} it (default {rb.iterations}), "
          f"rel diff {rel(xv, xb):.2e}")
    assert e_p <= 1e-5
    assert rv.converged and rb.converged
    assert rel(xv, xb) <= 1e-4


def test_solve_many_matches_solve(pkg, ctx):
    """GmresSaddle::solve_many (gmres.rs:801): each right-hand side's solution and report are the
    single solve's, bit for bit; an empty batch is a no-op."""
    p, f = pkg.generate.config_icosphere_l3()
    s = pkg.Surface(ctx, f, p)
    sd = pkg.Saddle(s)
    rng = np.random.default_rng(37)
    bs = rng.standard_normal((3, sd.n))
    xs, reps = sd.solve_many(bs)
    for i in range(3):
        x1, r1 = sd.solve(bs[i])
        assert np.array_equal(xs[i], x1)
        assert (reps[i].iterations, reps[i].cycles, reps[i].relative_residual) == (r1.iterations, r1.cycles, r1.relative_residual)
    xs0, reps0 = sd.solve_many(np.zeros((0, sd.n)))
    assert xs0.shape == (0, sd.n) and reps0 == []
    with pytest.raises(pkg.RsError):
        sd.solve_many(bs[:, :-1])


@pytest.mark.parametrize("name", ["icosphere_l3", "torus_16k"])
def test_lprime_device_factor_matches_host(pkg, ctx, name, monkeypatch):
    """L′ factored on the device (per-mesh symbolic analysis, fronts eliminated on the GPU) against
    the host multifrontal factorization of the same order: both solve L′ to the same backward
    error, and their solutions agree to the conditioning of L′ (ε = 1e-10)."""
    p, f = meshes(pkg)[name]()
    s = pkg.Surface(ctx, f, p)
    dev = pkg.Saddle(s)
    monkeypatch.setenv("DDGRS_LPRIME_ON_HOST", "1")
    host = pkg.Saddle(s)
    monkeypatch.delenv("DDGRS_LPRIME_ON_HOST")
    assert abs(dev.info()["eps"] - host.info()["eps"]) <= 1e-13 * host.info()["eps"]
    L, _ = lprime_axis_matrix(f, p, oracle.vertex_areas(f, p))
    rng = np.random.default_rng(23)
    for _ in range(2):
        r = rng.standard_normal(dev.n)
        xd, xh = dev.lprime_solve(r), host.lprime_solve(r)
        res_d, res_h = lprime_residual(L, r, xd), lprime_residual(L, r, xh)
        print(f"{name}: residual device {res_d:.2e} host {res_h:.2e}, rel diff {rel(xd, xh):.2e}")
        assert res_d <= max(10.0 * res_h, 1e-13)
        assert rel(xd, xh) <= 1e-7
        assert rel(dev.precondition(r), host.precondition(r)) <= 1e-7


def test_lprime_solve_paths_agree(pkg, ctx, monkeypatch):
    """The L′ solve's small-front path (matrices staged in LDS, one workgroup per block) and its
    large-front path (rows spread over workgroups; the forward's update vectors from P_J = L_SJ M_J)
    solve L′ to the same backward error, and their solutions agree to the conditioning of L′
    (different summation orders: up to ~1e-8, see above)."""
    p, f = meshes(pkg)["torus_16k"]()
    s = pkg.Surface(ctx, f, p)
    small = pkg.Saddle(s)
    monkeypatch.setenv("DDGRS_LPRIME_LARGE_PATH", "1")
    large = pkg.Saddle(s)
    monkeypatch.delenv("DDGRS_LPRIME_LARGE_PATH")
    rng = np.random.default_rng(24)
    r = rng.standard_normal(small.n)
    xs, xl = small.lprime_solve(r), large.lprime_solve(r)
    # the large path's default forms (u_J from P_J = L_SJ M_J, x_J from P_Jᵀ) against its two phases
    monkeypatch.setenv("DDGRS_LPRIME_SOLVE_TWO_PHASE", "1")
    xl2 = large.lprime_solve(r)
    monkeypatch.delenv("DDGRS_LPRIME_SOLVE_TWO_PHASE")
    L, _ = lprime_axis_matrix(f, p, oracle.vertex_areas(f, p))
    res = {k: lprime_residual(L, r, v) for k, v in (("small", xs), ("large", xl), ("two phases", xl2))}
    print(f"small/large path rel diff {rel(xl, xs):.2e}, one/two phases {rel(xl, xl2):.2e}, residuals {res}")
    assert res["large"] <= max(10.0 * res["small"], 1e-13)
    assert res["two phases"] <= max(10.0 * res["small"], 1e-13)
    assert rel(xl, xs) <= 1e-8
    assert rel(xl, xl2) <= 1e-8


@pytest.mark.parametrize("name", ["icosphere_l3", "torus_16k", "bumpy_l6"])
def test_lprime_subtree_launches_match_per_height(pkg, ctx, name, monkeypatch):
    """The lowest heights' solves in one launch per phase (one workgroup per subtree, a block per
    wave) against one launch per height and phase: each block's sums run over the same terms in the
    same order, so the solutions are bitwise equal.  Three right-hand sides in a row."""
    gen = {"bumpy_l6": pkg.generate.config_bumpy_l6, **meshes(pkg)}[name]
    p, f = gen()
    s = pkg.Surface(ctx, f, p)
    sub = pkg.Saddle(s)
    monkeypatch.setenv("DDGRS_LPRIME_SUBTREES", "0")
    per = pkg.Saddle(s)
    monkeypatch.delenv("DDGRS_LPRIME_SUBTREES")
    rng = np.random.default_rng(26)
    for i in range(3):
        r = rng.standard_normal(sub.n)
        xs, xp = sub.lprime_solve(r), per.lprime_solve(r)
        print(f"{name} rhs {i}: max |subtree - per height| {np.abs(xs - xp).max():.3e}", flush=True)
        assert np.all(np.isfinite(xs))
        assert np.array_equal(xs, xp)


def test_lprime_one_launch_is_refused(pkg, ctx, monkeypatch):
    """The one-launch L′ solve is disabled: asking for it raises InvalidOption before any launch."""
    p, f = pkg.generate.config_icosphere_l3()
    sd = pkg.Saddle(pkg.Surface(ctx, f, p))
    monkeypatch.setenv("DDGRS_LPRIME_SOLVE_ONE_LAUNCH", "1")
    with pytest.raises(pkg.RsError) as e:
        sd.lprime_solve(np.ones(sd.n))
    assert e.value.kind == "InvalidOption"


@pytest.mark.skip(reason="the one-launch L' solve hung on the device twice (the compiler merged thread 0's task "
                  "completion and next ticket across the loop's back edge, so the workgroup re-read the old "
                  "ticket); the restructured, now also iteration-bounded loop has not run on the device, and the "
                  "library refuses the opt-in until it has")
@pytest.mark.parametrize("name,large", [("torus_16k", False), ("torus_16k", True), ("bumpy_l6", False)])
def test_lprime_one_launch_matches_per_height(pkg, ctx, name, large, monkeypatch):
    """The one-launch L′ solve (workgroups take the blocks' tasks in ticket order and wait for the
    tasks they depend on) against one launch per height and phase: each task does the same
    arithmetic, so the solutions are bitwise equal.  Three right-hand sides in a row exercise the
    counters' reset between solves; ``large`` puts every block on the row-split path (y/u and t/x
    tasks)."""
    gen = {"bumpy_l6": pkg.generate.config_bumpy_l6, **meshes(pkg)}[name]
    p, f = gen()
    s = pkg.Surface(ctx, f, p)
    if large:
        monkeypatch.setenv("DDGRS_LPRIME_LARGE_PATH", "1")
    sd = pkg.Saddle(s)
    monkeypatch.delenv("DDGRS_LPRIME_LARGE_PATH", raising=False)
    rng = np.random.default_rng(25)
    for i in range(3):
        r = rng.standard_normal(sd.n)
        monkeypatch.setenv("DDGRS_LPRIME_SOLVE_ONE_LAUNCH", "1")
        x_one = sd.lprime_solve(r)
        monkeypatch.delenv("DDGRS_LPRIME_SOLVE_ONE_LAUNCH")
        x_per = sd.lprime_solve(r)
        print(f"{name} large={large} rhs {i}: max |one - per| {np.abs(x_one - x_per).max():.3e}")
        assert np.all(np.isfinite(x_one))
        assert np.array_equal(x_one, x_per)


def test_lean_saddle_octa(pkg, ctx):
    c, o = fx.read_case("rs-octa-saddle-none"), fx.read_out("rs-octa-saddle-none")
    p, f = np.array(c["positions"]), np.array(c["faces"])
    nv = len(p)
    g = np.array([o["rows"]["g"][i] for i in range(nv)]).reshape(-1)
    xr = np.array([o["rows"]["x"][i] for i in range(nv)]).reshape(-1)
    s = pkg.Surface(ctx, f, p)
    sd = pkg.Saddle(s, pkg.Exponents(*c["exponents"]), pkg.SaddleOptions(chi=0.0))
    b = np.zeros(3 * nv + 3)
    b[: 3 * nv] = g
    x, rep = sd.solve(b)
    err = np.abs(x[: 3 * nv] - xr).max() / np.abs(xr).max()
    print(f"octa saddle: {rep.iterations} iterations, x rel {err:.2e}, <g,x> {g @ x[:3 * nv]:.12g} / {o['gx']:.12g}")
    assert rep.converged
    assert err <= 1e-5
    assert abs(g @ x[: 3 * nv] - o["gx"]) <= 1e-5 * abs(o["gx"])


@pytest.mark.parametrize("name", ["icosphere_l3", "torus_16k"])
def test_saddle_solve(pkg, ctx, name):
    p, f = meshes(pkg)[name]()
    s = pkg.Surface(ctx, f, p)
    sd = pkg.Saddle(s)
    od = oracle.Saddle(f, p)
    e, d = pkg.exact_energy_and_differential(s) if name == "icosphere_l3" else (None, None)
    if d is None:
        d = pkg.BhTree(s, pkg.BhOptions(theta=0.25)).energy_and_differential(pkg.Exponents())[1]
    b = np.zeros(od.n)
    b[: 3 * len(p)] = np.asarray(d).reshape(-1)
    x, rep = sd.solve(b)
    x0, rep0 = od.solve(b)
    err = rel(x, x0)
    print(f"{name}: iterations {rep.iterations} (oracle {rep0['iterations']:.0f}), residual {rep.relative_residual:.2e}, "
          f"solution rel {err:.2e}")
    assert rep.converged and rep.relative_residual <= 1e-6
    assert abs(rep.iterations - rep0["iterations"]) <= 1
    assert err <= 1e-4
    # the device solution solves the oracle's FP64 system to the metric's FP32 accuracy
    assert rel(od.apply(x), b) <= 1e-4
    # the cycle residuals and Arnoldi estimates (gmres.rs:214-217) follow the oracle's while the
    # residual is above the FP32 products' rounding
    assert len(rep.history) == rep.iterations and len(rep.cycle_residuals) >= rep.cycles + 1
    assert abs(rep.cycle_residuals[0] - 1.0) <= 1e-12 and rep.cycle_residuals[-1] == rep.relative_residual
    h0 = rep0["history"]
    for k in range(min(len(h0), len(rep.history))):
        if h0[k] < 1e-4:
            break
        assert abs(rep.history[k] - h0[k]) <= 1e-2 * h0[k], (k, rep.history[k], h0[k])


@pytest.mark.parametrize("name,theta,steps", [("icosphere_l3", 0.5, 3), ("torus_16k", 0.25, 1)])
def test_step_parity(pkg, ctx, name, theta, steps):
    p, f = meshes(pkg)[name]()
    s = pkg.Surface(ctx, f, p)
    ses = pkg.Session(s, pkg.StepOptions(theta=theta))
    ors = oracle.Session(f, p, theta=theta)
    st = ses.state()
    for k in range(steps):
        r = ses.step()
        r0 = ors.step(threads=16)
        x0 = ors.positions()
        x = s.positions()
        print(f"{name} step {k}: E {r.energy_before:.12g} / {r0['energy0']:.12g} -> {r.energy_trial:.12g} / "
              f"{r0['energy1']:.12g}; gmres {r.gmres.iterations} / {r0['gmres_iterations']:.0f}; "
              f"trials {r.trials} / {r0['trials']:.0f}; tau {r.tau:.6g} / {r0['tau']:.6g}")
        assert abs(r.energy_before - r0["energy0"]) <= 1e-5 * abs(r0["energy0"])
        assert r.accepted == bool(r0["accepted"])
        assert r.trials == int(r0["trials"])
        assert abs(r.tau - r0["tau"]) <= 1e-4 * r0["tau"]
        assert abs(r.energy_trial - r0["energy1"]) <= 1e-5 * abs(r0["energy1"])
        assert abs(r.gmres.iterations - r0["gmres_iterations"]) <= 1
        assert r.gmres.converged and r.gmres.relative_residual <= 1e-6
        assert len(r.gmres.history) == r.gmres.iterations and r.gmres.cycle_residuals[-1] == r.gmres.relative_residual
        assert r.energy_trial < r.energy_before
        moved = np.abs(x0 - p).max()
        assert np.abs(x - x0).max() <= 1e-4 * moved
        p = x0
    # the barycenter target (area-weighted, fixed at creation) is kept
    a = oracle.vertex_areas(f, x)
    bc = (a[:, None] * x).sum(0) / a.sum()
    assert np.abs(bc - st["barycenter_target"]).max() <= 1e-12 * max(1.0, np.abs(x).max())


def test_step_l6(pkg, ctx):
    """The bench workload (81,920 faces): one step against the oracle's."""
    p, f = pkg.generate.config_bumpy_l6()
    s = pkg.Surface(ctx, f, p)
    ses = pkg.Session(s)
    ors = oracle.Session(f, p)
    r = ses.step()
    r0 = ors.step(threads=16)
    x, x0 = s.positions(), ors.positions()
    print(f"L6: E {r.energy_before:.12g} / {r0['energy0']:.12g}; gmres {r.gmres.iterations} / "
          f"{r0['gmres_iterations']:.0f}; tau {r.tau:.6g} / {r0['tau']:.6g}; seconds {r.seconds}")
    assert abs(r.energy_before - r0["energy0"]) <= 1e-5 * abs(r0["energy0"])
    assert abs(r.energy_trial - r0["energy1"]) <= 1e-5 * abs(r0["energy1"])
    assert r.trials == int(r0["trials"]) and abs(r.tau - r0["tau"]) <= 1e-4 * r0["tau"]
    assert abs(r.gmres.iterations - r0["gmres_iterations"]) <= 1
    assert np.abs(x - x0).max() <= 1e-4 * np.abs(x0 - p).max()


def test_step_errors(pkg, ctx):
    p, f = pkg.generate.config_icosphere_l3()
    s = pkg.Surface(ctx, f, p)
    with pytest.raises(pkg.RsError) as e:
        pkg.Session(s, pkg.StepOptions(gmres=pkg.GmresOptions(tolerance=0.0)))
    assert e.value.kind == "InvalidOption"
    with pytest.raises(pkg.RsError) as e:
        pkg.Saddle(s, options=pkg.SaddleOptions(leaf_size=0))
    assert e.value.kind == "InvalidOption"
    with pytest.raises(pkg.RsError) as e:
        pkg.Session(s, pkg.StepOptions(exponents=pkg.Exponents(alpha=1.0, beta=2.0)))
    assert e.value.kind == "InvalidExponents"


@pytest.mark.parametrize("switch", ["DDGRS_BLOCK_LEVELS_ON_HOST", "DDGRS_BVH_LEVELS_ON_HOST"])
def test_device_levels_equal_host_levels(pkg, ctx, monkeypatch, switch):
    """Steps whose block cluster trees (or BVHs) are built level by level on the device, sized from
    the previous step, give bitwise the positions of steps that read every level back to the host."""
    p, f = pkg.generate.config_torus_16k()
    out = {}
    for mode in ("device", "host"):
        if mode == "host":
            monkeypatch.setenv(switch, "1")
        s = pkg.Surface(ctx, f, p)
        ses = pkg.Session(s)
        for _ in range(3):
            ses.step()
        out[mode] = s.positions()
    monkeypatch.delenv(switch)
    assert np.array_equal(out["device"], out["host"])


@pytest.mark.parametrize("name,cut_size", [("torus_16k", None), ("icosphere_l3", None), ("bumpy_l6", None),
                                           ("torus_16k", "1e9")])
def test_bvh_subtrees_equal_level_loop(pkg, ctx, monkeypatch, name, cut_size):
    """From the second build on, the BVH's nodes below a cut depth are built one subtree per
    workgroup (the whole tree at L3); the steps are bitwise those of BVHs built level by level.  A
    cut placed above nodes too large for one workgroup (cut size 1e9: the root) falls back to the
    level loop."""
    p, f = meshes(pkg)[name]() if name != "bumpy_l6" else pkg.generate.config_bumpy_l6()
    out = {}
    for mode in ("subtrees", "levels"):
        if mode == "levels":
            monkeypatch.setenv("DDGRS_BVH_SUBTREES", "0")
        elif cut_size:
            monkeypatch.setenv("DDGRS_BVH_CUT_SIZE", cut_size)
        s = pkg.Surface(ctx, f, p)
        ses = pkg.Session(s)
        for _ in range(3):
            ses.step()
        out[mode] = s.positions()
        ses.close()
        s.close()
    assert np.array_equal(out["subtrees"], out["levels"])


def test_long_trajectory_converges(pkg, ctx):
    """Fourteen steps at L6: once the gradient is small and smooth, the FP32 products' rounding
    keeps the explicit residual just above 1e-6 (the Arnoldi estimate reaches it); the solve then
    switches to the extended near field and every step converges to the true relative residual."""
    p, f = pkg.generate.config_bumpy_l6()
    s = pkg.Surface(ctx, f, p)
    ses = pkg.Session(s)
    extended = 0
    for k in range(14):
        r = ses.step()
        extended += r.gmres.extended
        print(f"L6 step {k}: E {r.energy_before:.6g}, gmres {r.gmres.iterations} extended {r.gmres.extended} "
              f"residual {r.gmres.relative_residual:.2e}")
        assert r.gmres.converged and r.gmres.relative_residual <= 1e-6
        assert r.energy_trial < r.energy_before
    assert extended >= 1


def test_step_lean_stuck(pkg, ctx):
    c, o = fx.read_case("rs-icosa-step-stuck"), fx.read_out("rs-icosa-step-stuck")
    p, f = c["positions"], c["faces"]
    s = pkg.Surface(ctx, f, p)
    opts = pkg.StepOptions(exponents=pkg.Exponents(*c["exponents"]), energy_exact=True,
                           saddle=pkg.SaddleOptions(chi=0.0), armijo_c=c["armijo"], max_halvings=c["maxhalv"],
                           start=pkg.StartRule.absolute(c["start"][1]))
    ses = pkg.Session(s, opts)
    r = ses.step()
    relv = lambda a, b: abs(a - b) / abs(b)
    print(f"stuck: F0 {r.energy_before:.12g} / {o['energy0']:.12g}, slope {r.sobolev_norm_sq:.10g} / {o['slope']:.10g}, "
          f"|x| {r.direction_norm:.10g} / {o['xnorm']:.10g}, trials {r.trial_list}")
    assert relv(r.energy_before, o["energy0"]) <= 1e-5
    assert relv(r.sobolev_norm_sq, o["slope"]) <= 1e-4
    assert relv(r.direction_norm, o["xnorm"]) <= 1e-4
    assert relv(ses.state()["mean_edge_length0"], o["lbar"]) <= 1e-12
    assert r.tau0 == o["tau0"]
    want = {"accepted": 0, "armijo-failed": 1, "non-finite": 2}
    assert len(r.trial_list) == len(o["trials"]) == r.trials
    for (tau, fe, outcome), (_, wtau, wout, wf, _) in zip(r.trial_list, o["trials"]):
        assert outcome == want[wout] and tau == wtau
        assert relv(fe, wf) <= 1e-4
    assert not r.accepted and r.tau == 0.0 and r.halvings == o["result"][2]
    assert relv(r.energy_trial, o["energy1"]) <= 1e-5
    assert relv(r.energy_after, o["energy-final"]) <= 1e-5
    assert np.abs(s.positions() - p).max() <= 1e-12


@pytest.mark.parametrize("name", ["rs-icosa-step", "rs-mobius-step", "rs-torus-step-ref"])
def test_step_lean_constrained(pkg, ctx, name):
    """The reference's constrained step fixtures (area and volume rows through the Schur complement
    and the projection's Newton rounds, RS/src/solve/schur.rs:101,495) through the device session
    with exact energy and the exact metric, at the FP32 metric's accuracy: the direction, the
    trials, the constraint values after the projection and the positions."""
    c, o = fx.read_case(name), fx.read_out(name)
    p, f = c["positions"], c["faces"]
    rows = []
    for cc in c["constraints"]:
        if cc[0] in ("area", "volume"):
            rows.append((cc[0], fx.hexf(cc[2]) if len(cc) > 2 else None))
    st = c["start"]
    start = (pkg.StartRule.max_displacement(st[1]) if st[0] == "maxdisp" else
             pkg.StartRule.absolute(st[1]) if st[0] == "abs" else pkg.StartRule.reference())
    s = pkg.Surface(ctx, f, p)
    opts = pkg.StepOptions(exponents=pkg.Exponents(*c["exponents"]), energy_exact=True,
                           saddle=pkg.SaddleOptions(chi=0.0), armijo_c=c["armijo"], max_halvings=c["maxhalv"],
                           start=start, constraints=tuple(rows), newton=c["newton"])
    ses = pkg.Session(s, opts)
    r = ses.step()
    phi_before, phi = ses.constraints()
    relv = lambda a, b: abs(a - b) / abs(b)
    print(f"{name}: F0 {r.energy_before:.12g} / {o['energy0']:.12g}, slope {r.sobolev_norm_sq:.10g} / {o['slope']:.10g}, "
          f"|x| {r.direction_norm:.10g} / {o['xnorm']:.10g}, newton {r.newton_rounds} / {o['newton']}, "
          f"pivot {r.min_pivot:.4g}, phi {phi}")
    assert relv(r.energy_before, o["energy0"]) <= 1e-5
    assert relv(r.sobolev_norm_sq, o["slope"]) <= 1e-4
    assert relv(r.direction_norm, o["xnorm"]) <= 1e-4
    assert relv(r.tau0, o["tau0"]) <= 1e-4
    want = {"accepted": 0, "armijo-failed": 1, "non-finite": 2}
    assert len(r.trial_list) == len(o["trials"])
    for (tau, fe, outcome), (_, wtau, wout, wf, _) in zip(r.trial_list, o["trials"]):
        assert outcome == want[wout] and relv(tau, wtau) <= 1e-4
        assert relv(fe, wf) <= 1e-4
    assert r.accepted and r.halvings == o["result"][2]
    assert relv(r.energy_trial, o["energy1"]) <= 1e-4
    assert len(phi) == len(o["phi"]) == 3 + len(rows)
    # the rows' targets: given, or the values at creation (the oracle's rows at the input)
    v0, _ = oracle.constraint_rows(f, p, [k for k, _ in rows])
    targets = [t if t is not None else v for (_, t), v in zip(rows, v0)]
    # Φ before the projection follows the step's motion; after it the barycenter rows vanish and
    # the nonlinear rows reach their tolerance 1e-8 |target| (rounds until tolerance), or match the
    # fixture's after its fixed rounds
    scale = max(1.0, max(abs(x) for x in o["phi-before"]))
    assert np.max(np.abs(np.array(phi_before) - np.array(o["phi-before"]))) <= 1e-4 * scale
    assert np.max(np.abs(phi[:3])) <= 1e-12 * max(1.0, np.abs(p).max())
    for i, (kind, _) in enumerate(rows):
        w = o["phi"][3 + i]
        if c["newton"][0] == "tol":
            assert 1 <= r.newton_rounds <= c["newton"][1]
            assert abs(phi[3 + i]) <= 1e-8 * abs(targets[i]), (kind, phi[3 + i], w)
        else:
            assert r.newton_rounds == o["newton"]
            assert abs(phi[3 + i] - w) <= 1e-4 * abs(w), (kind, phi[3 + i], w)
    assert relv(r.energy_after, o["energy-final"]) <= 1e-4
    want_x = np.array([o["rows"]["x"][i] for i in range(len(o["rows"]["x"]))])
    moved = np.abs(want_x - p).max()
    assert np.abs(s.positions() - want_x).max() <= 1e-4 * moved
    ses.close()
    s.close()


def test_step_c1_exact_energy(pkg, ctx):
    p, f = pkg.generate.config_icosphere_l3()
    s = pkg.Surface(ctx, f, p)
    ses = pkg.Session(s, pkg.StepOptions(energy_exact=True))
    ors = oracle.Session(f, p, energy_exact=True)
    r = ses.step()
    r0 = ors.step(threads=16)
    x, x0 = s.positions(), ors.positions()
    print(f"C1: E {r.energy_before:.12g} / {r0['energy0']:.12g} -> {r.energy_after:.12g} / {r0['energy_after']:.12g}; "
          f"gmres {r.gmres.iterations} / {r0['gmres_iterations']:.0f}; trials {r.trial_list} / {r0['trial_list']}")
    assert abs(r.energy_before - r0["energy0"]) <= 1e-5 * abs(r0["energy0"])
    assert abs(r.sobolev_norm_sq - r0["slope"]) <= 1e-4 * abs(r0["slope"])
    assert abs(r.direction_norm - r0["direction_norm"]) <= 1e-4 * r0["direction_norm"]
    assert len(r.trial_list) == len(r0["trial_list"])
    for (tau, fe, out), (tau0, fe0, out0) in zip(r.trial_list, r0["trial_list"]):
        assert out == out0 and abs(tau - tau0) <= 1e-4 * tau0
        assert abs(fe - fe0) <= 1e-5 * abs(fe0)
    assert r.accepted and r0["accepted"] == 1.0
    assert abs(r.energy_trial - r0["energy1"]) <= 1e-5 * abs(r0["energy1"])
    assert abs(r.energy_after - r0["energy_after"]) <= 1e-5 * abs(r0["energy_after"])
    assert abs(r.gmres.iterations - r0["gmres_iterations"]) <= 1
    assert r.gmres.converged and r.gmres.relative_residual <= 1e-6
    assert np.abs(x - x0).max() <= 1e-4 * np.abs(x0 - p).max()


def test_recovery_chain_retry(pkg, ctx):
    """A diverged solve (forced from the host: one GMRES iteration, every unconverged solve counted
    as diverged) is retried with right GMRES at restart 60 and 800 iterations (direction.rs:277
    escalate, stage 1); the step then equals the ordinary one."""
    p, f = pkg.generate.config_icosphere_l3()
    ref = pkg.Session(pkg.Surface(ctx, f, p)).step()
    s = pkg.Surface(ctx, f, p)
    r = pkg.Session(s, pkg.StepOptions(gmres=pkg.GmresOptions(max_iterations=1), diverged_residual=1e-30)).step()
    print(f"retry: stage {r.fallback_stage}, {r.fallbacks}, gmres {r.gmres.iterations} residual "
          f"{r.gmres.relative_residual:.2e}")
    assert r.fallback_stage == 1 and len(r.fallbacks) == 1 and "retried" in r.fallbacks[0]
    assert r.gmres.converged and r.gmres.relative_residual <= 1e-6
    assert r.trials == ref.trials and abs(r.tau - ref.tau) <= 1e-4 * ref.tau
    assert abs(r.energy_trial - ref.energy_trial) <= 1e-6 * abs(ref.energy_trial)


def test_recovery_chain_dense(pkg, ctx):
    """When the retry diverges too, a mesh of at most 6,000 vertices is solved by a dense LU of
    the exact metric's saddle system on the device (direction.rs:277 escalate, stage 2): the step
    equals the oracle's step with the exact (χ = 0) metric solved to 1e-12."""
    p, f = pkg.generate.config_icosphere_l3()
    s = pkg.Surface(ctx, f, p)
    opts = pkg.StepOptions(gmres=pkg.GmresOptions(max_iterations=1), diverged_residual=1e-30, retry_max_iterations=1)
    r = pkg.Session(s, opts).step()
    r0 = oracle.Session(f, p, chi=0.0, tol=1e-12, maxit=500).step(threads=16)
    print(f"dense: stage {r.fallback_stage}, {r.fallbacks}; tau {r.tau:.8g} / {r0['tau']:.8g}, "
          f"E {r.energy_trial:.12g} / {r0['energy1']:.12g}")
    assert r.fallback_stage == 2 and len(r.fallbacks) == 1 and "dense LU" in r.fallbacks[0]
    assert r.gmres.converged and r.gmres.relative_residual <= 1e-10
    assert abs(r.sobolev_norm_sq - r0["slope"]) <= 1e-5 * abs(r0["slope"])
    assert r.trials == int(r0["trials"]) and abs(r.tau - r0["tau"]) <= 1e-5 * r0["tau"]
    assert abs(r.energy_trial - r0["energy1"]) <= 1e-5 * abs(r0["energy1"])


def test_recovery_chain_off(pkg, ctx):
    p, f = pkg.generate.config_icosphere_l3()
    s = pkg.Surface(ctx, f, p)
    ses = pkg.Session(s, pkg.StepOptions(gmres=pkg.GmresOptions(max_iterations=1), diverged_residual=1e-30,
                                         robust=False))
    with pytest.raises(pkg.RsError) as e:
        ses.step()
    assert e.value.kind == "SolverDiverged"
