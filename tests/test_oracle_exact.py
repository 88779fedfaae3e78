"""Pins the FP64 oracle's exact energy and differential (CPU only).

Pins, following the reference's tests (crates/ddg-repulsive-surfaces/tests/energy.rs,
oracle.rs, reference_fixtures.rs):
* the Lean golden fixtures ``rs-*-energy*`` (tolerance 1e-12 normwise, tests/oracle.rs:39-40) and
  the error cases (same barycenter, degenerate face, bad exponents);
* the authors'-program fixtures (energy 1e-12, differential 1e-11 max-relative,
  reference_fixtures.rs:118-119);
* closed forms (energy.rs:362 ``two_triangles_by_hand``), a plain double loop with powf
  (energy.rs:397), rigid motion / scaling / gradient identities (energy.rs:250) and central
  differences (energy.rs:196);
* a planted mistake in a copy of the formula must fail these checks (the checks are not vacuous).
"""
import numpy as np
import pytest

import oracle
from conftest import rel_max
from golden import fixtures as fx


@pytest.mark.parametrize("name", fx.cases("energy"))
def test_lean_energy_fixtures(name):
    c, o = fx.read_case(name), fx.read_out(name)
    a, b = c["exponents"]
    e, d, _ = oracle.exact(c["faces"], c["positions"], a, b, threads=2)
    dref = np.array([o["rows"]["d"][i] for i in range(len(o["rows"]["d"]))])
    scale = max(abs(o["energy"]), 1e-300)
    assert abs(e - o["energy"]) <= 1e-12 * scale or (o["energy"] == 0.0 and e == 0.0)
    gs = np.linalg.norm(dref)
    assert np.linalg.norm(d - dref) <= 1e-12 * max(gs, 1e-300) or gs == 0.0


@pytest.mark.parametrize(
    "name,kind",
    [("rs-error-same-barycenter", "NonFiniteEnergy"), ("rs-error-degenerate", "DegenerateFace"),
     ("rs-error-exponents", "InvalidExponents")],
)
def test_lean_error_fixtures(name, kind):
    c, o = fx.read_case(name), fx.read_out(name)
    assert o["error"][0] == kind
    a, b = c["exponents"]
    with pytest.raises(oracle.OracleError) as ei:
        oracle.exact(c["faces"], c["positions"], a, b)
    assert ei.value.kind == kind
    if kind == "NonFiniteEnergy":
        assert tuple(int(x) for x in ei.value.culprit) == (0, 1)


@pytest.mark.parametrize("name", fx.AUTHORS_NAMES)
def test_authors_program_all_pairs(name):
    f = fx.authors(name)
    kv = f["kv"]
    e, d, _ = oracle.exact(f["faces"], f["positions"], kv["alpha"], kv["beta"])
    assert abs(e - kv["tpe_all_pairs"]) <= 1e-12 * kv["tpe_all_pairs"]
    assert rel_max(d, f["d_all"]) < 1e-11


TWO = np.array([[0, 1, 2], [3, 4, 5]], dtype=np.uint32)


def test_two_triangles_by_hand():
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 0, 1], [2, 0, 1], [1, 1, 1]], dtype=float)
    e, _, _ = oracle.exact(TWO, p)
    assert e == 2.0 * 0.25 / 64.0
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [2, 0, 0], [3, 0, 0], [2, 0, 1]], dtype=float)
    e, _, _ = oracle.exact(TWO, p)
    want = 2.0 * 0.25 * (1.0 / 3.0) ** 6 * (9.0 / 38.0) ** 6
    assert abs(e - want) <= 1e-14 * want
    e, _, _ = oracle.exact(TWO, p, 4.5, 9.5)
    want = 2.0 * 0.25 * (1.0 / 3.0) ** 4.5 * (38.0 / 9.0) ** -4.75
    assert abs(e - want) <= 1e-14 * want


def naive_energy(faces, p, alpha, beta):
    """The plain formula: cross products and powf per ordered pair (energy.rs:397)."""
    a0, a1, a2 = p[faces[:, 0]], p[faces[:, 1]], p[faces[:, 2]]
    n = np.cross(a1 - a0, a2 - a0)
    ln = np.linalg.norm(n, axis=1)
    area, nrm, x = 0.5 * ln, n / ln[:, None], (a0 + a1 + a2) / 3.0
    v = x[:, None, :] - x[None, :, :]
    nu = np.abs(np.einsum("ik,ijk->ij", nrm, v))
    r = np.linalg.norm(v, axis=2)
    np.fill_diagonal(r, 1.0)
    k = nu**alpha / r**beta
    np.fill_diagonal(k, 0.0)
    return float((k * area[:, None] * area[None, :]).sum())


def small_meshes(pkg_gen):
    p, f = pkg_gen.icosphere(1)
    yield "icosphere1", pkg_gen.normal_jitter(p, f, 0.03, 11), f
    p, f = pkg_gen.torus(8, 5, 1.0, 0.4)
    yield "torus8x5", pkg_gen.iso_jitter(p, 0.02, 12), f


EXPS = [(6.0, 12.0), (4.0, 8.0), (3.0, 7.5), (4.5, 9.5)]


@pytest.fixture(scope="module")
def gen():
    from conftest import load_package

    return load_package().generate


def test_plain_double_loop(gen):
    for name, p, f in small_meshes(gen):
        for a, b in EXPS:
            e, _, _ = oracle.exact(f, p, a, b)
            naive = naive_energy(f, p, a, b)
            assert abs(e - naive) <= 1e-12 * naive, (name, a, b)


def rodrigues(axis, th):
    axis = np.asarray(axis) / np.linalg.norm(axis)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(th) * k + (1 - np.cos(th)) * (k @ k)


def test_invariances_scaling_and_gradient_identities(gen):
    r = rodrigues([0.3, -0.5, 0.81], 0.83)
    t = np.array([0.4, -1.3, 2.2])
    for name, p, f in small_meshes(gen):
        for a, b in EXPS:
            e0, g0, _ = oracle.exact(f, p, a, b)
            e1, g1, _ = oracle.exact(f, p @ r.T + t, a, b)
            assert abs(e1 - e0) <= 1e-12 * e0
            gs = np.abs(g0).max()
            assert np.abs(g1 - g0 @ r.T).max() <= 1e-10 * gs
            cs = 1.7
            e2, _, _ = oracle.exact(f, cs * p, a, b)
            want = cs ** (4.0 + a - b) * e0
            assert abs(e2 - want) <= 1e-12 * want
            # Σ g = 0, Σ f × g = 0, Σ f·g = (4 + α − β) E   (SPEC §3.3)
            scale = np.abs(g0).max() * np.abs(p).max() * len(p)
            assert np.abs(g0.sum(axis=0)).max() <= 1e-12 * scale
            assert np.abs(np.cross(p, g0).sum(axis=0)).max() <= 1e-12 * scale
            assert abs((p * g0).sum() - (4.0 + a - b) * e0) <= 1e-10 * max(e0, scale)


def test_central_differences(gen):
    name, p, f = next(iter(small_meshes(gen)))
    edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    h = 1e-6 * np.linalg.norm(p[edges[:, 0]] - p[edges[:, 1]], axis=1).mean()
    for a, b in [(6.0, 12.0), (4.5, 9.5)]:
        _, g, _ = oracle.exact(f, p, a, b)
        fd = np.zeros_like(g)
        for v in range(len(p)):
            for k in range(3):
                q = p.copy()
                q[v, k] += h
                ep = oracle.exact(f, q, a, b)[0]
                q[v, k] -= 2 * h
                em = oracle.exact(f, q, a, b)[0]
                fd[v, k] = (ep - em) / (2 * h)
        assert np.linalg.norm(fd - g) / np.linalg.norm(g) < 1e-6


def test_thread_count_independent(gen):
    p, f = gen.config_icosphere_l3()
    e1, g1, _ = oracle.exact(f, p, threads=1)
    e8, g8, _ = oracle.exact(f, p, threads=8)
    assert e1 == e8 and np.array_equal(g1, g8)


def numpy_pull_back(faces, p, fp, drop_normal=False):
    """geometry.rs:87 face_pull_back in numpy; `drop_normal` plants a dropped term."""
    out = np.zeros_like(p)
    for s_, t in enumerate(faces):
        q = p[t]
        n = np.cross(q[1] - q[0], q[2] - q[0])
        ln = np.linalg.norm(n)
        nrm = n / ln
        e = [q[2] - q[1], q[0] - q[2], q[1] - q[0]]
        for c in range(3):
            ga = np.cross(nrm, e[c]) * 0.5 * fp[s_, 0]
            gn = nrm * (np.cross(e[c], nrm) @ fp[s_, 4:7]) / ln
            out[t[c]] += ga + fp[s_, 1:4] / 3.0 + (0.0 if drop_normal else gn)
    return out


def test_checks_catch_a_planted_mistake(gen):
    """The central-difference check is not vacuous: the pull-back restated in numpy reproduces
    the oracle's differential, and the same pull-back with its normal-variation term dropped (a
    plausible slip) fails the check."""
    name, p, f = next(iter(small_meshes(gen)))
    _, g, fp = oracle.exact(f, p)
    good = numpy_pull_back(f, p, fp)
    assert np.abs(good - g).max() <= 1e-12 * np.abs(g).max()
    bad = numpy_pull_back(f, p, fp, drop_normal=True)
    assert np.linalg.norm(bad - g) / np.linalg.norm(g) > 1e-3
