"""Pins of the NumPy restatement of the learned preconditioner (oracle/learned.py) on the CPU:
the backpropagated gradient against central differences, the features against closed forms on a
sphere, and Q's linearity and symmetry."""
import numpy as np

from conftest import load_package
from oracle import learned as LN


def _mesh(level=1):
    gen = load_package().generate
    pos, faces = gen.icosphere(level)
    return np.asarray(pos, dtype=np.float64), np.asarray(faces, dtype=np.int64)


def _random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(LN.layout()["total"]) * scale


def test_gradient_matches_central_differences():
    pos, faces = _mesh(1)
    rng = np.random.default_rng(5)
    pos = pos * (1.0 + 0.05 * rng.standard_normal((pos.shape[0], 1)))
    ptr, nb = LN.neighbours(faces, pos.shape[0])
    f = LN.features(pos, faces)
    p = _random_params(1)
    n = 3 * pos.shape[0] + 3
    r, e = rng.standard_normal(n), rng.standard_normal(n)
    s = 0.7
    L, g = LN.loss_grad(p, f, ptr, nb, r, e, s)
    lay = LN.layout()
    h = 1e-6
    # a few parameters of every group (a dropped term or a transposed weight fails one of them)
    for name, v in lay.items():
        if name == "total":
            continue
        o, c = v
        for i in rng.choice(c, size=min(c, 6), replace=False):
            k = o + int(i)
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (LN.loss_grad(pp, f, ptr, nb, r, e, s)[0] - LN.loss_grad(pm, f, ptr, nb, r, e, s)[0]) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-7, (name, i, fd, g[k])


def test_features_on_a_sphere():
    pos, faces = _mesh(2)
    f = LN.features(pos, faces)
    # unit normals radial (outward: the generator's faces are counter-clockwise seen from outside)
    radial = pos / np.linalg.norm(pos, axis=1)[:, None]
    assert np.allclose(np.linalg.norm(f[:, 3:6], axis=1), 1.0)
    assert np.min((f[:, 3:6] * radial).sum(axis=1)) > 0.99
    # positions centred (the mesh is symmetric) and of RMS radius 1
    assert np.allclose(f[:, 0:3], radial, atol=1e-12)
    # areas near the mean; the umbrella Laplacian points inward on a sphere
    assert abs(f[:, 6].mean() - 1.0) < 1e-12 and f[:, 6].min() > 0.5
    assert np.all(f[:, 7] < 0.0)
    # scale invariance of every feature
    assert np.allclose(LN.features(3.0 * pos, faces), f, atol=1e-12)


def test_q_is_linear_and_symmetric_with_symmetric_vertex_blocks():
    """Q is linear; with symmetric vertex and edge blocks it is symmetric (each edge block serves
    both directions of its edge); a non-symmetric edge block breaks it."""
    pos, faces = _mesh(1)
    nv = pos.shape[0]
    ptr, nb = LN.neighbours(faces, nv)
    E, emap = LN.edges(ptr, nb)
    assert E.shape[0] == len(nb) // 2 and np.all(E[:, 0] < E[:, 1])
    rng = np.random.default_rng(3)
    o = rng.standard_normal((nv + E.shape[0], 9))
    M = o.reshape(-1, 3, 3)
    o = ((M + M.transpose(0, 2, 1)) / 2).reshape(-1, 9)
    x, y = rng.standard_normal((nv, 3)), rng.standard_normal((nv, 3))
    qx, qy = LN.q_apply(o, x, ptr, nb), LN.q_apply(o, y, ptr, nb)
    assert np.allclose(LN.q_apply(o, 2 * x - y, ptr, nb), 2 * qx - qy)
    assert abs((qx * y).sum() - (x * qy).sum()) < 1e-10 * np.abs(qx).sum()
    o2 = o.copy()
    o2[nv + 5, 1] += 1.0
    q2x, q2y = LN.q_apply(o2, x, ptr, nb), LN.q_apply(o2, y, ptr, nb)
    assert abs((q2x * y).sum() - (x * q2y).sum()) > 1e-3
    # at Q = 0 (the output layers zero) the loss is 1
    p = _random_params(2)
    lay = LN.layout()
    for k in ("wv", "bv", "we2", "be2"):
        o_, c = lay[k]
        p[o_:o_ + c] = 0.0
    f = LN.features(pos, faces)
    r, e = rng.standard_normal(3 * nv + 3), rng.standard_normal(3 * nv + 3)
    assert abs(LN.loss_grad(p, f, ptr, nb, r, e, 1.0)[0] - 1.0) < 1e-15


def test_bf16_rounding_hand_values():
    """BF16 keeps 7 fraction bits: half an ulp ties to the even neighbour, above it rounds up, and
    a tie below a power of two carries into the exponent."""
    u = 2.0 ** -7
    x = np.array([1.0, 1 + u / 2, 1 + u + u / 2, 1 + u / 2 + 2.0 ** -20, 2 - u / 2, -(1 + u + u / 2), 0.0, 3.0])
    want = np.array([1.0, 1.0, 1 + 2 * u, 1 + u, 2.0, -(1 + 2 * u), 0.0, 3.0])
    assert np.array_equal(LN.bf16(x), want)


def test_bf16_products_on_bf16_inputs_are_exact():
    """Products whose inputs are BF16 values already are the unrounded products; the BF16 loss and
    gradient stay within BF16 rounding of the unrounded ones."""
    rng = np.random.default_rng(3)
    a, b = LN.bf16(rng.standard_normal((40, 16))), LN.bf16(rng.standard_normal((16, 64)))
    assert np.array_equal(LN._mm(a, b, True), a @ b)
    pos, faces = _mesh(1)
    ptr, nb = LN.neighbours(faces, pos.shape[0])
    f = LN.features(pos, faces)
    p = _random_params(2)
    n = 3 * pos.shape[0] + 3
    r, e = rng.standard_normal(n), rng.standard_normal(n)
    L0, g0 = LN.loss_grad(p, f, ptr, nb, r, e, 0.7)
    L1, g1 = LN.loss_grad(p, f, ptr, nb, r, e, 0.7, bf16=True)
    assert abs(L1 - L0) <= 3e-2 * abs(L0)
    assert np.linalg.norm(g1 - g0) <= 5e-2 * np.linalg.norm(g0)
    assert not np.array_equal(g1, g0)
