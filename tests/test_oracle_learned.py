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
    pos, faces = _mesh(1)
    nv = pos.shape[0]
    ptr, nb = LN.neighbours(faces, nv)
    rng = np.random.default_rng(3)
    o = rng.standard_normal((nv, 18))
    B = o[:, :9].reshape(nv, 3, 3)
    o[:, :9] = (B + B.transpose(0, 2, 1)).reshape(nv, 9) / 2
    o[:, 9:] = ((o[:, 9:].reshape(nv, 3, 3) + o[:, 9:].reshape(nv, 3, 3).transpose(0, 2, 1)) / 2).reshape(nv, 9)
    x, y = rng.standard_normal((nv, 3)), rng.standard_normal((nv, 3))
    qx, qy = LN.q_apply(o, x, ptr, nb), LN.q_apply(o, y, ptr, nb)
    assert np.allclose(LN.q_apply(o, 2 * x - y, ptr, nb), 2 * qx - qy)
    assert abs((qx * y).sum() - (x * qy).sum()) < 1e-10 * np.abs(qx).sum()
    # at Q = 0 (the output layer zero) the loss is 1
    p = _random_params(2)
    lay = LN.layout()
    for k in ("wo", "bo"):
        o_, c = lay[k]
        p[o_:o_ + c] = 0.0
    f = LN.features(pos, faces)
    r, e = rng.standard_normal(3 * nv + 3), rng.standard_normal(3 * nv + 3)
    assert abs(LN.loss_grad(p, f, ptr, nb, r, e, 1.0)[0] - 1.0) < 1e-15
