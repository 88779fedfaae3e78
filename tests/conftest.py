"""Shared test setup: the `gpu` marker, the package loader and common meshes."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_package():
    """Imports ``ddg-tools_amd/`` (a directory name Python cannot import by name) as ``ddg_tools_amd``."""
    name = "ddg_tools_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg = os.path.join(ROOT, "ddg-tools_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def ctx(pkg):
    c = pkg.Context(0)
    yield c
    c.close()


def rel_max(a, b):
    """max |a − b| / max |b| over all components (reference_fixtures.rs:87 ``rel_max``)."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def lprime_axis_matrix(faces, positions, mass, eps_rel=1e-10):
    """L′ = [[W + εM, c], [cᵀ, 0]] of one axis, assembled independently in numpy/scipy: W the cotan
    Laplacian ½(cot α + cot β) per edge, M = diag(a_v), c = a/Σa, ε = eps_rel·tr W / tr M
    (RS/src/solve/precond.rs:705-760).  Returns (matrix, ε)."""
    import scipy.sparse as sp

    f = np.asarray(faces, dtype=np.int64)
    p = np.asarray(positions, dtype=np.float64)
    nv = p.shape[0]
    rows, cols, vals = [], [], []
    for c in range(3):
        a, b = f[:, (c + 1) % 3], f[:, (c + 2) % 3]
        u, w = p[a] - p[f[:, c]], p[b] - p[f[:, c]]
        h = 0.5 * (u * w).sum(1) / np.linalg.norm(np.cross(u, w), axis=1)
        rows += [a, b, a, b]
        cols += [b, a, a, b]
        vals += [-h, -h, h, h]
    W = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nv, nv))
    eps = eps_rel * W.diagonal().sum() / mass.sum()
    crow = mass / mass.sum()
    L = sp.bmat([[W + sp.diags(eps * mass), crow[:, None]], [crow[None, :], None]]).tocsr()
    return L, eps


def lprime_residual(L, r, x):
    """max over the three axes of ‖L′ x_k − r_k‖ / ‖r_k‖ for interleaved saddle vectors."""
    nv = L.shape[0] - 1
    worst = 0.0
    for k in range(3):
        rk = np.concatenate([r[k:3 * nv:3], [r[3 * nv + k]]])
        xk = np.concatenate([x[k:3 * nv:3], [x[3 * nv + k]]])
        worst = max(worst, float(np.linalg.norm(L @ xk - rk) / np.linalg.norm(rk)))
    return worst
