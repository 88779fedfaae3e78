"""FP64 NumPy restatement of the learned right preconditioner — TEST INFRASTRUCTURE.

The reference has no learned preconditioner (its solver uses the plain one only,
``crates/ddg-repulsive-surfaces/src/solve/precond.rs``); this module restates the network that the
library trains on the device, so that its features, forward pass, loss and gradient can be checked:

* features (per vertex, 8): position centred at the mean vertex and divided by the RMS radius;
  the unit sum of the incident faces' unnormalised normals; the vertex area (a third of the
  incident face areas) over the mean vertex area; the umbrella Laplacian (mean neighbour minus
  the vertex) along that normal over the mean squared edge length, times the RMS radius (a
  dimensionless mean curvature);
* network: h0 = relu(F W0 + b0); four layers h <- relu([h, mean_N(h)] W_l + b_l) of width 64;
  vertex head B_v = h_v Wv + bv (9, row-major); edge head per undirected edge {u < v} (ascending):
  t = relu((h_u + h_v) We1 + be1) (64), E_uv = t We2 + be2 (9, row-major);
* (Q x)_v = s (B_v x_v + sum_{u in N(v)} E_uv x_u), loss |s q - e|^2 / |e|^2 for a stored pair
  (x, e) (the library stores x = c P^-1 r and e = y - x); ``forward`` returns the blocks as one
  array: the nv vertex rows, then the edge rows;
* ``bf16=True`` rounds both inputs of the products that the library runs on BF16 matrix cores
  (every product with 64 output columns: the input, message and edge layers, the back-propagated
  activations through every layer and head, the weight and bias gradients of the 64-wide layers)
  to BF16, nearest with ties to even, and accumulates exactly; the 9-output heads and their weight
  gradients stay unrounded (FP32 on the device).

Parity pins (``tests/test_oracle_learned.py``): the analytic gradient against central differences
of the loss (every parameter group), the features against closed forms on a sphere (radial unit
normals, equal areas on a regular mesh, the Laplacian's sign), Q's linearity and symmetry
(Q is symmetric when B_v is, because the edge blocks are shared); the BF16 rounding against hand
values (ties to even, carries into the exponent), and the BF16 products against the unrounded ones
when every input is already a BF16 value.

Only tests may import this module.
"""
from __future__ import annotations

import numpy as np

FEAT, WIDTH, LAYERS, OUT = 8, 64, 4, 9


def bf16(x):
    """x rounded to FP32, then to BF16 (the upper 16 bits; nearest, ties to even), as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _mm(a, b, rounded):
    return bf16(a) @ bf16(b) if rounded else a @ b


def layout():
    """Offsets of the parameter groups in the flat FP32 array (the library's order)."""
    o = 0
    lay = {}

    def take(name, count):
        nonlocal o
        lay[name] = (o, count)
        o += count

    take("w0", FEAT * WIDTH)
    take("b0", WIDTH)
    for l in range(LAYERS):
        take(f"w{l + 1}", 2 * WIDTH * WIDTH)
        take(f"b{l + 1}", WIDTH)
    take("wv", WIDTH * OUT)
    take("bv", OUT)
    take("we1", WIDTH * WIDTH)
    take("be1", WIDTH)
    take("we2", WIDTH * OUT)
    take("be2", OUT)
    lay["total"] = o
    return lay


def unpack(p):
    lay = layout()
    p = np.asarray(p, dtype=np.float64)
    out = {}
    shapes = {"w0": (FEAT, WIDTH), "wv": (WIDTH, OUT), "we1": (WIDTH, WIDTH), "we2": (WIDTH, OUT)}
    for l in range(LAYERS):
        shapes[f"w{l + 1}"] = (2 * WIDTH, WIDTH)
    for k, v in lay.items():
        if k == "total":
            continue
        o, c = v
        out[k] = p[o:o + c].reshape(shapes.get(k, (c,)))
    return out


def neighbours(faces, nv):
    """CSR of the mesh's edges in both directions, ascending per vertex."""
    adj = [set() for _ in range(nv)]
    for a, b, c in np.asarray(faces):
        for x, y in ((a, b), (b, c), (c, a)):
            adj[x].add(int(y))
            adj[y].add(int(x))
    ptr = np.zeros(nv + 1, dtype=np.int64)
    nb = []
    for v in range(nv):
        s = sorted(adj[v])
        nb.extend(s)
        ptr[v + 1] = len(nb)
    return ptr, np.asarray(nb, dtype=np.int64)


def edges(ptr, nb):
    """The undirected edges {u < v} in ascending order (ne × 2) and, per neighbour entry of the
    CSR, its edge."""
    nv = ptr.shape[0] - 1
    src = np.repeat(np.arange(nv), np.diff(ptr))
    e = np.stack([np.minimum(src, nb), np.maximum(src, nb)], axis=1)
    uniq, emap = np.unique(e, axis=0, return_inverse=True)
    return uniq, emap.reshape(-1)


def features(pos, faces):
    pos = np.asarray(pos, dtype=np.float64)
    faces = np.asarray(faces)
    nv = pos.shape[0]
    ptr, nb = neighbours(faces, nv)
    c = pos.mean(axis=0)
    R = np.sqrt(((pos - c) ** 2).sum() / nv) + 1e-300
    nf = np.cross(pos[faces[:, 1]] - pos[faces[:, 0]], pos[faces[:, 2]] - pos[faces[:, 0]])
    a2 = np.linalg.norm(nf, axis=1)
    vn = np.zeros((nv, 3))
    area = np.zeros(nv)
    for k in range(3):
        np.add.at(vn, faces[:, k], nf)
        np.add.at(area, faces[:, k], a2 / 6.0)
    deg = np.diff(ptr)
    src = np.repeat(np.arange(nv), deg)
    el = (((pos[nb] - pos[src]) ** 2).sum(axis=1)).mean()
    nn = np.linalg.norm(vn, axis=1)
    nn[nn == 0] = 1.0
    un = vn / nn[:, None]
    nsum = np.zeros((nv, 3))
    np.add.at(nsum, src, pos[nb])
    lap = np.where(deg[:, None] > 0, nsum / np.maximum(deg, 1)[:, None] - pos, 0.0)
    h = (lap * un).sum(axis=1)
    f = np.zeros((nv, FEAT))
    f[:, 0:3] = (pos - c) / R
    f[:, 3:6] = un
    f[:, 6] = area / area.mean()
    f[:, 7] = h / el * R
    return f


def _mean(h, ptr, nb):
    nv = ptr.shape[0] - 1
    deg = np.diff(ptr)
    src = np.repeat(np.arange(nv), deg)
    m = np.zeros_like(h)
    np.add.at(m, src, h[nb])
    return m / np.maximum(deg, 1)[:, None]


def _mean_t(dm, ptr, nb):
    nv = ptr.shape[0] - 1
    deg = np.diff(ptr)
    src = np.repeat(np.arange(nv), deg)
    d = np.zeros_like(dm)
    np.add.at(d, nb, dm[src] / np.maximum(deg, 1)[src][:, None])
    return d


def forward(p, f, ptr, nb, keep=False, bf16=False):
    P = unpack(p)
    pre0 = _mm(f, P["w0"], bf16) + P["b0"]
    h = np.maximum(pre0, 0.0)
    cache = [(f, None, pre0)]
    for l in range(LAYERS):
        m = _mean(h, ptr, nb)
        x = np.concatenate([h, m], axis=1)
        pre = _mm(x, P[f"w{l + 1}"], bf16) + P[f"b{l + 1}"]
        cache.append((x, m, pre))
        h = np.maximum(pre, 0.0)
    B = h @ P["wv"] + P["bv"]
    E, _ = edges(ptr, nb)
    x = h[E[:, 0]] + h[E[:, 1]]
    epre = _mm(x, P["we1"], bf16) + P["be1"]
    t = np.maximum(epre, 0.0)
    Eo = t @ P["we2"] + P["be2"]
    o = np.concatenate([B, Eo])
    return (o, h, cache, (x, epre, t)) if keep else o


def q_apply(o, r3, ptr, nb):
    """q_v = B_v r_v + sum_u E_uv r_u (without the scale); o = the vertex rows, then the edge rows."""
    nv = ptr.shape[0] - 1
    B = o[:nv].reshape(nv, 3, 3)
    Eo = o[nv:].reshape(-1, 3, 3)
    _, emap = edges(ptr, nb)
    src = np.repeat(np.arange(nv), np.diff(ptr))
    q = np.einsum("vab,vb->va", B, r3)
    np.add.at(q, src, np.einsum("kab,kb->ka", Eo[emap], r3[nb]))
    return q


def loss_grad(p, f, ptr, nb, r, e, s, bf16=False):
    """Loss |s q - e|^2 / |e|^2 and its gradient (flat, the parameter layout) by backpropagation."""
    nv = f.shape[0]
    r3 = np.asarray(r, dtype=np.float64)[:3 * nv].reshape(nv, 3)
    e3 = np.asarray(e, dtype=np.float64)[:3 * nv].reshape(nv, 3)
    e2 = (e3 ** 2).sum()
    o, h4, cache, (ex, epre, et) = forward(p, f, ptr, nb, keep=True, bf16=bf16)
    rb = globals()["bf16"] if bf16 else (lambda a: a)
    q = s * q_apply(o, r3, ptr, nb)
    d = q - e3
    L = (d ** 2).sum() / e2
    qd = 2.0 * d / e2
    E, _ = edges(ptr, nb)
    u, v = E[:, 0], E[:, 1]
    dB = s * np.einsum("va,vb->vab", qd, r3).reshape(nv, 9)
    dE = s * (np.einsum("ka,kb->kab", qd[u], r3[v]) + np.einsum("ka,kb->kab", qd[v], r3[u])).reshape(-1, 9)
    P = unpack(p)
    grads = {"wv": h4.T @ dB, "bv": dB.sum(axis=0), "we2": et.T @ dE, "be2": dE.sum(axis=0)}
    dh = _mm(dB, P["wv"].T, bf16)
    dpe = _mm(dE, P["we2"].T, bf16) * (epre > 0)
    grads["we1"] = _mm(ex.T, dpe, bf16)
    grads["be1"] = rb(dpe).sum(axis=0)
    dx = _mm(dpe, P["we1"].T, bf16)
    np.add.at(dh, u, dx)
    np.add.at(dh, v, dx)
    for l in range(LAYERS, 0, -1):
        x, _, pre = cache[l]
        dpre = dh * (pre > 0)
        grads[f"w{l}"] = _mm(x.T, dpre, bf16)
        grads[f"b{l}"] = rb(dpre).sum(axis=0)
        W = P[f"w{l}"]
        dh = _mm(dpre, W[:WIDTH].T, bf16) + _mean_t(_mm(dpre, W[WIDTH:].T, bf16), ptr, nb)
    _, _, pre0 = cache[0]
    dpre = dh * (pre0 > 0)
    grads["w0"] = _mm(f.T, dpre, bf16)
    grads["b0"] = rb(dpre).sum(axis=0)
    lay = layout()
    g = np.zeros(lay["total"])
    for k, v in lay.items():
        if k != "total":
            g[v[0]:v[0] + v[1]] = grads[k].reshape(-1)
    return L, g
