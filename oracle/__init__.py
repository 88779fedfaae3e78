"""FP64 CPU oracle of the Hs-descent path of ``ddg-repulsive-surfaces`` — TEST INFRASTRUCTURE.

A C++ restatement of the reference's Rust code (``oracle/src/*.cpp``, each function citing the
reference file and line it follows), compiled with ``-ffp-contract=off`` by ``oracle/Makefile``
into ``oracle/_build/librs_oracle.so`` and loaded here through ctypes.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may use
this module, and only as the checker or the timed CPU baseline.  The GPU library never loads it.

Parity pins (see the oracle tests, ``tests/test_oracle_*.py``):
* exact energy / differential: the Lean golden fixtures ``rs-*-energy*`` (1e-12 normwise), the
  authors'-program fixtures (1e-12 / 1e-11), closed forms (two parallel / perpendicular
  triangles), brute force, the scaling and rigid-motion identities and central differences;
* BVH / face tree / Barnes-Hut lists: structural invariants (leaf bound, contiguous ranges, the
  quarter-balance window, the cost minimality of every split by brute force), cluster data equal
  to direct sums, the lists partitioning every ordered pair exactly once;
* Barnes-Hut energy / gradient: theta = 0 equals the exact evaluation, central differences
  with frozen lists, the O(theta^2) error and the authors' Barnes-Hut value within 1e-2;
* metric B, B0 (chi = 0): the Lean ``rs-*-metric`` fixtures; the null space; symmetry;
* preconditioner and saddle solve: the Lean ``rs-icosa-precond`` and ``rs-octa-saddle-none``
  fixtures, an independent numpy assembly of K and L', GMRES against a dense solve;
* the descent step (``Session.step``): the Lean ``rs-icosa-step-stuck`` fixture (barycenter only,
  exact energy, dense metric, absolute start, capped halvings: F0, slope, ||x||, lbar, tau0, every
  trial and the stuck result at the reference's 1e-10), plus the Armijo condition, the energy
  decrease and the kept barycenter of accepted steps.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "librs_oracle.so")
_lib = None

STATUS_NAMES = {
    0: "Ok",
    1: "InvalidOption",
    2: "InvalidExponents",
    3: "NonFiniteEnergy",
    4: "SolverDiverged",
    5: "SingularSaddle",
    6: "DegenerateFace",
    7: "NonFinitePosition",
}


class OracleError(RuntimeError):
    def __init__(self, status: int, detail: str = "", culprit=None):
        self.status = status
        self.kind = STATUS_NAMES.get(status, str(status))
        self.detail = detail
        self.culprit = culprit
        super().__init__(f"{self.kind}: {detail}")


def build() -> str:
    """Compiles the oracle with its Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_INT = ctypes.c_int


def _declare(L):
    L.orc_exact.argtypes = [_I64, _I64, _P, _P, _D, _D, _INT, _P, _P, _P, _P, ctypes.c_char_p, _I64]
    L.orc_exact.restype = _INT
    L.orc_bh_build.argtypes = [_I64, _I64, _P, _P, _D, _I64, _INT, ctypes.POINTER(_INT), ctypes.c_char_p, _I64]
    L.orc_bh_build.restype = _P
    L.orc_bh_free.argtypes = [_P]
    L.orc_bh_refit.argtypes = [_P, _P]
    L.orc_bh_refit.restype = _INT
    L.orc_bh_counts.argtypes = [_P, _P]
    L.orc_bh_tree.argtypes = [_P, _P, _P, _P, _P, _P]
    L.orc_bh_lists.argtypes = [_P, _P, _P, _P, _P]
    L.orc_bh_energy.argtypes = [_P, _D, _D, _INT, _P]
    L.orc_bh_energy.restype = _INT
    L.orc_bh_energy_and_differential.argtypes = [_P, _D, _D, _INT, _P, _P, _P]
    L.orc_bh_energy_and_differential.restype = _INT
    L.orc_metric_build.argtypes = [_I64, _I64, _P, _P, _D, _D, _D, _I64, _I64, _INT, ctypes.POINTER(_INT), ctypes.c_char_p, _I64]
    L.orc_metric_build.restype = _P
    L.orc_metric_free.argtypes = [_P]
    L.orc_metric_apply.argtypes = [_P, _D, _D, _I64, _P, _P, _INT]
    L.orc_metric_stats.argtypes = [_P, _P]
    L.orc_metric_ha.argtypes = [_P, _P]
    L.orc_vfrac_build.argtypes = [_I64, _P, _P, _D, _D, _I64, _I64, _INT, ctypes.POINTER(_INT)]
    L.orc_vfrac_build.restype = _P
    L.orc_vfrac_free.argtypes = [_P]
    L.orc_vfrac_apply.argtypes = [_P, _I64, _P, _P, _INT]
    L.orc_vertex_areas.argtypes = [_I64, _I64, _P, _P, _P]
    L.orc_saddle_build.argtypes = [_I64, _I64, _P, _P, _D, _D, _D, _I64, _I64, _D, _D, _INT, ctypes.POINTER(_INT),
                                   _I64, _I64, _I64]
    L.orc_saddle_build.restype = _P
    L.orc_saddle_free.argtypes = [_P]
    L.orc_saddle_eps.argtypes = [_P]
    L.orc_saddle_eps.restype = _D
    L.orc_saddle_frac_scale.argtypes = [_P]
    L.orc_saddle_frac_scale.restype = _D
    L.orc_saddle_apply.argtypes = [_P, _P, _P, _INT]
    L.orc_precond_apply.argtypes = [_P, _P, _P, _INT]
    L.orc_lprime_solve.argtypes = [_P, _P, _P]
    L.orc_saddle_solve.argtypes = [_P, _P, _P, _I64, _D, _I64, _INT, _P]
    L.orc_saddle_solve_history.argtypes = [_P, _P, _P, _I64, _D, _I64, _INT, _P, _P, _I64, _P, _I64, _P]
    L.orc_face_operators.argtypes = [_I64, _I64, _P, _P, _P, _P, _P]
    L.orc_metric_blocks.argtypes = [_P] * 10
    L.orc_vfrac_blocks.argtypes = [_P] * 10
    L.orc_session_create.argtypes = [_I64, _I64, _P, _P, _D, _D, _P, _P, ctypes.POINTER(_INT)]
    L.orc_session_create.restype = _P
    L.orc_session_free.argtypes = [_P]
    L.orc_session_step.argtypes = [_P, _INT, _P, _P, _I64]
    L.orc_session_step.restype = _INT
    L.orc_session_positions.argtypes = [_P, _P]
    L.orc_session_constraints.argtypes = [_P, _P, _I64]
    L.orc_constraint_rows.argtypes = [_I64, _I64, _P, _P, _P, _I64, _P, _P]
    L.orc_schur.argtypes = [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P]
    L.orc_session_ccd.restype = _INT
    L.orc_session_ccd.argtypes = [_P, _P, _P]
    L.orc_ccd_first_contact.restype = _INT
    L.orc_ccd_first_contact.argtypes = [_I64, _I64, _P, _P, _P, _D, _P, _P, _P, _P, _P]
    L.orc_intersecting_pairs.restype = _I64
    L.orc_intersecting_pairs.argtypes = [_I64, _I64, _P, _P, _P, _I64]
    L.orc_intersection_band.argtypes = [_I64, _I64, _P, _P, _P, _P]
    L.orc_min_distance.restype = _D
    L.orc_min_distance.argtypes = [_I64, _I64, _P, _P]
    L.orc_ccd_primitive.restype = _D
    L.orc_ccd_primitive.argtypes = [_INT, _P]
    L.orc_fractional_symbol_ratio.argtypes = [_D]
    L.orc_fractional_symbol_ratio.restype = _D
    L.orc_remesh.argtypes = [_I64, _I64, _P, _P, _I64, _P, _D, _P, _P, _D, _P, _P, ctypes.POINTER(_INT),
                             ctypes.c_char_p, _I64]
    L.orc_remesh.restype = _P
    L.orc_remesh_op.argtypes = [_I64, _I64, _P, _P, _I64, _P, _INT, ctypes.c_uint32, ctypes.c_uint32, _INT, _INT,
                                ctypes.POINTER(_INT), ctypes.POINTER(_INT), ctypes.POINTER(_INT), ctypes.c_char_p, _I64]
    L.orc_remesh_op.restype = _P
    L.orc_remeshed_sizes.argtypes = [_P, _P]
    L.orc_remeshed_read.argtypes = [_P] * 8
    L.orc_remeshed_free.argtypes = [_P]
    L.orc_remesh_geom.argtypes = [_INT, _P, _P]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _mesh(faces, positions):
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    return f, p


def exact(faces, positions, alpha=6.0, beta=12.0, threads=0):
    """All-pairs energy, differential (nv x 3) and per-face partials (nf x 7) in FP64
    (exact.rs:146 ``energy_and_differential``)."""
    f, p = _mesh(faces, positions)
    nv, nf = p.shape[0], f.shape[0]
    e = np.zeros(1)
    d = np.zeros((nv, 3))
    fp = np.zeros((nf, 7))
    cul = np.full(2, -1, dtype=np.int64)
    msg = ctypes.create_string_buffer(256)
    st = lib().orc_exact(nv, nf, _ptr(f), _ptr(p), alpha, beta, threads, _ptr(e), _ptr(d), _ptr(fp), _ptr(cul), msg, 256)
    if st != 0:
        raise OracleError(st, msg.value.decode(), tuple(cul))
    return float(e[0]), d, fp


class BhTree:
    """The oracle's Barnes-Hut evaluator (barnes_hut.rs:114 ``BhTree``)."""

    def __init__(self, faces, positions, theta=0.5, leaf_size=8, threads=0):
        f, p = _mesh(faces, positions)
        self.nv, self.nf = p.shape[0], f.shape[0]
        self.threads = threads
        st = _INT(0)
        msg = ctypes.create_string_buffer(256)
        self._h = lib().orc_bh_build(self.nv, self.nf, _ptr(f), _ptr(p), theta, leaf_size, threads, ctypes.byref(st), msg, 256)
        if not self._h:
            raise OracleError(st.value, msg.value.decode())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_bh_free(h)
            self._h = None

    def refit(self, positions):
        p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        st = lib().orc_bh_refit(self._h, _ptr(p))
        if st != 0:
            raise OracleError(st)

    def counts(self):
        c = np.zeros(6, dtype=np.int64)
        lib().orc_bh_counts(self._h, _ptr(c))
        keys = ["nodes", "leaves", "lumped", "exact_blocks", "pairs_exact", "partner_slots"]
        return dict(zip(keys, (int(x) for x in c)))

    def tree(self):
        c = self.counts()
        nn = c["nodes"]
        order = np.zeros(self.nf, dtype=np.uint32)
        start = np.zeros(nn, dtype=np.uint32)
        end = np.zeros(nn, dtype=np.uint32)
        right = np.zeros(nn, dtype=np.uint32)
        cl = np.zeros((nn, 17))
        lib().orc_bh_tree(self._h, _ptr(order), _ptr(start), _ptr(end), _ptr(right), _ptr(cl))
        return {"order": order, "start": start, "end": end, "right": right, "clusters": cl}

    def lists(self):
        c = self.counts()
        nl = c["leaves"]
        lo = np.zeros(nl + 1, dtype=np.uint64)
        lm = np.zeros(c["lumped"], dtype=np.uint32)
        eo = np.zeros(nl + 1, dtype=np.uint64)
        ex = np.zeros(c["exact_blocks"], dtype=np.uint32)
        lib().orc_bh_lists(self._h, _ptr(lo), _ptr(lm), _ptr(eo), _ptr(ex))
        return {"lumped_off": lo, "lumped": lm, "exact_off": eo, "exact": ex}

    def energy(self, alpha=6.0, beta=12.0):
        e = np.zeros(1)
        st = lib().orc_bh_energy(self._h, alpha, beta, self.threads, _ptr(e))
        if st != 0:
            raise OracleError(st)
        return float(e[0])

    def energy_and_differential(self, alpha=6.0, beta=12.0):
        e = np.zeros(1)
        d = np.zeros((self.nv, 3))
        fp = np.zeros((self.nf, 7))
        st = lib().orc_bh_energy_and_differential(self._h, alpha, beta, self.threads, _ptr(e), _ptr(d), _ptr(fp))
        if st != 0:
            raise OracleError(st)
        return float(e[0]), d, fp


def _block_lists(fn, h):
    c = np.zeros(4, dtype=np.int64)
    fn(h, _ptr(c), *([None] * 8))
    n, nn, nfar, nnear = (int(v) for v in c)
    names = ("order", "start", "end", "right", "far_ptr", "far_col", "near_ptr", "near_col")
    sizes = (n, nn, nn, nn, nn + 1, nfar, nn + 1, nnear)
    out = {k: np.zeros(m, dtype=np.uint32) for k, m in zip(names, sizes)}
    fn(h, _ptr(c), *[_ptr(out[k]) for k in names])
    return out


class FaceMetric:
    """The oracle's hierarchical B̃, B̃⁰ (hierarchical.rs:283 ``FaceMetricTree``)."""

    def __init__(self, faces, positions, alpha=6.0, beta=12.0, chi=0.5, leaf_size=8, leaf_pairs=64, threads=0):
        f, p = _mesh(faces, positions)
        self.nv, self.nf, self.threads = p.shape[0], f.shape[0], threads
        st = _INT(0)
        msg = ctypes.create_string_buffer(256)
        self._h = lib().orc_metric_build(self.nv, self.nf, _ptr(f), _ptr(p), alpha, beta, chi, leaf_size, leaf_pairs,
                                         threads, ctypes.byref(st), msg, 256)
        if not self._h:
            raise OracleError(st.value, msg.value.decode())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_metric_free(h)
            self._h = None

    def apply(self, x, w_hi=1.0, w_lo=1.0):
        """y = (w_hi B̃ + w_lo B̃⁰) x for x of shape (nb, 3 nv) or (3 nv,)."""
        xs = np.ascontiguousarray(x, dtype=np.float64)
        one = xs.ndim == 1
        xs = xs.reshape(-1, 3 * self.nv)
        ys = np.zeros_like(xs)
        lib().orc_metric_apply(self._h, w_hi, w_lo, xs.shape[0], _ptr(xs), _ptr(ys), self.threads)
        return ys[0] if one else ys

    def stats(self):
        o = np.zeros(7, dtype=np.int64)
        lib().orc_metric_stats(self._h, _ptr(o))
        keys = ["admissible_blocks", "inadmissible_blocks", "admissible_pairs", "near_entries", "nodes", "depth", "near_total"]
        return dict(zip(keys, (int(x) for x in o)))

    def blocks(self):
        """The face tree and the admissible / inadmissible block lists per row node in creation
        order (block_tree.rs:129-209)."""
        return _block_lists(lib().orc_metric_blocks, self._h)

    def ha(self):
        out = np.zeros((self.nf, 2))
        lib().orc_metric_ha(self._h, _ptr(out))
        return out


class VertexFractional:
    """The oracle's vertex-pair L̃^V_τ (hierarchical.rs:1003 ``VertexFractionalTree``)."""

    def __init__(self, positions, mass, tau, chi=0.5, leaf_size=8, leaf_pairs=64, threads=0):
        p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        m = np.ascontiguousarray(mass, dtype=np.float64)
        self.nv, self.threads = p.shape[0], threads
        st = _INT(0)
        self._h = lib().orc_vfrac_build(self.nv, _ptr(p), _ptr(m), tau, chi, leaf_size, leaf_pairs, threads, ctypes.byref(st))
        if not self._h:
            raise OracleError(st.value, "vertex fractional build")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_vfrac_free(h)
            self._h = None

    def apply(self, x):
        xs = np.ascontiguousarray(x, dtype=np.float64)
        one = xs.ndim == 1
        xs = xs.reshape(-1, 3 * self.nv)
        ys = np.zeros_like(xs)
        lib().orc_vfrac_apply(self._h, xs.shape[0], _ptr(xs), _ptr(ys), self.threads)
        return ys[0] if one else ys

    def blocks(self):
        """The vertex tree and its block lists (as ``FaceMetric.blocks``)."""
        return _block_lists(lib().orc_vfrac_blocks, self._h)


def vertex_areas(faces, positions):
    f, p = _mesh(faces, positions)
    out = np.zeros(p.shape[0])
    lib().orc_vertex_areas(p.shape[0], f.shape[0], _ptr(f), _ptr(p), _ptr(out))
    return out


class Saddle:
    """The oracle's hierarchical saddle A′ with the augmented vertex-pair preconditioner and a global
    barycenter row (solve/iterative.rs:126, solve/precond.rs:645)."""

    FORMS = {"augmented": 0, "congruence": 1}
    SCALES = {"symbol": 0, "unit": 1}

    def __init__(self, faces, positions, alpha=6.0, beta=12.0, chi=0.5, leaf_size=8, leaf_pairs=64, eps_rel=1e-10,
                 threads=0, c_lo=1.0, form="augmented", scale="symbol", refine=False):
        """``form``, ``scale`` and ``refine`` are PreconditionerOptions' (solve/precond.rs:466)."""
        f, p = _mesh(faces, positions)
        self.nv, self.threads = p.shape[0], threads
        self.n = 3 * self.nv + 3
        st = _INT(0)
        self._h = lib().orc_saddle_build(self.nv, f.shape[0], _ptr(f), _ptr(p), alpha, beta, chi, leaf_size, leaf_pairs,
                                         eps_rel, c_lo, threads, ctypes.byref(st), self.FORMS[form], self.SCALES[scale],
                                         1 if refine else 0)
        if not self._h:
            raise OracleError(st.value, "saddle build")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_saddle_free(h)
            self._h = None

    @property
    def eps(self):
        return lib().orc_saddle_eps(self._h)

    @property
    def frac_scale(self):
        return lib().orc_saddle_frac_scale(self._h)

    def _vec(self, fn, x, *extra):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        fn(self._h, _ptr(x), _ptr(y), *extra)
        return y

    def apply(self, x):
        return self._vec(lib().orc_saddle_apply, x, self.threads)

    def precond(self, r):
        return self._vec(lib().orc_precond_apply, r, self.threads)

    def lprime_solve(self, r):
        return self._vec(lib().orc_lprime_solve, r)

    def solve(self, rhs, restart=50, tol=1e-6, maxit=200):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros(self.n)
        rep = np.zeros(6)
        cap = int(maxit) + int(maxit) // max(int(restart), 1) + 4
        cyc, hist, counts = np.zeros(cap), np.zeros(cap), np.zeros(2, dtype=np.int64)
        lib().orc_saddle_solve_history(self._h, _ptr(rhs), _ptr(x), restart, tol, maxit, self.threads, _ptr(rep),
                                       _ptr(cyc), cap, _ptr(hist), cap, _ptr(counts))
        keys = ["iterations", "cycles", "converged", "relative_residual", "operator_applications",
                "preconditioner_applications"]
        out = dict(zip(keys, rep.tolist()))
        out["cycle_residuals"] = cyc[: min(int(counts[0]), cap)].tolist()
        out["history"] = hist[: min(int(counts[1]), cap)].tolist()
        return x, out


class Session:
    """The oracle's Hs descent session (step/session.rs:280) with Barnes–Hut or exact energy, the
    hierarchical saddle (χ = 0: the dense metric), the global barycenter constraint, remeshing and
    CCD off.

    ``start`` is the line search's first step (step/options.rs ``StartRule``, session.rs:1604):
    ``("grow", lengths)`` = GrowPrevious{2, MaxDisplacement{lengths}} (the default),
    ``("abs", tau)``, ``("maxdisp", lengths)`` or ``("reference",)`` (2/‖x‖).
    ``max_halvings`` < 0: no cap.

    ``constraints``: nonlinear rows after the barycenter's, at most two of ``("area", target)`` and
    ``("volume", target)`` (target None: the value at creation), enforced through the Schur
    complement (solve/schur.rs) and the projection's Newton rounds: ``newton`` = ``("tol", n)``
    (UntilTolerance{max: n}) or ``("fixed", n)``."""

    START_KINDS = {"grow": 0, "abs": 1, "maxdisp": 2, "reference": 3}
    ROW_KINDS = {"area": 1, "volume": 2}

    def __init__(self, faces, positions, alpha=6.0, beta=12.0, theta=0.5, chi=0.5, restart=50, tol=1e-6, maxit=200,
                 energy_exact=False, start=("grow", 1.0), armijo_c=1e-2, tau_min_rel=1e-12, max_halvings=-1,
                 c_lo=1.0, leaf_size=8, constraints=(), newton=("tol", 5), ccd=None):
        f, p = _mesh(faces, positions)
        self.nv = p.shape[0]
        st = _INT(0)
        kind = self.START_KINDS[start[0]]
        value = float(start[1]) if len(start) > 1 else 0.0
        if len(constraints) > 2:
            raise ValueError("at most two nonlinear rows")
        kinds = [self.ROW_KINDS[c[0]] for c in constraints] + [0] * (2 - len(constraints))
        targets = [np.nan if c[1] is None else float(c[1]) for c in constraints] + [np.nan] * (2 - len(constraints))
        self.n_rows = 3 + len(constraints)
        c = dict(gap=0.05, reference_length=None, max_iterations=100, close_factor=0.5, leaf_size=8)
        c.update(ccd or {})
        ref = np.nan if c["reference_length"] is None else float(c["reference_length"])
        of = np.array([theta, chi, tol, value, armijo_c, tau_min_rel, c_lo] + targets +
                      [c["gap"], ref, c["close_factor"]], dtype=np.float64)
        oi = np.array([restart, maxit, 1 if energy_exact else 0, kind, max_halvings, leaf_size] + kinds +
                      [1 if newton[0] == "fixed" else 0, int(newton[1]), 0 if ccd is None else 1,
                       c["max_iterations"], c["leaf_size"]], dtype=np.int64)
        self._h = lib().orc_session_create(self.nv, f.shape[0], _ptr(f), _ptr(p), alpha, beta, _ptr(of), _ptr(oi),
                                           ctypes.byref(st))
        if not self._h:
            raise OracleError(st.value, "session")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_session_free(h)
            self._h = None

    def step(self, threads=0):
        """One step; the report's ``trials`` is a list of (τ, F, outcome) with outcome 0 accepted,
        1 Armijo failed, 2 non-finite."""
        cap = 256
        rep = np.zeros(21)
        tr = np.zeros((cap, 3))
        st = lib().orc_session_step(self._h, threads, _ptr(rep), _ptr(tr), cap)
        if st != 0:
            raise OracleError(st, "step")
        keys = ["energy0", "energy1", "slope", "tau", "accepted", "trials", "gmres_iterations", "gmres_converged",
                "gmres_relative_residual", "energy_after", "tau0", "gradient_norm", "direction_norm", "direction_max",
                "halvings", "lbar"]
        out = dict(zip(keys, rep.tolist()))
        out["seconds"] = dict(zip(["trees_energy", "solver_build", "gmres", "line_search", "projection"],
                                  rep[16:21].tolist()))
        out["trial_list"] = [(float(t), float(e), int(o)) for t, e, o in tr[: min(cap, int(out["trials"]))]]
        c = np.zeros(3 + 2 * self.n_rows)
        lib().orc_session_constraints(self._h, _ptr(c), c.size)
        out["newton_rounds"] = int(c[1])
        out["min_pivot"] = float(c[2])
        out["phi_before"] = c[3:3 + self.n_rows].tolist()
        out["phi"] = c[3 + self.n_rows:].tolist()
        of, oi = np.zeros(3), np.zeros(10, dtype=np.int64)
        if lib().orc_session_ccd(self._h, _ptr(of), _ptr(oi)):
            kinds = ("vertex_triangle", "edge_edge")
            out["ccd"] = {"tau_max": float(of[0]), "gap_abs": float(of[1]),
                          "contact": (float(of[2]), kinds[int(oi[1])], int(oi[2]), int(oi[3])) if oi[0] else None,
                          "iterations": int(oi[4]), "pairs_tested": int(oi[5]), "windows": int(oi[6]),
                          "close_pairs": int(oi[7]), "touching_pairs": int(oi[8]), "exempt_pairs": int(oi[9])}
        else:
            out["ccd"] = None
        return out

    def positions(self):
        out = np.zeros((self.nv, 3))
        lib().orc_session_positions(self._h, _ptr(out))
        return out


def constraint_rows(faces, positions, kinds):
    """(values, gradients (k × 3 nv)) of the rows ``kinds`` ("area", "volume") at the positions
    (constraints/mod.rs:462-647)."""
    f, p = _mesh(faces, positions)
    k = np.array([Session.ROW_KINDS[x] for x in kinds], dtype=np.int64)
    vals, grads = np.zeros(len(k)), np.zeros((len(k), 3 * p.shape[0]))
    lib().orc_constraint_rows(p.shape[0], f.shape[0], _ptr(f), _ptr(p), _ptr(k), len(k), _ptr(vals), _ptr(grads))
    return vals, grads


def schur(c, z, y, phi):
    """(x = y − Z N⁻¹ C_N y, h = Z N⁻¹ Φ_N, S = −C_N Z, smallest kept pivot) for the rows c (k × nd)
    and the columns z = A′⁻¹ C_Nᵀ (k × nd, position blocks) (solve/schur.rs)."""
    c = np.ascontiguousarray(c, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    k, nd = c.shape
    y = np.ascontiguousarray(y, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    x, h, S, mp = np.zeros(nd), np.zeros(nd), np.zeros((k, k)), ctypes.c_double()
    lib().orc_schur(k, nd, _ptr(c), _ptr(z), _ptr(y), _ptr(phi), _ptr(x), _ptr(h), _ptr(S), ctypes.byref(mp))
    return x, h, S, mp.value


def fractional_symbol_ratio(s):
    return lib().orc_fractional_symbol_ratio(s)


def face_operators(faces, positions, u):
    """(ū per face, D_f u per face (nf × 3)) for one vertex function u (metric/mod.rs:200-225)."""
    f, p = _mesh(faces, positions)
    u = np.ascontiguousarray(u, dtype=np.float64)
    avg, grad = np.zeros(f.shape[0]), np.zeros((f.shape[0], 3))
    lib().orc_face_operators(p.shape[0], f.shape[0], _ptr(f), _ptr(p), _ptr(u), _ptr(avg), _ptr(grad))
    return avg, grad


# ---- continuous collision detection (ccd.rs) ------------------------------------------------


def ccd_first_contact(faces, positions, velocity, tau_end, gap=0.05, reference_length=None, max_iterations=100,
                      close_factor=0.5, leaf_size=8, band=None):
    """``first_contact_with_band`` (ccd.rs:781) of the motion positions + t velocity on [0, tau_end],
    mesh against mesh.  Returns a dict: tau_max, gap_abs, contact (None or (tau, kind, a, b): kind
    "vertex_triangle" (vertex a, triangle b) or "edge_edge" (edge ids a < b of the sorted edge
    list)), iterations, pairs_tested, windows, close_pairs, touching_pairs, exempt_pairs."""
    f, p = _mesh(faces, positions)
    u = np.ascontiguousarray(velocity, dtype=np.float64).reshape(-1, 3)
    nv = p.shape[0]
    if u.shape[0] != nv:
        raise OracleError(1, "velocity: one vector per vertex")
    of = np.array([gap, np.nan if reference_length is None else reference_length, close_factor])
    oi = np.array([max_iterations, leaf_size], dtype=np.int64)
    b = None if band is None else np.ascontiguousarray(band, dtype=np.uint8)
    out_f, out_i = np.zeros(3), np.zeros(10, dtype=np.int64)
    st = lib().orc_ccd_first_contact(nv, f.shape[0], _ptr(f), _ptr(p), _ptr(u), float(tau_end), _ptr(of), _ptr(oi),
                                     None if b is None else _ptr(b), _ptr(out_f), _ptr(out_i))
    if st != 0:
        raise OracleError(st, "ccd")
    kinds = ("vertex_triangle", "edge_edge")
    return {
        "tau_max": float(out_f[0]),
        "gap_abs": float(out_f[1]),
        "contact": (float(out_f[2]), kinds[int(out_i[1])], int(out_i[2]), int(out_i[3])) if out_i[0] else None,
        "iterations": int(out_i[4]),
        "pairs_tested": int(out_i[5]),
        "windows": int(out_i[6]),
        "close_pairs": int(out_i[7]),
        "touching_pairs": int(out_i[8]),
        "exempt_pairs": int(out_i[9]),
    }


def intersecting_pairs(faces, positions):
    """The sorted face pairs (k × 2) that intersect (ccd.rs:1229)."""
    f, p = _mesh(faces, positions)
    n = lib().orc_intersecting_pairs(p.shape[0], f.shape[0], _ptr(f), _ptr(p), None, 0)
    out = np.zeros((n, 2), dtype=np.uint32)
    if n:
        lib().orc_intersecting_pairs(p.shape[0], f.shape[0], _ptr(f), _ptr(p), _ptr(out), n)
    return out


def intersection_band(faces, positions):
    """(per-vertex flags, intersecting pairs, vertices in the band) of ``IntersectionBand::new``
    (ccd.rs:1292)."""
    f, p = _mesh(faces, positions)
    flags, counts = np.zeros(p.shape[0], dtype=np.uint8), np.zeros(2, dtype=np.int64)
    lib().orc_intersection_band(p.shape[0], f.shape[0], _ptr(f), _ptr(p), _ptr(flags), _ptr(counts))
    return flags, int(counts[0]), int(counts[1])


def min_distance(faces, positions):
    """The smallest distance over the non-adjacent vertex-triangle and edge-edge pairs (ccd.rs:1364)."""
    f, p = _mesh(faces, positions)
    return lib().orc_min_distance(p.shape[0], f.shape[0], _ptr(f), _ptr(p))


def ccd_primitive(kind, *points):
    """The curves crate's distances and the static tests on points (3-vectors): "point_triangle"
    (p, a, b, c), "segment_segment" (p1, q1, p2, q2), "point_segment" (p, a, b),
    "segment_triangle" (p, q, a, b, c: 1.0 if they meet), "triangles" (6 points: 1.0 if they
    intersect)."""
    k = {"point_triangle": 0, "segment_segment": 1, "point_segment": 2, "segment_triangle": 3, "triangles": 4}[kind]
    x = np.ascontiguousarray(np.array(points, dtype=np.float64).reshape(-1))
    return lib().orc_ccd_primitive(k, _ptr(x))


# ---- remeshing (remesh/mod.rs, remesh/mesh.rs) ----------------------------------------------

# mesh.rs:75 ``RefusalKind::ALL``
REFUSAL_KINDS = ("NoSuchEdge", "BothPinned", "LinkCondition", "Foldover", "RepeatedFace", "NonManifoldResult",
                 "BoundaryEdge", "NotDelaunay", "ExistingEdge", "Guard", "DegenerateFace")
_REPORT_HEAD = ("splits", "collapses", "flips", "splits_deferred", "smoothing_moves", "smoothing_halved",
                "smoothing_refused", "smoothing_skipped", "flip_cap_hits", "n_vertices", "n_faces", "guard_refused",
                "guard_tested", "guard_exempt")


def _remeshed(h) -> dict:
    """Reads and frees a Remeshed handle: positions, faces, face_parents, vertex map (CSR), pinned."""
    L = lib()
    try:
        n = np.zeros(4, dtype=np.int64)
        L.orc_remeshed_sizes(h, _ptr(n))
        nv, nf, nm, npin = (int(x) for x in n)
        out = {"positions": np.zeros((nv, 3)), "faces": np.zeros((nf, 3), dtype=np.uint32),
               "face_parents": np.zeros(nf, dtype=np.uint32), "map_offsets": np.zeros(nv + 1, dtype=np.uint32),
               "map_sources": np.zeros(nm, dtype=np.uint32), "map_weights": np.zeros(nm),
               "pinned": np.zeros(npin, dtype=np.uint32)}
        L.orc_remeshed_read(h, *[_ptr(out[k]) for k in ("positions", "faces", "face_parents", "map_offsets",
                                                         "map_sources", "map_weights", "pinned")])
        return out
    finally:
        L.orc_remeshed_free(h)


def remesh(faces, positions, l0, pinned=(), split=1.5, collapse=0.5, rounds=5, rho=0.5, flip_cap_per_vertex=100,
           order="flip_then_smooth", enabled=("split", "collapse", "flip", "smooth"), max_growth=2.0,
           guard_gap=None, band=None):
    """One remeshing pass (mod.rs:235 ``remesh``).  ``guard_gap``: the absolute gap of the CCD
    collapse guard (ccd.rs:1448; None: no guard); ``band``: its per-vertex exempt flags.  Returns
    (result dict as ``_remeshed``, report dict with the refusal counts per kind)."""
    f, p = _mesh(faces, positions)
    pins = np.ascontiguousarray(pinned, dtype=np.uint32)
    of = np.array([split, collapse, rho, max_growth], dtype=np.float64)
    oi = np.array([rounds, flip_cap_per_vertex, 0 if order == "flip_then_smooth" else 1,
                   "split" in enabled, "collapse" in enabled, "flip" in enabled, "smooth" in enabled], dtype=np.int64)
    b = None if band is None else np.ascontiguousarray(band, dtype=np.uint8)
    rep = np.zeros(47, dtype=np.int64)
    st = _INT(0)
    msg = ctypes.create_string_buffer(512)
    h = lib().orc_remesh(p.shape[0], f.shape[0], _ptr(f), _ptr(p), pins.size, _ptr(pins), float(l0), _ptr(of),
                         _ptr(oi), 0.0 if guard_gap is None else float(guard_gap), None if b is None else _ptr(b),
                         _ptr(rep), ctypes.byref(st), msg, 512)
    if not h:
        raise OracleError(st.value, msg.value.decode())
    report = {k: int(v) for k, v in zip(_REPORT_HEAD, rep[:14])}
    for j, name in enumerate(("refused_splits", "refused_collapses", "refused_flips")):
        report[name] = {k: int(v) for k, v in zip(REFUSAL_KINDS, rep[14 + 11 * j: 25 + 11 * j])}
    return _remeshed(h), report


def remesh_op(faces, positions, op, a, b, pinned=(), delaunay=False, foldover=False):
    """One ``EditMesh`` operation (mesh.rs:592 split, :679 collapse, :888 flip) and ``finish``.
    Returns (refusal kind or None, reordered, result dict)."""
    f, p = _mesh(faces, positions)
    pins = np.ascontiguousarray(pinned, dtype=np.uint32)
    ref, reo, st = _INT(0), _INT(0), _INT(0)
    msg = ctypes.create_string_buffer(512)
    h = lib().orc_remesh_op(p.shape[0], f.shape[0], _ptr(f), _ptr(p), pins.size, _ptr(pins),
                            {"split": 0, "collapse": 1, "flip": 2}[op], a, b, int(delaunay), int(foldover),
                            ctypes.byref(ref), ctypes.byref(reo), ctypes.byref(st), msg, 512)
    if not h:
        raise OracleError(st.value, msg.value.decode())
    return (None if ref.value == 0 else REFUSAL_KINDS[ref.value - 1]), bool(reo.value), _remeshed(h)


def remesh_geom(kind, *points):
    """The remesher's predicates (geom.rs): "circumcenter" (3 points → 3-vector), "delaunay"
    (p, q, c, d → bool), "folds" (before, after normals → bool), "triangle_distance" (6 points)."""
    k = {"circumcenter": 0, "delaunay": 1, "folds": 2, "triangle_distance": 3}[kind]
    x = np.ascontiguousarray(np.array(points, dtype=np.float64).reshape(-1))
    out = np.zeros(3)
    lib().orc_remesh_geom(k, _ptr(x), _ptr(out))
    if k == 0:
        return out
    if k in (1, 2):
        return bool(out[0])
    return float(out[0])
