"""ddg-tools_amd — the MI355X (gfx950) Hs-descent path of ``ddg-repulsive-surfaces``.

Python mirror of the reference's API for this path, over the C ABI of ``include/ddgrs.h``
(``lib/libddgrs_hip.so``, built by ``build.py``).  Names, argument meaning and errors follow the
reference (``crates/ddg-repulsive-surfaces``):

* ``Context`` — ``GpuContext`` / ``Device::Cuda`` (``src/gpu/mod.rs:47-151``);
* ``Surface`` — the validated mesh with device-resident positions (``src/surface.rs:132,513``);
* ``energy(surface, exponents, eval)`` — ``energy()`` (``src/energy/mod.rs:168``) for
  ``EnergyEval.Exact`` and ``EnergyEval.BarnesHut``;
* ``BhTree``, ``FaceMetricTree``, ``VertexFractionalTree`` — the Barnes–Hut tree and the block
  cluster tree products (``src/energy/barnes_hut.rs``, ``src/metric/hierarchical.rs``);
* ``Saddle`` — the step's saddle solver (``hierarchical_saddle``, ``src/solve/iterative.rs:455``);
* ``Session`` — the Hs descent session (``Session::new`` / ``step``, ``src/step/session.rs:371,724``);
* errors raise ``RsError`` whose ``kind`` names the reference's ``RsError`` variant.

The library is the only compute path: if ``lib/libddgrs_hip.so`` is missing, importing the
binding raises; there is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes
import dataclasses
import os
import weakref
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libddgrs_hip.so")

KINDS = {
    0: "Ok",
    1: "InvalidOption",
    2: "InvalidExponents",
    3: "NonFiniteEnergy",
    4: "SolverDiverged",
    5: "SingularSaddle",
    6: "DegenerateFace",
    7: "NonFinitePosition",
    8: "Gpu",
    # raised by the Python layer (io.py, api.py)
    9: "Parse",
    10: "Io",
    11: "PolygonFace",
    12: "Unsupported",
}


class RsError(RuntimeError):
    """A failed call: ``kind`` (the reference's RsError variant), ``operation``, ``detail``."""

    def __init__(self, status, operation, detail, a=-1, b=-1, value=0.0):
        self.status = status
        self.kind = KINDS.get(status, str(status))
        self.operation = operation
        self.detail = detail
        self.a, self.b, self.value = a, b, value
        super().__init__(f"{self.kind} [{operation}]: {detail}")


class _Err(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("operation", ctypes.c_char * 64),
        ("detail", ctypes.c_char * 256),
        ("a", ctypes.c_int64),
        ("b", ctypes.c_int64),
        ("value", ctypes.c_double),
    ]


_lib = None

# Live handles by teardown rank: sessions (which point at attached learned networks) first, then
# the other handles built on a surface, then surfaces, then contexts.  The atexit hook closes them
# in that order while the interpreter and the HIP runtime are still whole, so no destroy call is
# left to the finalizers of interpreter shutdown, whose order is arbitrary.
_LIVE = [weakref.WeakSet() for _ in range(4)]


def _register(obj, rank: int):
    _LIVE[rank].add(obj)


@atexit.register
def close_all():
    """Closes every live handle in dependency order (also usable before a process forks or ends)."""
    for live in _LIVE:
        for obj in list(live):
            obj.close()


def library_path() -> str:
    return LIB_PATH


def lib():
    """The loaded C ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with ddg-tools_amd/build.py (no CPU fallback exists)")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double
_E = ctypes.POINTER(_Err)

# name -> (restype, argtypes); every int-returning entry point takes a trailing ddgrs_error*.
SIGNATURES = {
    "ddgrs_ctx_create": (_I32, [_I32, _I32, ctypes.POINTER(_P), _E]),
    "ddgrs_ctx_destroy": (None, [_P]),
    "ddgrs_device_info": (_I32, [_P, ctypes.c_char_p, _I64, ctypes.POINTER(_I64), ctypes.POINTER(_I64), _E]),
    "ddgrs_ctx_synchronize": (_I32, [_P, _E]),
    "ddgrs_ctx_set_profiling": (_I32, [_P, _I32, _E]),
    "ddgrs_ctx_kernel_times": (_I32, [_P, _I64, _P, _P, _P, ctypes.POINTER(_I64), _E]),
    "ddgrs_surface_create": (_I32, [_P, _I64, _I64, _P, _P, ctypes.POINTER(_P), _E]),
    "ddgrs_surface_destroy": (None, [_P]),
    "ddgrs_surface_set_positions": (_I32, [_P, _P, _E]),
    "ddgrs_surface_get_positions": (_I32, [_P, _P, _E]),
    "ddgrs_surface_edges": (_I32, [_P, _P, _I64, ctypes.POINTER(_I64), _E]),
    "ddgrs_ccd_options_default": (None, [_P]),
    "ddgrs_ccd_first_contact": (_I32, [_P, _P, _P, _D, _P, _P, _P, _E]),
    "ddgrs_ccd_intersecting_pairs": (_I32, [_P, _P, _P, _I64, ctypes.POINTER(_I64), _E]),
    "ddgrs_ccd_band": (_I32, [_P, _P, _P, _P, _E]),
    "ddgrs_exact_energy_and_differential": (_I32, [_P, _D, _D, ctypes.POINTER(_D), _P, _P, _E]),
    "ddgrs_exact_energy_and_differential_noalloc": (_I32, [_P, _D, _D, ctypes.POINTER(_D), _P, _P, _E]),
    "ddgrs_bh_build": (_I32, [_P, _D, _I64, ctypes.POINTER(_P), _E]),
    "ddgrs_bh_destroy": (None, [_P]),
    "ddgrs_bh_refit": (_I32, [_P, _P, _E]),
    "ddgrs_bh_energy": (_I32, [_P, _D, _D, ctypes.POINTER(_D), _E]),
    "ddgrs_bh_energy_and_differential": (_I32, [_P, _D, _D, ctypes.POINTER(_D), _P, _P, _E]),
    "ddgrs_bh_counts": (_I32, [_P, _P, _E]),
    "ddgrs_bh_tree": (_I32, [_P, _P, _P, _P, _P, _P, _E]),
    "ddgrs_bh_lists": (_I32, [_P, _P, _P, _P, _P, _E]),
    "ddgrs_metric_build": (_I32, [_P, _D, _D, _D, _I64, _I64, ctypes.POINTER(_P), _E]),
    "ddgrs_metric_destroy": (None, [_P]),
    "ddgrs_metric_apply_batch": (_I32, [_P, _D, _D, _I64, _P, _P, _E]),
    "ddgrs_metric_stats": (_I32, [_P, _P, _E]),
    "ddgrs_metric_blocks": (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _E]),
    "ddgrs_vfrac_blocks": (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _E]),
    "ddgrs_vfrac_build": (_I32, [_P, _D, _D, _I64, _I64, ctypes.POINTER(_P), _E]),
    "ddgrs_vfrac_destroy": (None, [_P]),
    "ddgrs_vfrac_apply_batch": (_I32, [_P, _I64, _P, _P, _E]),
    "ddgrs_saddle_options_default": (None, [_P]),
    "ddgrs_saddle_build": (_I32, [_P, _D, _D, _P, ctypes.POINTER(_P), _E]),
    "ddgrs_saddle_destroy": (None, [_P]),
    "ddgrs_saddle_info": (_I32, [_P, _P, _E]),
    "ddgrs_saddle_apply_batch": (_I32, [_P, _I64, _P, _P, _E]),
    "ddgrs_saddle_precond_batch": (_I32, [_P, _I64, _P, _P, _E]),
    "ddgrs_saddle_lprime_batch": (_I32, [_P, _I64, _P, _P, _E]),
    "ddgrs_saddle_solve": (_I32, [_P, _P, _P, _I64, _D, _I64, _P, _E]),
    "ddgrs_saddle_solve_many": (_I32, [_P, _I64, _P, _P, _I64, _D, _I64, _P, _E]),
    "ddgrs_step_options_default": (None, [_P]),
    "ddgrs_session_create": (_I32, [_P, _P, ctypes.POINTER(_P), _E]),
    "ddgrs_session_destroy": (None, [_P]),
    "ddgrs_session_step": (_I32, [_P, _P, _E]),
    "ddgrs_session_state": (_I32, [_P, _P, _E]),
    "ddgrs_session_set_state": (_I32, [_P, _P, _E]),
    "ddgrs_session_row_targets": (_I32, [_P, _P, _E]),
    "ddgrs_session_trials": (_I32, [_P, _I64, _P, _P, _P, ctypes.POINTER(_I64), _E]),
    "ddgrs_session_fallbacks": (_I32, [_P, ctypes.c_char_p, _I64, ctypes.POINTER(_I64), _E]),
    "ddgrs_session_gmres_history": (_I32, [_P, _P, _I64, _P, _I64, _P, _E]),
    "ddgrs_session_constraints": (_I32, [_P, _P, _P, _I64, ctypes.POINTER(_I64), _E]),
    "ddgrs_schur": (_I32, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _E]),
    "ddgrs_saddle_gmres_history": (_I32, [_P, _P, _I64, _P, _I64, _P, _E]),
    "ddgrs_learned_create": (_I32, [_P, ctypes.c_uint64, _I64, ctypes.POINTER(_P), _E]),
    "ddgrs_learned_destroy": (None, [_P]),
    "ddgrs_learned_info": (_I32, [_P, _P, _P, _E]),
    "ddgrs_learned_params": (_I32, [_P, _P, _P, _E]),
    "ddgrs_learned_set_scale": (_I32, [_P, _D, _E]),
    "ddgrs_learned_add_pair": (_I32, [_P, _P, _P, _P, _P, _E]),
    "ddgrs_learned_set_krylov_pairs": (_I32, [_P, _I32, _E]),
    "ddgrs_learned_features": (_I32, [_P, _P, _P, _E]),
    "ddgrs_learned_blocks": (_I32, [_P, _P, _P, _E]),
    "ddgrs_learned_loss_grad": (_I32, [_P, _I64, _P, _P, _E]),
    "ddgrs_learned_train": (_I32, [_P, _I64, _I64, _D, _P, _E]),
    "ddgrs_learned_apply_batch": (_I32, [_P, _P, _I64, _P, _P, _E]),
    "ddgrs_session_set_learned": (_I32, [_P, _P, _I32, _E]),
    "ddgrs_remesh_options_default": (None, [_P]),
    "ddgrs_remesh": (_I32, [_P, _I64, _I64, _P, _P, _I64, _P, _D, _P, _P, ctypes.POINTER(_P), _P, _E]),
    "ddgrs_remesh_op": (_I32, [_P, _I64, _I64, _P, _P, _I64, _P, _I32, ctypes.c_uint32, ctypes.c_uint32, _I32, _I32,
                                ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_P), _E]),
    "ddgrs_remeshed_sizes": (_I32, [_P, _P, _E]),
    "ddgrs_remeshed_read": (_I32, [_P, _P, _P, _P, _P, _P, _P, _P, _E]),
    "ddgrs_remeshed_destroy": (None, [_P]),
}


def _declare(L):
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _call(fn, *args):
    err = _Err()
    st = fn(*args, ctypes.byref(err))
    if st != 0:
        raise RsError(st, err.operation.decode(errors="replace"), err.detail.decode(errors="replace"), err.a, err.b, err.value)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a, shape=None) -> np.ndarray:
    x = np.ascontiguousarray(a, dtype=np.float64)
    return x.reshape(shape) if shape is not None else x


class Context:
    """One device and one in-order stream (``GpuContext``, src/gpu/mod.rs:129)."""

    def __init__(self, ordinal: int = 0, determinism: int = 1):
        h = _P()
        _call(lib().ddgrs_ctx_create, ordinal, determinism, ctypes.byref(h))
        self._h = h
        _register(self, 3)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().ddgrs_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def device_info(self) -> dict:
        name = ctypes.create_string_buffer(256)
        cu, mem = _I64(), _I64()
        _call(lib().ddgrs_device_info, self._h, name, 256, ctypes.byref(cu), ctypes.byref(mem))
        return {"name": name.value.decode(), "compute_units": cu.value, "memory_bytes": mem.value}

    def synchronize(self):
        _call(lib().ddgrs_ctx_synchronize, self._h)

    def set_profiling(self, enabled: bool):
        _call(lib().ddgrs_ctx_set_profiling, self._h, 1 if enabled else 0)

    def kernel_times(self) -> dict:
        cap = 256
        names = ctypes.create_string_buffer(64 * cap)
        launches = np.zeros(cap, dtype=np.int64)
        ms = np.zeros(cap)
        n = _I64()
        _call(lib().ddgrs_ctx_kernel_times, self._h, cap, names, _ptr(launches), _ptr(ms), ctypes.byref(n))
        out = {}
        for i in range(min(n.value, cap)):
            nm = names.raw[64 * i: 64 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = {"launches": int(launches[i]), "ms": float(ms[i])}
        return out


class Surface:
    """A triangle mesh with its positions on the device (``Surface``, src/surface.rs:132)."""

    def __init__(self, ctx: Context, faces, positions):
        self.ctx = ctx
        self.faces = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        p = _f64(positions).reshape(-1, 3)
        self.nv, self.nf = p.shape[0], self.faces.shape[0]
        h = _P()
        _call(lib().ddgrs_surface_create, ctx.handle, self.nv, self.nf, _ptr(self.faces), _ptr(p), ctypes.byref(h))
        self._h = h
        _register(self, 2)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().ddgrs_surface_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_positions(self, positions):
        p = _f64(positions).reshape(-1, 3)
        if p.shape[0] != self.nv:
            raise RsError(1, "set_positions", f"{p.shape[0]} positions for {self.nv} vertices")
        _call(lib().ddgrs_surface_set_positions, self._h, _ptr(p))

    def positions(self) -> np.ndarray:
        out = np.zeros((self.nv, 3))
        _call(lib().ddgrs_surface_get_positions, self._h, _ptr(out))
        return out

    def edges(self) -> np.ndarray:
        """The sorted unique edges (ne × 2, a < b) (``Surface::edges``, src/surface.rs:564)."""
        ne = _I64()
        _call(lib().ddgrs_surface_edges, self._h, None, 0, ctypes.byref(ne))
        out = np.zeros((ne.value, 2), dtype=np.uint32)
        _call(lib().ddgrs_surface_edges, self._h, _ptr(out), ne.value, ctypes.byref(ne))
        return out


# ---- continuous collision detection (src/ccd.rs) --------------------------------------------


class _CcdOptionsC(ctypes.Structure):
    _fields_ = [("gap", _D), ("reference_length", _D), ("max_iterations", _I64), ("close_factor", _D),
                ("leaf_size", _I64)]


class _CcdReportC(ctypes.Structure):
    _fields_ = [
        ("tau_max", _D),
        ("has_contact", _I32),
        ("contact_kind", _I32),
        ("contact_a", _I64),
        ("contact_b", _I64),
        ("gap_abs", _D),
        ("iterations", _I64),
        ("pairs_tested", _I64),
        ("close_pairs", _I64),
        ("touching_pairs", _I64),
        ("exempt_pairs", _I64),
        ("windows", _I32),
        ("seconds", _D),
    ]


@dataclasses.dataclass(frozen=True)
class CcdOptions:
    """``CcdOptions`` (src/ccd.rs:86): gap relative to the reference length (None: the mean edge
    length of the start positions), advances per pair and window, the closing fraction of pairs
    that start inside the gap, the broad-phase leaf size."""

    gap: float = 0.05
    reference_length: Optional[float] = None
    max_iterations: int = 100
    close_factor: float = 0.5
    leaf_size: int = 8

    def _c(self) -> _CcdOptionsC:
        ref = float("nan") if self.reference_length is None else float(self.reference_length)
        return _CcdOptionsC(self.gap, ref, self.max_iterations, self.close_factor, self.leaf_size)


@dataclasses.dataclass(frozen=True)
class CcdReport:
    """``CcdReport`` (src/ccd.rs:214); ``contact`` is None or (τ, kind, a, b) with kind
    "vertex_triangle" (vertex a, triangle b) or "edge_edge" (edge ids a < b of ``Surface.edges``)."""

    tau_max: float
    contact: Optional[tuple]
    gap_abs: float
    iterations: int
    pairs_tested: int
    windows: int
    close_pairs: int
    touching_pairs: int
    exempt_pairs: int
    seconds: float


def first_contact(surface: Surface, velocity, tau_end: float, options: CcdOptions = CcdOptions(), start=None,
                  band=None) -> CcdReport:
    """``first_contact_with_band`` (src/ccd.rs:781) of start + t·velocity on [0, tau_end] on the
    device (start None: the surface's positions; band None or per-vertex flags)."""
    u = _f64(velocity).reshape(-1, 3)
    if u.shape[0] != surface.nv:
        raise RsError(1, "ccd_first_contact", f"{u.shape[0]} velocities for {surface.nv} vertices")
    st = None if start is None else _f64(start).reshape(-1, 3)
    if st is not None and st.shape[0] != surface.nv:
        raise RsError(1, "ccd_first_contact", f"{st.shape[0]} positions for {surface.nv} vertices")
    b = None if band is None else np.ascontiguousarray(band, dtype=np.uint8).reshape(-1)
    if b is not None and b.shape[0] != surface.nv:
        raise RsError(1, "ccd_first_contact", f"band: {b.shape[0]} flags for {surface.nv} vertices")
    o = options._c()
    r = _CcdReportC()
    _call(lib().ddgrs_ccd_first_contact, surface.handle, None if st is None else _ptr(st), _ptr(u), float(tau_end),
          ctypes.byref(o), None if b is None else _ptr(b), ctypes.byref(r))
    return _ccd_report(r)


def _ccd_report(r: _CcdReportC) -> CcdReport:
    kinds = ("vertex_triangle", "edge_edge")
    contact = (float(r.tau_max), kinds[r.contact_kind], int(r.contact_a), int(r.contact_b)) if r.has_contact else None
    return CcdReport(float(r.tau_max), contact, float(r.gap_abs), int(r.iterations), int(r.pairs_tested),
                     int(r.windows), int(r.close_pairs), int(r.touching_pairs), int(r.exempt_pairs), float(r.seconds))


def intersecting_pairs(surface: Surface, positions=None) -> np.ndarray:
    """The sorted face pairs (k × 2, f < g) that intersect (``intersecting_pairs``, src/ccd.rs:1229)."""
    p = None if positions is None else _f64(positions).reshape(-1, 3)
    n = _I64()
    _call(lib().ddgrs_ccd_intersecting_pairs, surface.handle, None if p is None else _ptr(p), None, 0, ctypes.byref(n))
    out = np.zeros((n.value, 2), dtype=np.uint32)
    if n.value:
        _call(lib().ddgrs_ccd_intersecting_pairs, surface.handle, None if p is None else _ptr(p), _ptr(out), n.value,
              ctypes.byref(n))
    return out


def intersection_band(surface: Surface, positions=None):
    """``IntersectionBand::new`` (src/ccd.rs:1292): (per-vertex flags, intersecting pairs, vertices
    in the band)."""
    p = None if positions is None else _f64(positions).reshape(-1, 3)
    flags = np.zeros(surface.nv, dtype=np.uint8)
    counts = np.zeros(2, dtype=np.int64)
    _call(lib().ddgrs_ccd_band, surface.handle, None if p is None else _ptr(p), _ptr(flags), _ptr(counts))
    return flags, int(counts[0]), int(counts[1])


# ---- remeshing (src/remesh/) -------------------------------------------------------------------


class _RemeshOptionsC(ctypes.Structure):
    _fields_ = [
        ("split", _D),
        ("collapse", _D),
        ("rounds", _I32),
        ("rho", _D),
        ("flip_cap_per_vertex", _I64),
        ("order", _I32),
        ("split_on", _I32),
        ("collapse_on", _I32),
        ("flip_on", _I32),
        ("smooth_on", _I32),
        ("max_growth", _D),
        ("guard_gap", _D),
    ]


class _RemeshReportC(ctypes.Structure):
    _fields_ = [
        ("splits", _I64),
        ("collapses", _I64),
        ("flips", _I64),
        ("splits_deferred", _I64),
        ("refused_splits", _I64 * 11),
        ("refused_collapses", _I64 * 11),
        ("refused_flips", _I64 * 11),
        ("smoothing_moves", _I64),
        ("smoothing_halved", _I64),
        ("smoothing_refused", _I64),
        ("smoothing_skipped", _I64),
        ("flip_cap_hits", _I64),
        ("n_vertices", _I64),
        ("n_faces", _I64),
        ("smoothing_iterations", _I64),
        ("seconds", _D),
    ]


# src/remesh/mesh.rs:75 ``RefusalKind::ALL``
REFUSAL_KINDS = ("NoSuchEdge", "BothPinned", "LinkCondition", "Foldover", "RepeatedFace", "NonManifoldResult",
                 "BoundaryEdge", "NotDelaunay", "ExistingEdge", "Guard", "DegenerateFace")


@dataclasses.dataclass(frozen=True)
class EnabledOps:
    """``EnabledOps`` (src/remesh/mod.rs:80)."""

    split: bool = True
    collapse: bool = True
    flip: bool = True
    smooth: bool = True


@dataclasses.dataclass(frozen=True)
class RemeshOptions:
    """``RemeshOptions`` (src/remesh/mod.rs:96): split / collapse thresholds as multiples of L0,
    flip and smooth rounds, smoothing step ρ, flip cap per vertex, order "flip_then_smooth" (the
    paper, default) or "smooth_then_flip", the parts that run, the growth cap (``math.inf``
    disables it).  ``guard_collapses`` is the session's choice (the CCD gap rule on collapses, with
    ``CcdOptions().gap``); ``remesh`` takes the guard explicitly."""

    split: float = 1.5
    collapse: float = 0.5
    rounds: int = 5
    rho: float = 0.5
    flip_cap_per_vertex: int = 100
    order: str = "flip_then_smooth"
    enabled_ops: EnabledOps = EnabledOps()
    max_growth: float = 2.0
    guard_collapses: bool = True

    def _c(self, guard_gap: float) -> _RemeshOptionsC:
        if self.order not in ("flip_then_smooth", "smooth_then_flip"):
            raise RsError(1, "remesh", f"order: {self.order!r}: flip_then_smooth or smooth_then_flip")
        e = self.enabled_ops
        return _RemeshOptionsC(self.split, self.collapse, self.rounds, self.rho, self.flip_cap_per_vertex,
                               0 if self.order == "flip_then_smooth" else 1, int(e.split), int(e.collapse),
                               int(e.flip), int(e.smooth), self.max_growth, guard_gap)


@dataclasses.dataclass(frozen=True)
class VertexMap:
    """``VertexMap`` (src/remesh/transport.rs:15): new vertex → old vertices and weights (CSR)."""

    offsets: np.ndarray
    sources: np.ndarray
    weights: np.ndarray
    n_old: int

    def entry(self, v: int):
        lo, hi = int(self.offsets[v]), int(self.offsets[v + 1])
        return self.sources[lo:hi], self.weights[lo:hi]

    def transport(self, old) -> np.ndarray:
        """new[v] = Σ w_i old[i] for per-vertex data (transport.rs:83)."""
        a = np.asarray(old, dtype=np.float64)
        if a.shape[0] != self.n_old:
            raise RsError(1, "transport", f"{a.shape[0]} values for {self.n_old} old vertices")
        counts = np.diff(self.offsets.astype(np.int64))
        rows = np.repeat(np.arange(len(counts)), counts)
        w = self.weights.reshape((-1,) + (1,) * (a.ndim - 1))
        out = np.zeros((len(counts),) + a.shape[1:])
        np.add.at(out, rows, w * a[self.sources])
        return out


@dataclasses.dataclass(frozen=True)
class Remeshed:
    """``Remeshed`` (src/remesh/mesh.rs:210): faces, positions, the vertex map, the face parents
    and the new ids of the input pins."""

    faces: np.ndarray
    positions: np.ndarray
    vertex_map: VertexMap
    face_parents: np.ndarray
    pinned: np.ndarray


@dataclasses.dataclass(frozen=True)
class RemeshReport:
    """``RemeshReport`` (src/remesh/mod.rs:187); refusals as {kind: count}."""

    splits: int
    collapses: int
    flips: int
    refused_splits: dict
    splits_deferred: int
    refused_collapses: dict
    refused_flips: dict
    smoothing_moves: int
    smoothing_halved: int
    smoothing_refused: int
    smoothing_skipped: int
    flip_cap_hits: int
    n_vertices: int
    n_faces: int
    seconds: float
    smoothing_iterations: int = 0


def _read_remeshed(h, n_old: int) -> Remeshed:
    L = lib()
    try:
        n = np.zeros(4, dtype=np.int64)
        _call(L.ddgrs_remeshed_sizes, h, _ptr(n))
        nv, nf, nm, npin = (int(x) for x in n)
        pos = np.zeros((nv, 3))
        faces = np.zeros((nf, 3), dtype=np.uint32)
        parents = np.zeros(nf, dtype=np.uint32)
        off = np.zeros(nv + 1, dtype=np.uint32)
        src = np.zeros(nm, dtype=np.uint32)
        w = np.zeros(nm)
        pins = np.zeros(npin, dtype=np.uint32)
        _call(L.ddgrs_remeshed_read, h, _ptr(pos), _ptr(faces), _ptr(parents), _ptr(off), _ptr(src), _ptr(w),
              _ptr(pins))
        return Remeshed(faces, pos, VertexMap(off, src, w, n_old), parents, pins)
    finally:
        L.ddgrs_remeshed_destroy(h)


def remesh(ctx: "Context", faces, positions, l0: float, options: RemeshOptions = RemeshOptions(), pinned=(),
           guard_gap: Optional[float] = None, band=None):
    """``remesh`` (src/remesh/mod.rs:235) on the device: one pass of splits, collapses, Delaunay
    flips and tangential smoothing.  ``guard_gap``: the absolute gap of the collapse guard (the
    session's ``CcdCollapseGuard``, src/ccd.rs:1448; None: no guard); ``band``: its exempt vertices
    (``intersection_band``).  Returns (Remeshed, RemeshReport)."""
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    p = _f64(positions).reshape(-1, 3)
    pins = np.ascontiguousarray(pinned, dtype=np.uint32).reshape(-1)
    b = None if band is None else np.ascontiguousarray(band, dtype=np.uint8).reshape(-1)
    if b is not None and b.shape[0] != p.shape[0]:
        raise RsError(1, "remesh", f"band: {b.shape[0]} flags for {p.shape[0]} vertices")
    o = options._c(0.0 if guard_gap is None else float(guard_gap))
    h = _P()
    r = _RemeshReportC()
    _call(lib().ddgrs_remesh, ctx.handle, p.shape[0], f.shape[0], _ptr(f), _ptr(p), pins.size,
          _ptr(pins) if pins.size else None, float(l0), ctypes.byref(o), None if b is None else _ptr(b),
          ctypes.byref(h), ctypes.byref(r))
    out = _read_remeshed(h, p.shape[0])
    kinds = lambda arr: {k: int(v) for k, v in zip(REFUSAL_KINDS, arr)}
    rep = RemeshReport(int(r.splits), int(r.collapses), int(r.flips), kinds(r.refused_splits), int(r.splits_deferred),
                       kinds(r.refused_collapses), kinds(r.refused_flips), int(r.smoothing_moves),
                       int(r.smoothing_halved), int(r.smoothing_refused), int(r.smoothing_skipped),
                       int(r.flip_cap_hits), int(r.n_vertices), int(r.n_faces), float(r.seconds),
                       int(r.smoothing_iterations))
    return out, rep


def remesh_op(ctx: "Context", faces, positions, op: str, a: int, b: int, pinned=(), delaunay: bool = False,
              foldover: bool = False):
    """One ``EditMesh`` operation on the device (src/remesh/mesh.rs:592 split, :679 collapse, :888
    flip) and ``finish``: (refusal kind or None, reordered, Remeshed)."""
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    p = _f64(positions).reshape(-1, 3)
    pins = np.ascontiguousarray(pinned, dtype=np.uint32).reshape(-1)
    ref, reo, h = _I32(), _I32(), _P()
    _call(lib().ddgrs_remesh_op, ctx.handle, p.shape[0], f.shape[0], _ptr(f), _ptr(p), pins.size,
          _ptr(pins) if pins.size else None, {"split": 0, "collapse": 1, "flip": 2}[op], a, b, int(delaunay),
          int(foldover), ctypes.byref(ref), ctypes.byref(reo), ctypes.byref(h))
    out = _read_remeshed(h, p.shape[0])
    return (None if ref.value == 0 else REFUSAL_KINDS[ref.value - 1]), bool(reo.value), out


# ---- energy (src/energy/mod.rs) ----------------------------------------------------------------


@dataclasses.dataclass(frozen=True)
class Exponents:
    """``Exponents`` (src/kernel.rs:26); validated by the library."""

    alpha: float = 6.0
    beta: float = 12.0


@dataclasses.dataclass(frozen=True)
class BhOptions:
    """``BhOptions`` (src/energy/mod.rs:62): θ = 0.5, leaf size 8 by default."""

    theta: float = 0.5
    leaf_size: int = 8


class EnergyEval:
    """``EnergyEval::{Exact, BarnesHut}`` (src/energy/mod.rs:106)."""

    Exact = "exact"

    @staticmethod
    def BarnesHut(opts: BhOptions = BhOptions()):
        return ("barnes_hut", opts)


@dataclasses.dataclass
class EnergyReport:
    """``EnergyReport`` (src/energy/mod.rs:130) without penalties."""

    energy: float
    differential: np.ndarray
    pairs_exact: int
    nodes_lumped: int


def exact_energy_and_differential(surface: Surface, exponents: Exponents = Exponents(), partials: bool = False):
    """``exact::energy_and_differential`` (src/energy/exact.rs:146): (E, dE (nv×3)[, partials (nf×7)])."""
    e = _D()
    d = np.zeros((surface.nv, 3))
    fp = np.zeros((surface.nf, 7)) if partials else None
    _call(lib().ddgrs_exact_energy_and_differential, surface.handle, exponents.alpha, exponents.beta, ctypes.byref(e),
          _ptr(d), _ptr(fp) if partials else None)
    return (e.value, d, fp) if partials else (e.value, d)


def exact_energy_and_differential_noalloc(surface: Surface, exponents: Exponents = Exponents(), partials: bool = False):
    """As ``exact_energy_and_differential`` through the entry point that allocates nothing (its
    scratch is a fixed per-device array: at most 131,072 faces and vertices; overlapping calls on
    one device race)."""
    e = _D()
    d = np.zeros((surface.nv, 3))
    fp = np.zeros((surface.nf, 7)) if partials else None
    _call(lib().ddgrs_exact_energy_and_differential_noalloc, surface.handle, exponents.alpha, exponents.beta,
          ctypes.byref(e), _ptr(d), _ptr(fp) if partials else None)
    return (e.value, d, fp) if partials else (e.value, d)


class BhTree:
    """The Barnes–Hut evaluator of one step (``BhTree``, src/energy/barnes_hut.rs:114): face BVH and
    frozen lists built at the surface's current positions; ``refit`` moves it for line-search
    trials."""

    def __init__(self, surface: Surface, opts: BhOptions = BhOptions()):
        self.surface = surface
        self.opts = opts
        h = _P()
        _call(lib().ddgrs_bh_build, surface.handle, opts.theta, opts.leaf_size, ctypes.byref(h))
        self._h = h
        _register(self, 1)

    def close(self):
        if getattr(self, "_h", None):
            lib().ddgrs_bh_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def refit(self, positions=None):
        if positions is None:
            _call(lib().ddgrs_bh_refit, self._h, None)
        else:
            p = _f64(positions).reshape(-1, 3)
            if p.shape[0] != self.surface.nv:
                raise RsError(1, "bh_refit", f"{p.shape[0]} positions for {self.surface.nv} vertices")
            _call(lib().ddgrs_bh_refit, self._h, _ptr(p))

    def energy(self, exponents: Exponents = Exponents()) -> float:
        e = _D()
        _call(lib().ddgrs_bh_energy, self._h, exponents.alpha, exponents.beta, ctypes.byref(e))
        return e.value

    def energy_and_differential(self, exponents: Exponents = Exponents(), partials: bool = False):
        e = _D()
        d = np.zeros((self.surface.nv, 3))
        fp = np.zeros((self.surface.nf, 7)) if partials else None
        _call(lib().ddgrs_bh_energy_and_differential, self._h, exponents.alpha, exponents.beta, ctypes.byref(e), _ptr(d),
              _ptr(fp) if partials else None)
        return (e.value, d, fp) if partials else (e.value, d)

    def counts(self) -> dict:
        c = np.zeros(6, dtype=np.int64)
        _call(lib().ddgrs_bh_counts, self._h, _ptr(c))
        keys = ["nodes", "leaves", "lumped", "exact_blocks", "pairs_exact", "partner_slots"]
        return dict(zip(keys, (int(x) for x in c)))

    def tree(self) -> dict:
        c = self.counts()
        nn = c["nodes"]
        order = np.zeros(self.surface.nf, dtype=np.uint32)
        start, end, right = (np.zeros(nn, dtype=np.uint32) for _ in range(3))
        cl = np.zeros((nn, 17))
        _call(lib().ddgrs_bh_tree, self._h, _ptr(order), _ptr(start), _ptr(end), _ptr(right), _ptr(cl))
        return {"order": order, "start": start, "end": end, "right": right, "clusters": cl}

    def lists(self) -> dict:
        c = self.counts()
        nl = c["leaves"]
        lo = np.zeros(nl + 1, dtype=np.uint32)
        lm = np.zeros(c["lumped"], dtype=np.uint32)
        eo = np.zeros(nl + 1, dtype=np.uint32)
        ex = np.zeros(c["exact_blocks"], dtype=np.uint32)
        _call(lib().ddgrs_bh_lists, self._h, _ptr(lo), _ptr(lm), _ptr(eo), _ptr(ex))
        return {"lumped_off": lo, "lumped": lm, "exact_off": eo, "exact": ex}


@dataclasses.dataclass(frozen=True)
class ClusterTreeOptions:
    """``ClusterTreeOptions`` (src/metric/hierarchical.rs:65): χ = 0.5, leaf 8, leaf_pairs 64."""

    chi: float = 0.5
    leaf_size: int = 8
    leaf_pairs: int = 64


class _Handle:
    _destroy = None
    _rank = 1  # teardown order (see close_all)

    def __setattr__(self, name, value):
        object.__setattr__(self, name, value)
        if name == "_h" and value is not None:
            _register(self, self._rank)

    def close(self):
        if getattr(self, "_h", None):
            getattr(lib(), self._destroy)(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _batch(fn, h, n, x, *extra):
    xs = _f64(x)
    one = xs.ndim == 1
    xs = np.ascontiguousarray(xs.reshape(-1, n))
    ys = np.zeros_like(xs)
    _call(fn, h, *extra, xs.shape[0], _ptr(xs), _ptr(ys))
    return ys[0] if one else ys


def _block_lists(fn, h) -> dict:
    """The tree and block lists of a block cluster tree handle (``ddgrs_*_blocks``)."""
    c = np.zeros(4, dtype=np.int64)
    _call(fn, h, _ptr(c), *([None] * 8))
    n, nn, nfar, nnear = (int(v) for v in c)
    out = {k: np.zeros(m, dtype=np.uint32) for k, m in (("order", n), ("start", nn), ("end", nn), ("right", nn),
                                                        ("far_ptr", nn + 1), ("far_col", nfar),
                                                        ("near_ptr", nn + 1), ("near_col", nnear))}
    _call(fn, h, _ptr(c), *[_ptr(out[k]) for k in ("order", "start", "end", "right", "far_ptr", "far_col",
                                                      "near_ptr", "near_col")])
    return out


class FaceMetricTree(_Handle):
    """B̃ and B̃⁰ of one configuration (``FaceMetricTree``, src/metric/hierarchical.rs:283)."""

    _destroy = "ddgrs_metric_destroy"

    def __init__(self, surface: Surface, exponents: Exponents = Exponents(), options: ClusterTreeOptions = ClusterTreeOptions()):
        self.surface = surface
        h = _P()
        _call(lib().ddgrs_metric_build, surface.handle, exponents.alpha, exponents.beta, options.chi, options.leaf_size,
              options.leaf_pairs, ctypes.byref(h))
        self._h = h

    def apply_batch(self, x, w_hi: float = 1.0, w_lo: float = 1.0):
        """y = (w_hi B̃ + w_lo B̃⁰) x for x of shape (nb, 3 nv) or (3 nv,)."""
        return _batch(lib().ddgrs_metric_apply_batch, self._h, 3 * self.surface.nv, x, w_hi, w_lo)

    def stats(self) -> dict:
        o = np.zeros(7, dtype=np.int64)
        _call(lib().ddgrs_metric_stats, self._h, _ptr(o))
        keys = ["admissible_blocks", "inadmissible_blocks", "admissible_pairs", "near_entries", "nodes", "depth", "near_total"]
        return dict(zip(keys, (int(v) for v in o)))

    def blocks(self) -> dict:
        """The face tree (order, start, end, right) and the admissible (far) and inadmissible (near)
        block lists per row node in creation order (src/spatial/block_tree.rs:129-209)."""
        return _block_lists(lib().ddgrs_metric_blocks, self._h)


class VertexFractionalTree(_Handle):
    """L̃^V_τ over vertex pairs (``VertexFractionalTree``, src/metric/hierarchical.rs:1003)."""

    _destroy = "ddgrs_vfrac_destroy"

    def __init__(self, surface: Surface, tau: float, options: ClusterTreeOptions = ClusterTreeOptions()):
        self.surface = surface
        h = _P()
        _call(lib().ddgrs_vfrac_build, surface.handle, tau, options.chi, options.leaf_size, options.leaf_pairs, ctypes.byref(h))
        self._h = h

    def apply_batch(self, x):
        return _batch(lib().ddgrs_vfrac_apply_batch, self._h, 3 * self.surface.nv, x)

    def blocks(self) -> dict:
        """The vertex tree and block lists, as ``FaceMetricTree.blocks`` (hierarchical.rs:1017)."""
        return _block_lists(lib().ddgrs_vfrac_blocks, self._h)


class _SaddleOptionsC(ctypes.Structure):
    _fields_ = [
        ("chi", _D),
        ("leaf_size", _I64),
        ("leaf_pairs", _I64),
        ("w_tp", _D),
        ("c_lo", _D),
        ("eps_rel", _D),
        ("form", _I32),
        ("scale", _I32),
        ("refine", _I32),
    ]


class _GmresReportC(ctypes.Structure):
    _fields_ = [
        ("iterations", _I64),
        ("cycles", _I64),
        ("operator_applications", _I64),
        ("preconditioner_applications", _I64),
        ("converged", _I32),
        ("extended", _I32),
        ("relative_residual", _D),
    ]


class _StepOptionsC(ctypes.Structure):
    _fields_ = [
        ("alpha", _D),
        ("beta", _D),
        ("energy_exact", _I32),
        ("theta", _D),
        ("saddle", _SaddleOptionsC),
        ("restart", _I64),
        ("tolerance", _D),
        ("max_iterations", _I64),
        ("armijo_c", _D),
        ("tau_min_rel", _D),
        ("max_halvings", _I64),
        ("start_kind", _I32),
        ("start_value", _D),
        ("robust", _I32),
        ("retry_max_iterations", _I64),
        ("dense_max_vertices", _I64),
        ("diverged_residual", _D),
        ("rows", _I32 * 2),
        ("row_targets", _D * 2),
        ("newton_fixed", _I32),
        ("newton_count", _I32),
        ("ccd", _I32),
        ("ccd_options", _CcdOptionsC),
    ]


class _StepReportC(ctypes.Structure):
    _fields_ = [
        ("energy_before", _D),
        ("energy_trial", _D),
        ("gradient_norm", _D),
        ("sobolev_norm_sq", _D),
        ("tau", _D),
        ("tau0", _D),
        ("accepted", _I32),
        ("trials", _I32),
        ("gmres", _GmresReportC),
        ("pairs_exact", _I64),
        ("nodes_lumped", _I64),
        ("seconds_trees", _D),
        ("seconds_energy", _D),
        ("seconds_metric", _D),
        ("seconds_precond", _D),
        ("seconds_solve", _D),
        ("seconds_line_search", _D),
        ("seconds_projection", _D),
        ("seconds_total", _D),
        ("learned_mode", _I32),
        ("learned_fallback", _I32),
        ("learned_iterations", _I64),
        ("seconds_learned", _D),
        ("energy_after", _D),
        ("direction_norm", _D),
        ("direction_max", _D),
        ("halvings", _I64),
        ("fallback_stage", _I32),
        ("newton_rounds", _I32),
        ("min_pivot", _D),
        ("has_ccd", _I32),
        ("tau_ccd", _D),
        ("ccd", _CcdReportC),
        ("seconds_ccd", _D),
    ]


@dataclasses.dataclass(frozen=True)
class SaddleOptions:
    """The step's hierarchical solver options (src/solve/iterative.rs:350 ``HierarchicalOptions``,
    src/solve/precond.rs:466 ``PreconditionerOptions``; metric weights of ``SobolevWeights``)."""

    chi: float = 0.5
    leaf_size: int = 8
    leaf_pairs: int = 64
    w_tp: float = 1.0
    c_lo: float = 1.0
    eps_rel: float = 1e-10
    # PreconditionerOptions::form ("augmented" or "congruence"), ::scale of the fractional factor
    # ("symbol" or "unit") and ::refine (one step of iterative refinement per L' solve)
    form: str = "augmented"
    scale: str = "symbol"
    refine: bool = False

    def _c(self) -> _SaddleOptionsC:
        forms, scales = {"augmented": 0, "congruence": 1}, {"symbol": 0, "unit": 1}
        if self.form not in forms:
            raise RsError(1, "options", f"form {self.form!r}: 'augmented' or 'congruence'")
        if self.scale not in scales:
            raise RsError(1, "options", f"scale {self.scale!r}: 'symbol' or 'unit'")
        return _SaddleOptionsC(self.chi, self.leaf_size, self.leaf_pairs, self.w_tp, self.c_lo, self.eps_rel,
                               forms[self.form], scales[self.scale], int(bool(self.refine)))


@dataclasses.dataclass(frozen=True)
class GmresOptions:
    """``GmresOptions`` (src/solve/gmres.rs:138): restart 50, relative tolerance 1e-6, 200 iterations."""

    restart: int = 50
    tolerance: float = 1e-6
    max_iterations: int = 200


@dataclasses.dataclass(frozen=True)
class GmresReport:
    """``GmresReport`` (src/solve/gmres.rs:200)."""

    iterations: int
    cycles: int
    operator_applications: int
    preconditioner_applications: int
    converged: bool
    relative_residual: float
    extended: bool = False  # the products switched to the extended near field (FP32 residual floor reached)
    cycle_residuals: tuple = ()  # the explicit relative residual at every cycle start and at the end
    history: tuple = ()  # the Arnoldi estimate after every iteration


def _gmres_history(fn, h) -> tuple:
    """(cycle_residuals, history) of the last solve through ``fn`` (a ``*_gmres_history`` entry)."""
    counts = np.zeros(2, dtype=np.int64)
    _call(fn, h, None, 0, None, 0, _ptr(counts))
    c = np.zeros(int(counts[0]))
    hist = np.zeros(int(counts[1]))
    _call(fn, h, _ptr(c), c.size, _ptr(hist), hist.size, _ptr(counts))
    return tuple(float(v) for v in c), tuple(float(v) for v in hist)


def _gmres_report(c: _GmresReportC, history: tuple = ((), ())) -> GmresReport:
    return GmresReport(int(c.iterations), int(c.cycles), int(c.operator_applications), int(c.preconditioner_applications),
                       bool(c.converged), float(c.relative_residual), bool(c.extended), *history)


class Saddle(_Handle):
    """The step's saddle solver: A′ = [[w(B̃ + c_lo B̃⁰), C_Qᵀ], [C_Q, 0]] with the global barycenter
    rows, its augmented preconditioner and GMRES (``hierarchical_saddle``, src/solve/iterative.rs:455;
    ``GmresSaddle``, src/solve/gmres.rs:717).  Vectors have 3 nv + 3 entries."""

    _destroy = "ddgrs_saddle_destroy"

    def __init__(self, surface: Surface, exponents: Exponents = Exponents(), options: SaddleOptions = SaddleOptions()):
        self.surface = surface
        self.n = 3 * surface.nv + 3
        h = _P()
        o = options._c()
        _call(lib().ddgrs_saddle_build, surface.handle, exponents.alpha, exponents.beta, ctypes.byref(o), ctypes.byref(h))
        self._h = h

    def info(self) -> dict:
        o = np.zeros(2)
        _call(lib().ddgrs_saddle_info, self._h, _ptr(o))
        return {"eps": float(o[0]), "fractional_scale": float(o[1])}

    def apply_batch(self, x):
        return _batch(lib().ddgrs_saddle_apply_batch, self._h, self.n, x)

    def precondition(self, r):
        return _batch(lib().ddgrs_saddle_precond_batch, self._h, self.n, r)

    def lprime_solve(self, r):
        return _batch(lib().ddgrs_saddle_lprime_batch, self._h, self.n, r)

    def solve(self, b, options: GmresOptions = GmresOptions()):
        """x = A′⁻¹ b and the ``GmresReport``."""
        bb = _f64(b, (self.n,))
        x = np.zeros(self.n)
        rep = _GmresReportC()
        _call(lib().ddgrs_saddle_solve, self._h, _ptr(bb), _ptr(x), options.restart, options.tolerance,
              options.max_iterations, ctypes.byref(rep))
        return x, _gmres_report(rep, _gmres_history(lib().ddgrs_saddle_gmres_history, self._h))

    def solve_many(self, bs, options: GmresOptions = GmresOptions()):
        """``GmresSaddle::solve_many`` (src/solve/gmres.rs:801): x_i = A′⁻¹ b_i for the rows of bs
        (nb × (3 nv + 3)) and one ``GmresReport`` per row (histories left empty; the history
        entry point holds the last solve's)."""
        bb = np.ascontiguousarray(bs, dtype=np.float64)
        if bb.ndim != 2 or bb.shape[1] != self.n:
            raise RsError(1, "saddle_solve_many", f"right-hand sides of shape {bb.shape}, want (nb, {self.n})")
        nb = bb.shape[0]
        xs = np.zeros((nb, self.n))
        reps = (_GmresReportC * max(nb, 1))()
        _call(lib().ddgrs_saddle_solve_many, self._h, nb, _ptr(bb), _ptr(xs), options.restart, options.tolerance,
              options.max_iterations, reps)
        return xs, [_gmres_report(reps[i]) for i in range(nb)]


def schur(c, z, y, phi):
    """The step's Schur complement on the host (``SchurComplement``, src/solve/schur.rs:63): from the
    rows c (k × nd) and the columns z = A′⁻¹ C_Nᵀ (k × nd), (x = y − Z N⁻¹ C_N y, h = Z N⁻¹ Φ_N,
    S = −C_N Z, the smallest kept pivot)."""
    c = np.ascontiguousarray(c, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    k, nd = c.shape
    y = _f64(y, (nd,))
    phi = _f64(phi, (k,))
    x, h, S, mp = np.zeros(nd), np.zeros(nd), np.zeros((k, k)), ctypes.c_double()
    _call(lib().ddgrs_schur, k, nd, _ptr(c), _ptr(z), _ptr(y), _ptr(phi), _ptr(x), _ptr(h), _ptr(S), ctypes.byref(mp))
    return x, h, S, float(mp.value)


class StartRule:
    """The line search's first step (src/step/options.rs ``StartRule``; session.rs:1604): values
    are ``(kind, value)`` pairs for ``StepOptions.start``."""

    GROW_PREVIOUS, ABSOLUTE, MAX_DISPLACEMENT, REFERENCE = 0, 1, 2, 3

    @staticmethod
    def grow_previous(lengths: float = 1.0):
        """``GrowPrevious{factor: 2, first: MaxDisplacement{lengths}}`` (the default)."""
        return (StartRule.GROW_PREVIOUS, float(lengths))

    @staticmethod
    def absolute(tau: float):
        return (StartRule.ABSOLUTE, float(tau))

    @staticmethod
    def max_displacement(lengths: float):
        return (StartRule.MAX_DISPLACEMENT, float(lengths))

    @staticmethod
    def reference():
        """``Reference``: τ₀ = 2 / ‖x‖."""
        return (StartRule.REFERENCE, 0.0)


class TrialOutcome:
    """``TrialOutcome`` (src/step/session.rs)."""

    ACCEPTED, ARMIJO_FAILED, NON_FINITE = 0, 1, 2


@dataclasses.dataclass(frozen=True)
class StepOptions:
    """A step of the Hs flow (src/step/session.rs ``FlowOptions`` with remeshing off, the global
    barycenter constraint; ``start`` defaults to ``StartRule::GrowPrevious{2,
    MaxDisplacement{1}}``, ``max_halvings`` < 0 is no cap; ``ccd`` None is
    ``LineSearchOptions::ccd`` None, else the CCD cap of τ₀ with those options)."""

    exponents: Exponents = Exponents()
    energy_exact: bool = False
    theta: float = 0.5
    saddle: SaddleOptions = SaddleOptions()
    gmres: GmresOptions = GmresOptions()
    armijo_c: float = 1e-2
    tau_min_rel: float = 1e-12
    max_halvings: int = -1
    start: tuple = (0, 1.0)
    # the recovery chain of a diverged solve (src/step/direction.rs:259 RobustHierarchical)
    robust: bool = True
    retry_max_iterations: int = 0  # 0: max(4 max_iterations, 800)
    dense_max_vertices: int = 6000
    diverged_residual: float = 1e-1  # an unconverged solve above it has diverged (the reference's 0.1)
    # nonlinear rows after the barycenter's: up to two of ("area", target) and ("volume", target),
    # target None for the value at creation (src/constraints/mod.rs Constraint::Area, ::Volume);
    # Newton rounds of the projection: ("tol", n) = UntilTolerance{max: n}, ("fixed", n)
    constraints: tuple = ()
    newton: tuple = ("tol", 5)
    ccd: Optional[CcdOptions] = None

    _ROW_KINDS = {"area": 1, "volume": 2}

    def _c(self) -> _StepOptionsC:
        if len(self.constraints) > 2:
            raise RsError(1, "options", "constraints: at most the area and the volume rows")
        rows = [self._ROW_KINDS[k] for k, _ in self.constraints] + [0] * (2 - len(self.constraints))
        targets = [float("nan") if t is None else float(t) for _, t in self.constraints]
        targets += [float("nan")] * (2 - len(targets))
        if self.newton[0] not in ("tol", "fixed"):
            raise RsError(1, "options", f"newton = {self.newton!r}: ('tol', n) or ('fixed', n)")
        return _StepOptionsC(self.exponents.alpha, self.exponents.beta, 1 if self.energy_exact else 0, self.theta,
                             self.saddle._c(), self.gmres.restart, self.gmres.tolerance, self.gmres.max_iterations,
                             self.armijo_c, self.tau_min_rel, self.max_halvings, int(self.start[0]),
                             float(self.start[1]), 1 if self.robust else 0, self.retry_max_iterations,
                             self.dense_max_vertices, self.diverged_residual, (_I32 * 2)(*rows),
                             (_D * 2)(*targets), 1 if self.newton[0] == "fixed" else 0, int(self.newton[1]),
                             0 if self.ccd is None else 1, (self.ccd or CcdOptions())._c())


@dataclasses.dataclass(frozen=True)
class StepReport:
    """The report of one step (src/step/session.rs:146 ``StepReport``, the fields of this path)."""

    energy_before: float
    energy_trial: float
    gradient_norm: float
    sobolev_norm_sq: float
    tau: float
    tau0: float
    accepted: bool
    trials: int
    gmres: GmresReport
    pairs_exact: int
    nodes_lumped: int
    seconds: dict
    learned_mode: int = 0
    learned_fallback: int = 0  # 1: stagnated, plain used; 2: Q off after consecutive misses
    learned_iterations: int = 0
    energy_after: float = 0.0  # F at the projected positions
    direction_norm: float = 0.0
    direction_max: float = 0.0
    halvings: int = 0
    trial_list: tuple = ()  # (τ, F, TrialOutcome) per trial, in order
    fallback_stage: int = 0  # 1: the GMRES retry solved the step, 2: the dense LU did
    fallbacks: tuple = ()  # the escalations as text (LinearReport::fallbacks)
    newton_rounds: int = 0  # Newton rounds kept by the projection (nonlinear rows)
    min_pivot: float = 1.0  # the smallest kept equilibrated Schur pivot
    tau_ccd: Optional[float] = None  # the CCD's τ_max (tau0 is after its cap)
    ccd: Optional["CcdReport"] = None


class Session(_Handle):
    """A descent session on a device-resident surface (``Session``, src/step/session.rs:280): ``step()``
    moves the surface (``Session::step``, :724)."""

    _destroy = "ddgrs_session_destroy"
    _rank = 0

    def __init__(self, surface: Surface, options: StepOptions = StepOptions()):
        self.surface = surface
        self.options = options
        h = _P()
        o = options._c()
        _call(lib().ddgrs_session_create, surface.handle, ctypes.byref(o), ctypes.byref(h))
        self._h = h

    def step(self) -> StepReport:
        r = _StepReportC()
        _call(lib().ddgrs_session_step, self._h, ctypes.byref(r))
        sec = {k[len("seconds_"):]: float(getattr(r, k)) for k, _ in _StepReportC._fields_ if k.startswith("seconds_")}
        return StepReport(float(r.energy_before), float(r.energy_trial), float(r.gradient_norm), float(r.sobolev_norm_sq),
                          float(r.tau), float(r.tau0), bool(r.accepted), int(r.trials),
                          _gmres_report(r.gmres, _gmres_history(lib().ddgrs_session_gmres_history, self._h)),
                          int(r.pairs_exact), int(r.nodes_lumped), sec, int(r.learned_mode), int(r.learned_fallback),
                          int(r.learned_iterations), float(r.energy_after), float(r.direction_norm),
                          float(r.direction_max), int(r.halvings), self.trials(), int(r.fallback_stage),
                          self.fallbacks(), int(r.newton_rounds), float(r.min_pivot),
                          float(r.tau_ccd) if r.has_ccd else None, _ccd_report(r.ccd) if r.has_ccd else None)

    def constraints(self) -> tuple:
        """(Φ before the projection, Φ after it) of the last step: the barycenter's x, y, z, then
        the nonlinear rows; empty without nonlinear rows."""
        n = _I64()
        _call(lib().ddgrs_session_constraints, self._h, None, None, 0, ctypes.byref(n))
        a, b = np.zeros(int(n.value)), np.zeros(int(n.value))
        if n.value:
            _call(lib().ddgrs_session_constraints, self._h, _ptr(a), _ptr(b), int(n.value), ctypes.byref(n))
        return tuple(a.tolist()), tuple(b.tolist())

    def fallbacks(self) -> tuple:
        """The last step's escalations of the recovery chain, as text."""
        n = _I64()
        buf = ctypes.create_string_buffer(4096)
        _call(lib().ddgrs_session_fallbacks, self._h, buf, 4096, ctypes.byref(n))
        return tuple(x for x in buf.value.decode().split("\n") if x)

    def trials(self) -> tuple:
        """The last step's trials as (τ, F, ``TrialOutcome``) in order (``StepReport::trials``)."""
        n = _I64()
        _call(lib().ddgrs_session_trials, self._h, 0, None, None, None, ctypes.byref(n))
        k = int(n.value)
        tau, f, o = np.zeros(k), np.zeros(k), np.zeros(k, dtype=np.int32)
        if k:
            _call(lib().ddgrs_session_trials, self._h, k, _ptr(tau), _ptr(f), _ptr(o), ctypes.byref(n))
        return tuple((float(a), float(b), int(c)) for a, b, c in zip(tau, f, o))

    def set_learned(self, learned: "Learned | None", mode: int):
        """Attach a learned preconditioner: mode 1 records a training pair per step (plain solve),
        2 solves with (I + Q) P⁻¹ (the plain solve on a miss), 0 detaches."""
        _call(lib().ddgrs_session_set_learned, self._h, learned._h if learned is not None else None, mode)
        self._learned = learned  # keeps it alive while attached

    def state(self) -> dict:
        o = np.zeros(5)
        _call(lib().ddgrs_session_state, self._h, _ptr(o))
        return {"mean_edge_length0": float(o[0]), "last_tau": float(o[1]), "barycenter_target": o[2:5].copy()}

    def set_state(self, mean_edge_length0: float, last_tau: float, barycenter_target) -> None:
        """Continues a run on a new surface (after a remeshing): ℓ̄₀, the last accepted τ and the
        barycenter target of the session it replaces."""
        o = np.array([mean_edge_length0, last_tau, *np.asarray(barycenter_target, dtype=np.float64).reshape(3)])
        _call(lib().ddgrs_session_set_state, self._h, _ptr(o))

    def row_targets(self) -> tuple:
        """The area / volume rows' targets (NaN for no row), as given or at creation."""
        o = np.zeros(2)
        _call(lib().ddgrs_session_row_targets, self._h, _ptr(o))
        return float(o[0]), float(o[1])

    def positions(self) -> np.ndarray:
        return self.surface.positions()


class Learned(_Handle):
    """The learned right preconditioner of a surface's mesh (no counterpart in the reference): a
    mesh network giving per-vertex and per-edge 3×3 blocks, Q, added to the plain preconditioner
    (z = (I + Q) P⁻¹ r).  Training, the forward pass and Q x run on the device."""

    _destroy = "ddgrs_learned_destroy"

    def __init__(self, surface: Surface, seed: int = 0, max_pairs: int = 64):
        self.surface = surface
        h = _P()
        _call(lib().ddgrs_learned_create, surface.handle, seed, max_pairs, ctypes.byref(h))
        self._h = h

    def info(self) -> dict:
        o = np.zeros(4, dtype=np.int64)
        sc = ctypes.c_double()
        _call(lib().ddgrs_learned_info, self._h, _ptr(o), ctypes.byref(sc))
        return {"parameters": int(o[0]), "pairs": int(o[1]), "adam_steps": int(o[2]), "features": int(o[3]),
                "scale": float(sc.value)}

    def params(self) -> np.ndarray:
        p = np.zeros(self.info()["parameters"], dtype=np.float32)
        _call(lib().ddgrs_learned_params, self._h, _ptr(p), None)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (self.info()["parameters"],):
            raise ValueError("wrong parameter count")
        _call(lib().ddgrs_learned_params, self._h, None, _ptr(p))

    def set_scale(self, s: float):
        _call(lib().ddgrs_learned_set_scale, self._h, float(s))

    def set_krylov_pairs(self, k: int):
        """Krylov pairs a recording step adds after its own (0 to 8, default 3)."""
        _call(lib().ddgrs_learned_set_krylov_pairs, self._h, int(k))

    def add_pair(self, positions, r, pr, y):
        n = 3 * self.surface.nv + 3
        a = [_f64(positions, (self.surface.nv, 3))] + [_f64(v, (n,)) for v in (r, pr, y)]
        _call(lib().ddgrs_learned_add_pair, self._h, *[_ptr(v) for v in a])

    def features(self, positions) -> np.ndarray:
        f = np.zeros((self.surface.nv, 8), dtype=np.float32)
        _call(lib().ddgrs_learned_features, self._h, _ptr(_f64(positions, (self.surface.nv, 3))), _ptr(f))
        return f

    def edges(self) -> np.ndarray:
        """The mesh's undirected edges {u < v} in the network's order (ascending), ne × 2."""
        f = np.asarray(self.surface.faces, dtype=np.int64).reshape(-1, 3)
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e = np.unique(np.sort(e, axis=1), axis=0)
        return e

    def blocks(self, positions) -> np.ndarray:
        """The network's blocks at the positions: B_v (nv rows of 9, row-major), then E_uv (one row
        of 9 per edge of ``edges()``)."""
        b = np.zeros((self.surface.nv + self.edges().shape[0], 9), dtype=np.float32)
        _call(lib().ddgrs_learned_blocks, self._h, _ptr(_f64(positions, (self.surface.nv, 3))), _ptr(b))
        return b

    def loss_grad(self, pair: int, grad: bool = True):
        g = np.zeros(self.info()["parameters"], dtype=np.float32) if grad else None
        l = ctypes.c_double()
        _call(lib().ddgrs_learned_loss_grad, self._h, pair, ctypes.byref(l), _ptr(g) if grad else None)
        return float(l.value), g

    def train(self, iterations: int, batch: int = 8, lr: float = 1e-3):
        o = np.zeros(2)
        _call(lib().ddgrs_learned_train, self._h, iterations, batch, lr, _ptr(o))
        return float(o[0]), float(o[1])

    def train_validated(self, rounds: int, updates: int, validate, batch: int = 8, lr: float = 1e-3):
        """Adam in ``rounds`` rounds of ``updates`` updates, judged by GMRES: ``validate(net)``
        returns the iterations of held-out solves with (I + Q) P⁻¹ (``validate(None)``: with the
        plain preconditioner), measured before training and after every round, and the parameters
        with the fewest iterations are kept — the starting ones (Q = 0: the plain preconditioner)
        unless a round does strictly better.  The pair loss is not the judge: on the benchmark
        surfaces lowering it can raise GMRES's iteration counts.  Returns (kept round, per-round
        iterations, per-round (loss before, loss after))."""
        best_p, best_it, kept = self.params(), validate(None), 0
        its, losses = [best_it], []
        for r in range(rounds):
            losses.append(self.train(updates, batch, lr))
            it = validate(self)
            its.append(it)
            if it < best_it:
                best_p, best_it, kept = self.params(), it, r + 1
        self.set_params(best_p)
        return kept, its, losses

    def apply(self, positions, rs) -> np.ndarray:
        n = 3 * self.surface.nv + 3
        r = _f64(rs).reshape(-1, n)
        z = np.zeros_like(r)
        _call(lib().ddgrs_learned_apply_batch, self._h, _ptr(_f64(positions, (self.surface.nv, 3))), r.shape[0], _ptr(r),
              _ptr(z))
        return z.reshape(np.shape(rs))


def energy(surface: Surface, exponents: Exponents = Exponents(), eval=EnergyEval.Exact) -> EnergyReport:
    """``energy()`` (src/energy/mod.rs:168) without penalties."""
    if eval == EnergyEval.Exact:
        e, d = exact_energy_and_differential(surface, exponents)
        return EnergyReport(e, d, surface.nf * (surface.nf - 1), 0)
    if isinstance(eval, tuple) and eval[0] == "barnes_hut":
        bh = BhTree(surface, eval[1])
        e, d = bh.energy_and_differential(exponents)
        c = bh.counts()
        bh.close()
        return EnergyReport(e, d, c["pairs_exact"], c["lumped"])
    raise RsError(1, "energy", f"unsupported evaluation {eval!r}")


from . import generate  # noqa: E402,F401
from . import io  # noqa: E402,F401
from . import api  # noqa: E402,F401
