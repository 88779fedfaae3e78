"""The Python API of the reference's Repulsive Surfaces flow for this back end
(``crates/ddg-py/src/python/repulsive_surfaces.rs:404-700``, option parsing
``crates/ddg-py/src/repulsive_surfaces.rs:786-1050``), with ``device="rocm"`` selecting the MI355X
path (the reference's ``device`` takes ``"cpu"``, ``"cuda"`` or ``"auto"``, :829-842).

* ``RepulsiveSurfaceSession(positions, faces, **options)`` with ``step()``, ``run(max_steps,
  progress, time_limit)``, ``snapshot()``, ``reset()``, ``positions``, ``faces``, ``energy``,
  ``iteration``; step dicts carry the reference's keys for the fields this path produces.
* ``surface_energy(positions, faces, **options)`` → (total, differential, report).
* ``gpu_info()`` → ``available``, ``devices`` (ordinal, name, memory_bytes), ``reason``.

Positions and faces may be flat (3 per vertex / face) or (n, 3).  The options are the reference's
names; those that select something this path does not run (other metrics, energies, optimizers,
penalties, pins, obstacles, per-component rows, target schedules) raise ``Unsupported`` instead of
being ignored, and an unknown name is ``InvalidOption``.  Remeshing is on by default as in the
reference (``remesh``, ``remesh_split``, ``remesh_collapse``, ``remesh_rounds``, ``remesh_rho``,
``remesh_flip_cap``, ``remesh_order``, ``remesh_ops``; ``crates/ddg-py/src/repulsive_surfaces.rs:
1062-1111``): after every step the device remeshes the projected positions with the collapse
guard (the line search's CCD gap, or the default one with the band of the faces that intersect
then), and a remeshing that changed the mesh continues the run on a new device session with the
previous one's ℓ̄₀, last τ and targets.  ``constraints`` takes the
reference's list of dicts (``crates/ddg-py/src/repulsive_surfaces.rs:629``): the global
``{"type": "barycenter"}`` (required: it is the saddle's own row) and up to one each of
``{"type": "area", "target": t}`` and ``{"type": "volume", "target": t}`` (target optional: the
value at creation); ``newton_steps`` (fixed rounds) or ``newton_max`` (rounds until tolerance,
default 5) set the projection's Newton rounds.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import (CcdOptions, Context, EnabledOps, EnergyEval, Exponents, GmresOptions, RemeshOptions, RsError, SaddleOptions,
               Session, StartRule, StepOptions, Surface, BhOptions, energy as _energy, intersection_band, remesh as _remesh)

UNSUPPORTED = 12

# option -> default (crates/ddg-py/src/repulsive_surfaces.rs:786-1050), for the options of this path
DEFAULTS = {
    "alpha": 6.0, "beta": 12.0, "energy": "tangent_point", "tp_weight": 1.0, "low_order": True, "metric": "sobolev",
    "chi": 0.5, "cluster_leaf_size": 8, "gmres_restart": 50, "gmres_tolerance": 1e-6, "gmres_max_iterations": 200,
    "gmres_side": "right", "device": "rocm", "rocm_ordinal": 0, "gpu_determinism": "run_to_run",
    "backend": "hierarchical", "theta": 0.5, "leaf_size": 8, "energy_eval": "barnes_hut", "method":
    "projected_gradient", "barycenter_mode": "auto", "barycenter_shift": False, "armijo_c": 1e-2,
    "armijo": "standard", "start_lengths": 1.0, "start_factor": 2.0, "start_tau": None, "start": "grow_previous",
    "tau_min_rel": 1e-12, "max_halvings": None, "remesh": True, "remesh_split": 1.5, "remesh_collapse": 0.5,
    "remesh_rounds": 5, "remesh_rho": 0.5, "remesh_flip_cap": 100, "remesh_order": "flip_then_smooth", "remesh_ops": None,
    "ccd": True, "ccd_gap": 0.05, "ccd_max_iterations": 100,
    "ccd_close_factor": 0.5, "max_iterations": 1000,
    "stuck_limit": 5, "stop": "energy_stall", "stop_relative": 1e-7, "stop_window": 10, "threads": None,
    "constraints": None, "barycenter": True, "newton_steps": None, "newton_max": 5,
}
# the reference's names that only select paths outside this one: allowed at their default only
_ONLY_DEFAULT = {"energy", "metric", "gmres_side", "gpu_determinism", "method", "barycenter_mode", "barycenter_shift",
                 "armijo", "start_factor"}


def _unsupported(name, value, what):
    return RsError(UNSUPPORTED, "options", f"{name} = {value!r}: {what} is not on the MI355X path")


def _options(options: dict):
    o = dict(DEFAULTS)
    for k, v in options.items():
        if k == "cuda_ordinal":
            k = "rocm_ordinal"
        if k in ("penalties", "pins", "pin_targets", "pin_move", "ccd_obstacles", "schedules"):
            if v:
                raise _unsupported(k, v, "penalties, extra constraints and obstacles")
            continue
        if k not in DEFAULTS:
            raise RsError(1, "options", f"unknown option {k!r}")
        o[k] = v
    for k in _ONLY_DEFAULT:
        ok = o[k] == DEFAULTS[k] or (k == "barycenter_mode" and o[k] in ("auto", "global"))
        if not ok:
            raise _unsupported(k, o[k], "this choice")
    if o["device"] not in ("rocm", "auto"):
        raise _unsupported("device", o["device"], "a device other than the ROCm GPU")
    if o["backend"] not in ("hierarchical", "auto"):
        raise _unsupported("backend", o["backend"], "the dense back end")
    if o["energy_eval"] not in ("barnes_hut", "exact"):
        raise RsError(1, "options", f"energy_eval = {o['energy_eval']!r}: 'barnes_hut' or 'exact'")
    if o["stop"] not in ("energy_stall", "off"):
        raise _unsupported("stop", o["stop"], "this stop rule")
    o["_remesh"] = _remesh_options(o) if o["remesh"] else None
    rows = _constraints(o["constraints"], o["barycenter"])
    if o["newton_steps"] is not None:
        newton = ("fixed", int(o["newton_steps"]))
    else:
        newton = ("tol", int(o["newton_max"]))
    kinds = {"grow_previous": StartRule.grow_previous(o["start_lengths"]),
             "max_displacement": StartRule.max_displacement(o["start_lengths"]), "reference": StartRule.reference()}
    if o["start"] == "absolute":
        if o["start_tau"] is None:
            raise RsError(1, "options", "start_tau: required with start='absolute'")
        start = StartRule.absolute(o["start_tau"])
    elif o["start"] in kinds:
        start = kinds[o["start"]]
    else:
        raise RsError(1, "options", f"start = {o['start']!r}")
    step = StepOptions(
        exponents=Exponents(float(o["alpha"]), float(o["beta"])),
        energy_exact=o["energy_eval"] == "exact",
        theta=float(o["theta"]),
        saddle=SaddleOptions(chi=float(o["chi"]), leaf_size=int(o["cluster_leaf_size"]), w_tp=float(o["tp_weight"]),
                             c_lo=1.0 if o["low_order"] else 0.0),
        gmres=GmresOptions(int(o["gmres_restart"]), float(o["gmres_tolerance"]), int(o["gmres_max_iterations"])),
        armijo_c=float(o["armijo_c"]),
        tau_min_rel=float(o["tau_min_rel"]),
        max_halvings=-1 if o["max_halvings"] is None else int(o["max_halvings"]),
        start=start,
        constraints=tuple(rows),
        newton=newton,
        # LineSearchOptions::ccd (repulsive_surfaces.rs:1042-1048): on by default
        ccd=CcdOptions(gap=float(o["ccd_gap"]), max_iterations=int(o["ccd_max_iterations"]),
                       close_factor=float(o["ccd_close_factor"])) if o["ccd"] else None,
    )
    return o, step


def _remesh_options(o) -> RemeshOptions:
    """RemeshOptions from the remesh_* options (repulsive_surfaces.rs:1062-1111)."""
    if o["remesh_order"] not in ("flip_then_smooth", "smooth_then_flip"):
        raise RsError(1, "options", f"remesh_order = {o['remesh_order']!r}: flip_then_smooth or smooth_then_flip")
    ops = o["remesh_ops"]
    if ops is None:
        enabled = EnabledOps()
    else:
        if not isinstance(ops, (list, tuple)):
            raise RsError(1, "options", f"remesh_ops: expected a list of names, got {type(ops).__name__}")
        bad = [x for x in ops if x not in ("split", "collapse", "flip", "smooth")]
        if bad:
            raise RsError(1, "options", f"remesh_ops: {bad[0]!r} is not valid; use split, collapse, flip, smooth")
        enabled = EnabledOps(*(k in ops for k in ("split", "collapse", "flip", "smooth")))
    return RemeshOptions(split=float(o["remesh_split"]), collapse=float(o["remesh_collapse"]),
                         rounds=int(o["remesh_rounds"]), rho=float(o["remesh_rho"]),
                         flip_cap_per_vertex=int(o["remesh_flip_cap"]), order=o["remesh_order"], enabled_ops=enabled)


def _constraints(value, barycenter):
    """The nonlinear rows of the reference's ``constraints`` list (repulsive_surfaces.rs:629)."""
    if value is None:
        if not barycenter:
            raise _unsupported("barycenter", False, "a step without the barycenter row")
        return []
    if not isinstance(value, (list, tuple)):
        raise RsError(1, "options", f"constraints: expected a list of dicts, got {type(value).__name__}")
    rows, bary, seen = [], False, set()
    for i, c in enumerate(value):
        what = f"constraints[{i}]"
        if not isinstance(c, dict):
            raise RsError(1, "options", f"{what}: expected a dict")
        kind = c.get("type")
        if kind not in ("barycenter", "pins", "area", "volume"):
            raise RsError(1, "options", f"{what}.type = {kind!r}: 'barycenter', 'pins', 'area' or 'volume'")
        if kind == "pins":
            raise _unsupported(what, c, "pinned vertices")
        if c.get("component") is not None:
            raise _unsupported(what, c, "a per-component row")
        if kind == "barycenter":
            if c.get("target") is not None or set(c) - {"type", "component", "target"}:
                raise _unsupported(what, c, "a barycenter target")
            bary = True
            continue
        extra = set(c) - {"type", "component", "target"}
        if extra:
            raise _unsupported(what, c, f"{sorted(extra)} (growth schedules)")
        if kind in seen:
            raise _unsupported(what, c, f"a second {kind} row")
        seen.add(kind)
        t = c.get("target")
        rows.append((kind, None if t is None else float(t)))
    if not bary:
        raise _unsupported("constraints", value, "a step without the barycenter row")
    return rows


def _mesh(positions, faces):
    p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    return p, f


class RepulsiveSurfaceSession:
    """``RepulsiveSurfaceSession`` (crates/ddg-py/src/python/repulsive_surfaces.rs:404) on the
    device: the barycenter row (and optional area / volume rows), the CCD cap in the line search
    and remeshing after every step (session.rs:1107-1187), as the reference's defaults."""

    def __init__(self, positions, faces, **options):
        self._options, self._step_options = _options(options)
        self._p0, self._faces0 = _mesh(positions, faces)
        self._ctx = Context(int(self._options["rocm_ordinal"]))
        self.reset()

    def reset(self):
        """Back to the initial mesh with the same options."""
        self._faces = self._faces0
        self._face_origin = np.arange(self._faces.shape[0], dtype=np.int32)
        self._surface = Surface(self._ctx, self._faces, self._p0)
        self._session = Session(self._surface, self._step_options)
        self._iteration, self._stuck, self._report = 0, 0, None
        self._accepted = []
        self._status = "running"
        self._energy = None

    def _remesh_after_step(self, band):
        """Step 5 of Session::step_inner (session.rs:1107-1187): the device remeshes the projected
        positions with the collapse guard; a changed mesh gets a new surface and session that keep
        ℓ̄₀, the last τ and the targets (the face origins compose through the face parents)."""
        st = self._session.state()
        l0 = st["mean_edge_length0"]
        q = self._surface.positions()
        if self._options["ccd"]:
            gap = float(self._options["ccd_gap"]) * l0
        else:
            gap = CcdOptions().gap * l0
            band = intersection_band(self._surface, q)[0]
        out, rep = _remesh(self._ctx, self._faces, q, l0, self._options["_remesh"], guard_gap=gap, band=band)
        if out.faces.shape == self._faces.shape and np.array_equal(out.faces, self._faces):
            self._surface.set_positions(out.positions)
        else:
            targets = self._session.row_targets()
            rows = tuple((kind, targets[i]) for i, (kind, _) in enumerate(self._step_options.constraints))
            self._session.close()
            self._surface.close()
            self._faces = out.faces
            self._face_origin = self._face_origin[out.face_parents]
            self._surface = Surface(self._ctx, self._faces, out.positions)
            self._session = Session(self._surface, dataclasses.replace(self._step_options, constraints=rows))
            self._session.set_state(l0, st["last_tau"], st["barycenter_target"])
        return {"splits": rep.splits, "collapses": rep.collapses, "flips": rep.flips,
                "refused_splits": rep.refused_splits, "splits_deferred": rep.splits_deferred,
                "refused_collapses": rep.refused_collapses, "refused_flips": rep.refused_flips,
                "smoothing_moves": rep.smoothing_moves, "smoothing_halved": rep.smoothing_halved,
                "smoothing_refused": rep.smoothing_refused, "smoothing_skipped": rep.smoothing_skipped,
                "flip_cap_hits": rep.flip_cap_hits, "n_vertices": rep.n_vertices, "n_faces": rep.n_faces,
                "seconds": rep.seconds}

    def step(self) -> dict:
        remeshing = self._options["_remesh"] is not None
        band = None
        if remeshing and self._options["ccd"]:
            # the faces that intersect at the start of the step (session.rs:920-928)
            band = intersection_band(self._surface, self._surface.positions())[0]
        r = self._session.step()
        self._iteration += 1
        self._stuck = 0 if r.accepted else self._stuck + 1
        self._energy = r.energy_after
        if r.accepted:
            self._accepted.append(r.energy_after)
        w = int(self._options["stop_window"])
        converged = (self._options["stop"] == "energy_stall" and len(self._accepted) > w
                     and self._accepted[-w - 1] - self._accepted[-1] <= self._options["stop_relative"] * abs(self._accepted[-1]))
        if self._stuck >= int(self._options["stuck_limit"]):
            self._status = "stuck"
        elif converged:
            self._status = "converged"
        elif self._iteration >= int(self._options["max_iterations"]):
            self._status = "max_iterations"
        g = r.gmres
        phi_before, phi = self._session.constraints()
        remesh_report = self._remesh_after_step(band) if remeshing else None
        self._report = {
            "iteration": self._iteration, "status": "accepted" if r.accepted else "stuck",
            "session_status": self._status, "n_vertices": self._surface.nv, "n_faces": self._surface.nf,
            "energy_before": r.energy_before, "energy_after": r.energy_after, "tangent_point_after": r.energy_after,
            "penalties_after": [], "gradient_norm": r.gradient_norm, "sobolev_norm_sq": r.sobolev_norm_sq,
            "tau": r.tau, "tau0": r.tau0, "tau_ccd": r.tau_ccd, "halvings": r.halvings, "newton_steps": r.newton_rounds,
            "phi_before": list(phi_before), "phi": list(phi), "schur_pivot": r.min_pivot,
            "trials": [{"tau": t, "energy": e, "outcome": ("accepted", "armijo_failed", "non_finite")[o]}
                       for t, e, o in r.trial_list],
            "linear": {"iterations": g.iterations, "cycles": g.cycles, "converged": g.converged,
                       "relative_residual": g.relative_residual, "cycle_residuals": list(g.cycle_residuals),
                       "history": list(g.history), "fallbacks": list(r.fallbacks)},
            "remesh": remesh_report, "seconds": dict(r.seconds),
        }
        return self._report

    def run(self, max_steps=None, progress=None, time_limit=None) -> dict:
        t0 = time.perf_counter()
        e0 = self._energy
        steps, end = [], "final"
        while self._status == "running":
            if max_steps is not None and len(steps) >= max_steps:
                end = "max_steps"
                break
            if time_limit is not None and time.perf_counter() - t0 >= time_limit:
                end = "time_limit"
                break
            rep = self.step()
            steps.append(rep)
            if e0 is None:
                e0 = rep["energy_before"]
            if progress is not None and progress(rep) is False:
                end = "callback"
                break
        return {"end": end, "status": self._status, "steps_taken": len(steps), "iteration": self._iteration,
                "energy_before": e0, "energy_after": self._energy, "seconds": time.perf_counter() - t0, "steps": steps}

    def snapshot(self) -> dict:
        st = self._session.state()
        return {"positions": self.positions, "faces": self.faces, "pinned": np.zeros(0, dtype=np.int32),
                "face_origin": self._face_origin.copy(), "iteration": self._iteration,
                "energy": self._energy, "status": self._status, "error": None,
                "mean_edge_length0": st["mean_edge_length0"], "report": self._report}

    @property
    def positions(self) -> np.ndarray:
        return self._surface.positions().reshape(-1)

    @property
    def faces(self) -> np.ndarray:
        return self._faces.reshape(-1).astype(np.int32)

    @property
    def energy(self):
        return self._energy

    @property
    def iteration(self) -> int:
        return self._iteration


def surface_energy(positions, faces, **options):
    """``surface_energy`` (crates/ddg-py/src/python/repulsive_surfaces.rs:655): (total, the
    differential (3 per vertex), report) with the options alpha, beta, energy_eval, theta,
    leaf_size."""
    o, step = _options(options)
    p, f = _mesh(positions, faces)
    ctx = Context(int(o["rocm_ordinal"]))
    s = Surface(ctx, f, p)
    t0 = time.perf_counter()
    ev = EnergyEval.Exact if step.energy_exact else EnergyEval.BarnesHut(BhOptions(step.theta, int(o["leaf_size"])))
    rep = _energy(s, step.exponents, ev)
    seconds = time.perf_counter() - t0
    s.close()
    ctx.close()
    return rep.energy, np.asarray(rep.differential).reshape(-1), {
        "energy": rep.energy, "total": rep.energy, "penalties": [], "subcritical": False,
        "pairs_exact": rep.pairs_exact, "nodes_lumped": rep.nodes_lumped, "seconds": seconds}


def gpu_info() -> dict:
    """``gpu_info`` (crates/ddg-py/src/python/repulsive_surfaces.rs:683) for the ROCm path; never
    raises."""
    devices, reason = [], None
    try:
        k = 0
        while True:
            try:
                ctx = Context(k)
            except RsError:
                break
            info = ctx.device_info()
            devices.append({"ordinal": k, "name": info["name"], "compute_units": info["compute_units"],
                            "memory_bytes": info["memory_bytes"]})
            ctx.close()
            k += 1
        if not devices:
            reason = "no ROCm device can be opened"
    except Exception as e:  # the library or the runtime is missing
        reason = str(e)
    return {"rocm_compiled": True, "available": bool(devices), "devices": devices, "reason": reason}
