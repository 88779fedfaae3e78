"""Benchmark of the Hs-descent hot path on one MI355X.

``python bench.py [--gpus 1] [--steps K] [--warmup W]`` prints one JSON line.

Workload: the plain Hs descent step on the bumpy icosphere at level 6 (40,962 vertices, 81,920
faces, 16 seeded bumps; α = 6, β = 12): Barnes–Hut energy and differential (θ = 0.5, leaf 8), the
block cluster tree metric (χ = 0.5, 64-pair dense blocks, FP32 values), the reference
preconditioner (L′ = sparse LDLᵀ with the global barycenter row, augmented with L̃^V_{2−s}), GMRES
(restart 50) to relative residual 1e-6 in FP64, the Armijo line search and the barycenter
projection.  Positions live in HBM for the whole run; the timed region is K consecutive steps of
one session started from the seeded input (after W warm-up steps of another session from the same
input), bracketed by device synchronisations.

The roofline object describes the kernel with the largest device time per step among the two
modelled ones:

* ``bh_gradient`` (FP32 VALU bound): 63 FLOP per interaction — an ordered exact face pair (S, T)
  or a lumped (face, node) pair — value and gradient in the shared (6, 12) form;
* ``bt_panel_c12`` (HBM bound), the near-field panels of the metric product B̃/B̃⁰ (12 channels):
  8 bytes per near entry (two FP32 values), plus 96 bytes in and 96 bytes out per face (12
  channels as FP32 hi/lo pairs).

Both are reported (``rooflines``); their launch durations come from HIP events on the library's own
stream during a profiled replay of the first timed step.  ``traffic`` is the HBM bytes per launch
from ``rocprofv3 --pmc`` FETCH_SIZE and WRITE_SIZE passes over a one-step run, each scaled by the
counter's reading for the 256 MiB copy of ``ddg-tools_amd/lib/hbm_copy`` (run as child processes
before this process touches the GPU; ``--no-traffic`` skips them).  ``cpu_baseline`` runs the FP64
oracle (a C++ restatement of the reference's Rust step) on the host for the same step from the
same input.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_INTERACTION = 63
BYTES_PER_NEAR_ENTRY = 8
BYTES_PER_FACE_ROWS = 192
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0  # dense BF16 MFMA
PEAK_HBM_GBS = 8000.0
METRIC = "Hs descent steps/sec at 81,920 faces; GMRES iters, learned vs plain precond"


def lumped_face_node_pairs(tree: dict, lists: dict) -> int:
    """Σ over leaves of (lumped nodes) × (faces of the leaf): the (face, node) interactions."""
    start, end = tree["start"].astype(np.int64), tree["end"].astype(np.int64)
    nn = start.size
    leaf = np.ones(nn, dtype=bool)
    leaf[:-1] = start[1:] != start[:-1]  # an internal node's next pre-order node is its left child
    sizes = (end - start)[leaf]
    counts = np.diff(lists["lumped_off"].astype(np.int64))
    return int((counts * sizes).sum())


KERNEL_NAMES = {"bt_panel_c12": "bt_panel_kernel<12, 9, 2, float>", "bh_gradient": "bh_gradient_kernel<0>"}


def _counter_kib(directory: str, counter: str, substr: str):
    """Mean counter value (KiB) per dispatch of the kernels whose name contains `substr`."""
    vals = []
    for path in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] == counter and substr in r["Kernel_Name"]:
                    vals.append(float(r["Counter_Value"]))
    return sum(vals) / len(vals) if vals else None


def measure_traffic(timeout_s: int = 240):
    """HBM bytes per launch of the modelled kernels from `rocprofv3 --pmc` (FETCH_SIZE, then
    WRITE_SIZE, each in its own run of a one-step bench and of the 256 MiB copy probe, scaled by
    what the counter reads for the copy).  Runs as child processes before this process touches
    the GPU; returns {} when rocprofv3 or the probe is missing or a pass fails."""
    prof = shutil.which("rocprofv3")
    probe = os.path.join(ROOT, "ddg-tools_amd", "lib", "hbm_copy")
    if not prof or not os.path.exists(probe):
        return {}
    known_kib = 256 * 1024
    env = {**os.environ, "TMPDIR": "/tmp"}
    res = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        scaled = {k: 0.0 for k in KERNEL_NAMES}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            runs = {
                "bench": [sys.executable, os.path.abspath(__file__), "--steps", "1", "--warmup", "0",
                          "--no-cpu-baseline", "--no-traffic", "--no-learned"],
                "copy": [probe, "256", "3"],
            }
            for tag, cmd in runs.items():
                d = os.path.join(tmp, f"{counter}_{tag}")
                r = subprocess.run(["timeout", "-s", "KILL", str(timeout_s), prof, "--pmc", counter, "--output-format",
                                    "csv", "-d", d, "-o", "run", "--", *cmd], cwd="/tmp", env=env,
                                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                if r.returncode != 0:
                    return {"error": f"{tag} under --pmc {counter} exited with {r.returncode}"}
            copy_kib = _counter_kib(os.path.join(tmp, f"{counter}_copy"), counter, "copy_kernel")
            if not copy_kib:
                return {}
            factor = known_kib / copy_kib
            res[counter + "_scale"] = factor
            for key, name in KERNEL_NAMES.items():
                kib = _counter_kib(os.path.join(tmp, f"{counter}_bench"), counter, name)
                if kib is None:
                    scaled[key] = None
                elif scaled[key] is not None:
                    scaled[key] += kib * factor * 1024.0
        res.update({k: v for k, v in scaled.items() if v is not None})
    return res


def scale_run(args) -> int:
    """The level-7 scale run: consecutive steps of one session from the seeded input (Barnes-Hut
    theta 0.25, plain preconditioner), timed; then the same trajectory again with the positions read
    after steps 1, 10 and 20, whose energies (the device's Barnes-Hut energy on fresh lists) are
    compared with the FP64 oracle's at the same positions."""
    from __graft_entry__ import load_package

    pkg = load_package()
    ctx = pkg.Context(0)
    p, f = pkg.generate.config_bumpy_l7()
    steps = args.steps if args.steps != 5 else 20
    opts = pkg.StepOptions(theta=0.25)
    surf = pkg.Surface(ctx, f, p)
    ses = pkg.Session(surf, opts)
    for _ in range(args.warmup):
        ses.step()
    ses.close()
    surf.set_positions(p)
    ses = pkg.Session(surf, opts)
    ctx.synchronize()
    print(f"bench: L7, timing {steps} step(s)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    reps = [ses.step() for _ in range(steps)]
    ctx.synchronize()
    t1 = time.perf_counter()
    ses.close()
    checks = [k for k in (1, 10, 20) if k <= steps]
    surf.set_positions(p)
    ses = pkg.Session(surf, opts)
    drift = []
    import oracle

    threads = min(16, os.cpu_count() or 1)
    for k in range(1, steps + 1):
        ses.step()
        if k in checks:
            x = surf.positions()
            bh = pkg.BhTree(surf, pkg.BhOptions(theta=0.25, leaf_size=opts.saddle.leaf_size))
            e_dev = bh.energy(opts.exponents)
            bh.close()
            e_orc = oracle.BhTree(f, x, theta=0.25, threads=threads).energy()
            drift.append({"step": k, "energy": e_dev, "oracle_energy": e_orc, "rel": (e_dev - e_orc) / e_orc})
            print(f"bench: L7 step {k}: energy {e_dev:.12g}, oracle {e_orc:.12g}", file=sys.stderr, flush=True)
    ses.close()
    ms = (t1 - t0) * 1e3 / steps
    learned = None
    if not args.no_learned:
        # retrained on trajectories at this size (fields of 64 bumps like the timed one): unless set
        # otherwise, 8 fields x 10 plain steps with three Krylov pairs each (320 pairs; the level-6
        # recipe's 2,000 steps would take minutes at this size)
        surf.set_positions(p)
        la = argparse.Namespace(**vars(args))
        if "--learned-fields" not in sys.argv:
            la.learned_fields = 8
        if "--learned-steps" not in sys.argv:
            la.learned_steps = 10
        if "--learned-krylov" not in sys.argv:
            la.learned_krylov = 3
        learned = learned_leg(pkg, ctx, la, opts, p, f, ms, [r.gmres.iterations for r in reps], level=7,
                              bumps=(64, (0.0, 0.05), (0.05, 0.2)), steps=steps)
    line = {
        "metric": "Hs descent steps/sec at 327,680 faces (level-7 scale run)",
        "value": 1e3 / ms,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair and metric kernels, fp64 solver",
        "data": "synthetic (seeded bumpy icosphere)",
        "config": {
            "workload": "plain Hs descent step, bumpy icosphere L7 (163,842 vertices, 327,680 faces, 64 bumps), "
                        "Barnes-Hut theta 0.25, chi 0.5, GMRES restart 50 to 1e-6",
            "faces": int(f.shape[0]),
            "preconditioner": "plain (reference); the learned leg in `learned`",
            "gmres_iterations": [r.gmres.iterations for r in reps],
            # steps whose solve switched to the extended near field (the FP32 products' rounding held the
            # explicit residual above 1e-6), and steps that ended unconverged
            "extended_solve_rate": sum(1 for r in reps if r.gmres.extended) / steps,
            "unconverged_steps": sum(1 for r in reps if not r.gmres.converged),
            "energy_drift": drift,
        },
        "learned": learned,
    }
    print(json.dumps(line))
    return 0


def learned_leg(pkg, ctx, args, opts, p, f, plain_ms, plain_its, level=6, bumps=(16, (0.0, 0.1), (0.1, 0.4)),
                steps=None):
    """The learned preconditioner: pairs from plain descent on `fields` bump fields other than the
    timed one (`steps` each), Adam on the device, then the timed steps again from the same start
    with (I + Q) P⁻¹ (GMRES to the same 1e-6; the plain solve on a miss)."""
    gen = pkg.generate
    steps = steps or args.steps
    base, _ = gen.icosphere(level)
    surf = pkg.Surface(ctx, f, p)
    # per plain step: (g, A′⁻¹ g) and `learned_krylov` Krylov pairs (none for the default recipe: 2,000
    # (residual, plain-preconditioned solve) pairs from 50 steps on each of 40 bump fields)
    npairs = args.learned_pairs or (1 + args.learned_krylov) * args.learned_fields * args.learned_steps
    net = pkg.Learned(surf, seed=7, max_pairs=npairs)
    net.set_krylov_pairs(args.learned_krylov)
    t0 = time.perf_counter()
    for k in range(args.learned_fields):
        surf.set_positions(gen.bumps(base, bumps[0], bumps[1], bumps[2], 1000 + k))
        ses = pkg.Session(surf, opts)
        ses.set_learned(net, 1)
        for _ in range(args.learned_steps):
            ses.step()
        ses.close()
    t_pairs = time.perf_counter() - t0
    per_epoch = max(1, net.info()["pairs"] // args.learned_batch)
    iters = args.learned_epochs * per_epoch
    print(f"bench: learned: {net.info()['pairs']} pairs in {t_pairs:.1f} s, training {iters} updates", file=sys.stderr,
          flush=True)
    # the judge of the kept parameters: GMRES iterations of the first two steps on a held-out bump
    # field (neither a training field nor the timed one)
    p_val = gen.bumps(base, bumps[0], bumps[1], bumps[2], 999)

    def validate(model):
        surf.set_positions(p_val)
        ses = pkg.Session(surf, opts)
        if model is not None:
            ses.set_learned(model, 2)
        total = 0
        for _ in range(2):
            r = ses.step()
            total += r.learned_iterations if (model is not None and not r.learned_fallback) else \
                r.gmres.iterations + (10 ** 6 if model is not None else 0)
        ses.close()
        return total

    t0 = time.perf_counter()
    kept, val_its, losses = net.train_validated(args.learned_epochs, per_epoch, validate, batch=args.learned_batch,
                                                lr=args.learned_lr)
    before, after = losses[0][0], min(l[1] for l in losses)
    ctx.synchronize()
    t_train = time.perf_counter() - t0
    # the plain steps once more right before the learned ones, in the same conditions (a fresh
    # session on this surface), so that the ratio compares like with like
    surf.set_positions(p)
    ses = pkg.Session(surf, opts)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ses.step()
    ctx.synchronize()
    plain_ms_here = (time.perf_counter() - t0) * 1e3 / steps
    ses.close()
    surf.set_positions(p)
    ses = pkg.Session(surf, opts)
    ses.set_learned(net, 2)
    ctx.synchronize()
    t0 = time.perf_counter()
    reps = [ses.step() for _ in range(steps)]
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ses.close()
    # one profiled step: the network's forward layers of width 64 on BF16 MFMA against its dense peak
    surf.set_positions(p)
    ses = pkg.Session(surf, opts)
    ses.set_learned(net, 2)
    ctx.set_profiling(True)
    ses.step()
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    ses.close()
    nv = p.shape[0]
    ne = net.edges().shape[0]
    flop = 2.0 * (nv * (8 * 64 + 4 * 128 * 64) + ne * 64 * 64)
    rows = kt.get("ln_rows_mfma")
    head = kt.get("ln_rows_valu")
    fwd = None
    if rows:
        a = flop / (rows["ms"] * 1e-3) / 1e12
        # the layers of width 64 (BF16 inputs, FP32 accumulation); the two 64 x 9 heads run on VALU
        fwd = {"bound": "mfma", "achieved": a, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s", "frac": a / PEAK_BF16_TFLOPS,
               "kernel_ms_per_step": rows["ms"], "launches": rows["launches"],
               "head_valu_ms_per_step": head["ms"] if head else None,
               "work": f"2 x ({nv} vertices x (8x64 + 4 x 128x64) + {ne} edges x 64x64) multiply-adds, BF16 inputs"}
    its = [r.learned_iterations if not r.learned_fallback else r.gmres.iterations for r in reps]
    return {
        "pairs": net.info()["pairs"],
        "fields": args.learned_fields,
        "steps_per_field": args.learned_steps,
        "seconds_pairs": t_pairs,
        "training_updates": iters,
        "batch": args.learned_batch,
        "seconds_training": t_train,
        "loss_before": before,
        "loss_after": after,
        # validated training: GMRES iterations of two held-out steps before training (plain) and after
        # each epoch; the kept parameters are those of `kept_epoch` (0: Q = 0, the plain preconditioner)
        "validation_iterations": val_its,
        "kept_epoch": kept,
        "gmres_iterations_learned": [r.learned_iterations for r in reps],
        "gmres_iterations_plain": plain_its,
        "fallbacks": sum(1 for r in reps if r.learned_fallback == 1),
        "fallback_rate": sum(1 for r in reps if r.learned_fallback) / len(reps),
        "skipped_after_misses": sum(1 for r in reps if r.learned_fallback == 2),
        "converged": all(r.gmres.converged for r in reps),
        "ms_per_step_learned": ms,
        "ms_per_step_plain": plain_ms_here,
        "ms_per_step_plain_timed_run": plain_ms,
        "time_ratio_learned_over_plain": ms / plain_ms_here,
        "mean_iterations_used": float(np.mean(its)),
        "forward_ms_per_step": float(np.mean([r.seconds["learned"] for r in reps]) * 1e3),
        "network_roofline": fwd,
    }


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, f, p, opts) -> dict:
    """The FP64 oracle's step (this repository's C++ restatement of the reference's Rust CPU path,
    which cannot be built here) on the host cores, on a bounded sample of the same workload: on all
    the threads the GPU's share of the box gives (16 at most), `cpu_warmup` warm-up steps then the
    median of `cpu_steps` consecutive steps from the seeded input; one step each on 4 threads and
    on 1 thread; per-phase medians."""
    import oracle

    nproc = os.cpu_count() or 1
    full = min(16, nproc)
    out = {"unit": "steps/s", "kind": "port", "nproc": nproc, "cpu_model": _cpu_model()}
    runs = {}
    for threads, warm, steps in ((full, args.cpu_warmup, args.cpu_steps), (4, 0, 1), (1, 0, 1)):
        if threads > nproc or threads in runs:
            continue
        print(f"bench: oracle baseline, {warm} + {steps} step(s) on {threads} threads", file=sys.stderr, flush=True)
        ors = oracle.Session(f, p, theta=opts.theta)
        for _ in range(warm):
            ors.step(threads=threads)
        times, phases = [], []
        for _ in range(steps):
            c0 = time.perf_counter()
            rep = ors.step(threads=threads)
            times.append(time.perf_counter() - c0)
            phases.append(rep["seconds"])
        med = float(np.median(times))
        runs[threads] = {"steps_per_s": 1.0 / med, "median_s": med, "steps": steps, "warmup": warm,
                         "phases_s": {k: float(np.median([ph[k] for ph in phases])) for k in phases[0]}}
    out["value"] = runs[full]["steps_per_s"]
    out["cores"] = full
    out["runs"] = {str(k): v for k, v in runs.items()}
    out["sample"] = (f"FP64 oracle Hs step (same workload, Barnes-Hut theta 0.5): {full} threads, {args.cpu_warmup} "
                     f"warm-up + median of {args.cpu_steps} steps from the seeded input; 4 and 1 threads, one step each")
    return out


def ccd_leg(pkg, ctx, args, p, f) -> dict:
    """The continuous collision detection at the same mesh (first_contact_with_band, ccd.rs:781):
    a smooth synthetic motion of four mean edge lengths for its fastest vertex over [0, 4] (three
    windows), the device's median of 5 calls after one warm-up, the FP64 oracle once on one host
    thread (the CPU baseline of this leg; its report must equal the device's bit for bit), and the
    step with the CCD cap on (3 steps after one warm-up: ms per step and the CCD's share)."""
    vel = np.stack([np.sin(3 * p[:, 1]), np.cos(2 * p[:, 2]), -2.0 * p[:, 2] * np.abs(p[:, 0])], axis=1)
    e = {(min(a, b), max(a, b)) for t in f.tolist() for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))}
    ea = np.array(sorted(e))
    l0 = float(np.mean(np.linalg.norm(p[ea[:, 0]] - p[ea[:, 1]], axis=1)))
    vel *= l0 / np.max(np.linalg.norm(vel, axis=1))
    surf = pkg.Surface(ctx, f, p)
    pkg.first_contact(surf, vel, 4.0)
    times, rep = [], None
    for _ in range(5):
        c0 = time.perf_counter()
        rep = pkg.first_contact(surf, vel, 4.0)
        times.append(time.perf_counter() - c0)
    out = {"workload": "first_contact_with_band, 81,920 faces, smooth synthetic motion of 4 mean edge lengths",
           "device_ms": float(np.median(times) * 1e3), "windows": rep.windows, "pairs_tested": rep.pairs_tested,
           "iterations": rep.iterations, "contact": rep.contact}
    if not args.no_cpu_baseline:
        import oracle

        print("bench: oracle CCD baseline on one thread", file=sys.stderr, flush=True)
        c0 = time.perf_counter()
        ref = oracle.ccd_first_contact(f, p, vel, 4.0)
        out["cpu_ms"] = (time.perf_counter() - c0) * 1e3
        out["cpu_kind"] = "port"
        out["cpu_cores"] = 1
        out["matches_oracle"] = (rep.tau_max == ref["tau_max"] and rep.contact == ref["contact"]
                                 and (rep.iterations, rep.pairs_tested, rep.windows)
                                 == (ref["iterations"], ref["pairs_tested"], ref["windows"]))
    opts = pkg.StepOptions(ccd=pkg.CcdOptions())
    ses = pkg.Session(surf, opts)
    ses.step()
    ctx.synchronize()
    c0 = time.perf_counter()
    reps = [ses.step() for _ in range(3)]
    ctx.synchronize()
    out["step_ms_with_ccd"] = (time.perf_counter() - c0) * 1e3 / 3
    out["step_ccd_ms"] = float(np.mean([r.seconds["ccd"] for r in reps]) * 1e3)
    out["step_tau_ccd"] = [r.tau_ccd for r in reps]
    ses.close()
    surf.close()
    return out


def remesh_leg(pkg, ctx, args, p, f, opts) -> dict:
    """One remeshing pass as the reference's step runs it (session.rs:1109-1187): the positions after
    one descent step, L0 the input's mean edge length, default RemeshOptions and the collapse guard
    at the default CCD gap (0.05 L0) with the band of the faces that intersect now.  The device's
    median of 5 passes after one warm-up against the FP64 oracle's pass on one host thread (the CPU
    baseline of this leg; the device's faces, parents, vertex map and report must equal the
    oracle's, positions to the printed error); then a stress pass of a harder deformation (16
    bumps of up to 0.3 on the sphere) with thousands of splits and flips."""
    e = {(min(a, b), max(a, b)) for t in f.tolist() for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))}
    ea = np.array(sorted(e))
    l0 = float(np.mean(np.linalg.norm(p[ea[:, 0]] - p[ea[:, 1]], axis=1)))
    surf = pkg.Surface(ctx, f, p)
    ses = pkg.Session(surf, opts)
    ses.step()
    q = ses.positions()
    ses.close()
    band = pkg.intersection_band(surf, q)[0]
    surf.close()
    gap = pkg.CcdOptions().gap * l0
    out = {}
    for name, pos in (("after_step", q), ("stress", None)):
        if pos is None:
            p6, f6 = pkg.generate.icosphere(6)
            pos = pkg.generate.bumps(np.asarray(p6), 16, (0.0, 0.3), (0.1, 0.4), 3)
            band = None
        pkg.remesh(ctx, f, pos, l0, pkg.RemeshOptions(), guard_gap=gap, band=band)
        times, res = [], None
        for _ in range(5):
            c0 = time.perf_counter()
            res = pkg.remesh(ctx, f, pos, l0, pkg.RemeshOptions(), guard_gap=gap, band=band)
            times.append(time.perf_counter() - c0)
        dev, rep = res
        o = {"device_ms": float(np.median(times) * 1e3), "splits": rep.splits, "collapses": rep.collapses,
             "flips": rep.flips, "smoothing_moves": rep.smoothing_moves, "smoothing_iterations": rep.smoothing_iterations,
             "faces_out": rep.n_faces}
        if not args.no_cpu_baseline:
            import oracle

            print(f"bench: oracle remeshing pass ({name}) on one thread", file=sys.stderr, flush=True)
            c0 = time.perf_counter()
            ref, rref = oracle.remesh(f, pos, l0, guard_gap=gap, band=band)
            o["cpu_ms"] = (time.perf_counter() - c0) * 1e3
            o["cpu_kind"] = "port"
            o["cpu_cores"] = 1
            same = (dev.faces.shape == ref["faces"].shape and bool((dev.faces == ref["faces"]).all())
                    and bool((dev.face_parents == ref["face_parents"]).all())
                    and bool((dev.vertex_map.sources == ref["map_sources"]).all())
                    and all(getattr(rep, k) == rref[k] for k in ("splits", "collapses", "flips", "smoothing_moves")))
            o["matches_oracle"] = same
            o["max_position_error"] = (float(np.abs(dev.positions - ref["positions"]).max())
                                       if dev.positions.shape == ref["positions"].shape else None)
        out[name] = o
    out["workload"] = ("remesh (remesh/mod.rs:235) of 81,920 faces, L0 = the input's mean edge length, default "
                       "options, collapse guard at gap 0.05 L0; after_step: the positions after one Hs step; "
                       "stress: 16 bumps of up to 0.3")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5, help="timed oracle steps on all threads (median)")
    ap.add_argument("--cpu-warmup", type=int, default=2, help="oracle warm-up steps on all threads")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 --pmc traffic passes")
    ap.add_argument("--config", choices=["l6", "l7"], default="l6",
                    help="l6: the headline step (default); l7: the level-7 scale run (theta 0.25, energy drift "
                         "against the FP64 oracle after steps 1, 10 and 20, fallback rate)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned preconditioner leg")
    ap.add_argument("--no-ccd", action="store_true", help="skip the collision detection leg")
    ap.add_argument("--no-remesh", action="store_true", help="skip the remeshing leg")
    ap.add_argument("--learned-fields", type=int, default=40, help="bump fields of the training trajectories")
    ap.add_argument("--learned-steps", type=int, default=50, help="plain steps per field")
    ap.add_argument("--learned-krylov", type=int, default=0,
                    help="Krylov pairs per training step after the step's own (residual, solve) pair")
    ap.add_argument("--learned-pairs", type=int, default=0, help="pairs kept (0: every recorded pair)")
    ap.add_argument("--learned-epochs", type=int, default=10)
    ap.add_argument("--learned-batch", type=int, default=8)
    ap.add_argument("--learned-lr", type=float, default=2e-3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU (asked for {args.gpus}); how the step shards over GPUs is described, not built.",
              file=sys.stderr)
        return 2
    if args.config == "l7":
        return scale_run(args)

    traffic = {} if args.no_traffic else measure_traffic()

    from __graft_entry__ import load_package

    pkg = load_package()
    ctx = pkg.Context(0)
    p, f = pkg.generate.config_bumpy_l6()
    nf = f.shape[0]
    opts = pkg.StepOptions()
    surf = pkg.Surface(ctx, f, p)

    print(f"bench: warmup {args.warmup} step(s)", file=sys.stderr, flush=True)
    ses = pkg.Session(surf, opts)
    for _ in range(args.warmup):
        ses.step()
    ses.close()

    surf.set_positions(p)
    ses = pkg.Session(surf, opts)
    ctx.synchronize()
    print(f"bench: timing {args.steps} step(s)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    reps = [ses.step() for _ in range(args.steps)]
    ctx.synchronize()
    t1 = time.perf_counter()
    ses.close()
    ms = (t1 - t0) * 1e3 / args.steps

    # Profiled replay of the first step (same input): per-kernel device times.
    surf.set_positions(p)
    bh = pkg.BhTree(surf, pkg.BhOptions(theta=opts.theta, leaf_size=opts.saddle.leaf_size))
    counts = bh.counts()
    face_node = lumped_face_node_pairs(bh.tree(), bh.lists())
    bh.close()
    fm = pkg.FaceMetricTree(surf, opts.exponents, pkg.ClusterTreeOptions(opts.saddle.chi, opts.saddle.leaf_size,
                                                                          opts.saddle.leaf_pairs))
    mstats = fm.stats()
    fm.close()
    ses = pkg.Session(surf, opts)
    ctx.set_profiling(True)
    r1 = ses.step()
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    ses.close()
    device_ms = sum(v["ms"] for v in kt.values())
    launches = sum(v["launches"] for v in kt.values())
    print(f"bench: profiled step: {launches} launches, {device_ms:.2f} ms of kernels", file=sys.stderr)
    for k, v in sorted(kt.items(), key=lambda kv: -kv[1]["ms"]):
        print(f"bench:   {k:24s} {v['launches']:6d} launches {v['ms']:9.3f} ms", file=sys.stderr)

    rooflines = {}
    g = kt.get("bh_gradient")
    if g:
        kms = g["ms"] / g["launches"]
        flop = FLOP_PER_INTERACTION * (counts["pairs_exact"] + face_node)
        a = flop / (kms * 1e-3) / 1e12
        rooflines["bh_gradient"] = {
            "bound": "valu", "achieved": a, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": a / PEAK_FP32_TFLOPS,
            "traffic": traffic.get("bh_gradient"), "kernel_ms": kms, "ms_per_step": g["ms"],
            "work": f"{FLOP_PER_INTERACTION} FLOP x ({counts['pairs_exact']} exact ordered pairs + {face_node} "
                    f"lumped (face, node) pairs)",
        }
    m = kt.get("bt_panel_c12")
    if m:
        kms = m["ms"] / m["launches"]
        nbytes = BYTES_PER_NEAR_ENTRY * mstats["near_total"] + BYTES_PER_FACE_ROWS * nf
        a = nbytes / (kms * 1e-3) / 1e9
        rooflines["bt_panel_c12"] = {
            "bound": "hbm", "achieved": a, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": a / PEAK_HBM_GBS,
            "traffic": traffic.get("bt_panel_c12"), "algorithmic_bytes": nbytes, "kernel_ms": kms,
            "ms_per_step": m["ms"], "launches_per_step": m["launches"],
            "work": f"{BYTES_PER_NEAR_ENTRY} B x {mstats['near_total']} near entries + {BYTES_PER_FACE_ROWS} B x {nf} faces",
        }
    dominant = max(rooflines, key=lambda k: rooflines[k]["ms_per_step"]) if rooflines else None
    roof = dict(rooflines[dominant], kernel=dominant) if dominant else None

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, f, p, opts)

    its = [r.gmres.iterations for r in reps]
    stage = {k: float(np.mean([r.seconds[k] for r in reps]) * 1e3) for k in reps[0].seconds}
    learned = None
    if not args.no_learned:
        learned = learned_leg(pkg, ctx, args, opts, p, f, ms, its)
    ccd = None if args.no_ccd else ccd_leg(pkg, ctx, args, np.asarray(p), np.asarray(f, dtype=np.uint32))
    remesh = None if args.no_remesh else remesh_leg(pkg, ctx, args, np.asarray(p), np.asarray(f, dtype=np.uint32),
                                                     opts)
    line = {
        "metric": METRIC,
        "value": 1e3 / ms,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair and metric kernels, fp64 solver",
        "data": "synthetic (seeded bumpy icosphere)",
        "config": {
            "workload": "plain Hs descent step, bumpy icosphere L6 (40,962 vertices, 81,920 faces), Barnes-Hut "
                        "theta 0.5, chi 0.5, GMRES restart 50 to 1e-6",
            "faces": nf,
            "preconditioner": "plain (reference); the learned leg in `learned`",
            "gmres_iterations": its,
            "extended_solves": sum(1 for r in reps if r.gmres.extended),
            "energy": [r.energy_before for r in reps] + [reps[-1].energy_trial],
            "stage_ms": stage,
            "device_kernel_ms_step1": device_ms,
        },
        "roofline": roof,
        "rooflines": rooflines,
        "cpu_baseline": cpu,
        "learned": learned,
        "ccd": ccd,
        "remesh": remesh,
        "traffic_scale": {k: v for k, v in traffic.items() if k.endswith("_scale")},
        "traffic_error": traffic.get("error"),
        "kernel_ms_step1": {k: round(v["ms"], 4) for k, v in sorted(kt.items(), key=lambda kv: -kv[1]["ms"])[:16]},
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
