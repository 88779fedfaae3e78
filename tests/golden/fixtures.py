"""Readers for the golden fixtures kept under ``tests/golden/``.

* ``lean_rs/``: the Lean oracle's ``rs`` fixtures (``lean/oracle/fixtures/rs`` in the reference;
  protocol in ``lean/oracle/README.md`` "Repulsive surfaces"; the reference reads them with
  ``crates/ddg-repulsive-surfaces/tests/wp2_fixture/mod.rs``).  Floats are 16-hex-digit f64 bit
  patterns.
* ``authors/``: the authors'-program fixtures (``crates/ddg-repulsive-surfaces/tests/fixtures/
  reference``; read by ``tests/reference_fixtures.rs``): an OBJ mesh, ``NAME.txt`` key/value
  pairs, ``NAME.dE_all.txt`` / ``NAME.dE_bh.txt`` differentials and ``NAME.metric.txt``.
"""
from __future__ import annotations

import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LEAN_DIR = os.path.join(HERE, "lean_rs")
AUTHORS_DIR = os.path.join(HERE, "authors")


def hexf(tok: str) -> float:
    return struct.unpack(">d", bytes.fromhex(tok))[0]


def _lines(path):
    with open(path) as fh:
        return [ln.split() for ln in fh if ln.strip()]


def read_case(name: str) -> dict:
    """Parses ``NAME.in``: mesh, exponents, task options (as wp2_fixture/mod.rs:99 ``parse_case``)."""
    c = {"name": name, "positions": [], "faces": [], "r": [], "u": [], "constraints": [], "lines": []}
    for t in _lines(os.path.join(LEAN_DIR, name + ".in")):
        c["lines"].append(t)
        k = t[0]
        if k == "task":
            c["task"] = t[1]
        elif k == "exponents":
            c["exponents"] = (hexf(t[1]), hexf(t[2]))
        elif k == "tau":
            c["tau"] = hexf(t[1])
        elif k == "weights":
            c["weights"] = tuple(hexf(x) for x in t[1:4])
        elif k == "eps-rel":
            c["eps_rel"] = hexf(t[1])
        elif k == "r":
            c["r"].append([hexf(x) for x in t[2:5]])
        elif k == "u":
            c["u"].append([hexf(x) for x in t[1:]])
        elif k == "p":
            c["positions"].append([hexf(x) for x in t[1:4]])
        elif k == "f":
            c["faces"].append([int(x) for x in t[1:4]])
        elif k == "c":
            c["constraints"].append(t[1:])
        elif k == "armijo":
            c["armijo"] = hexf(t[1])
        elif k == "start":
            c["start"] = (t[1],) + tuple(hexf(x) for x in t[2:])
        elif k == "maxhalv":
            c["maxhalv"] = int(t[1])
        elif k == "newton":
            c["newton"] = (t[1], int(t[2]))
        elif k == "op":
            c["op"] = (t[1], int(t[2]), int(t[3]))
        elif k == "pinned":
            c["pinned"] = [int(x) for x in t[1:]]
        elif k == "legality":
            c["legality"] = set(t[1:])
    c.setdefault("exponents", (6.0, 12.0))
    c["positions"] = np.array(c["positions"], dtype=np.float64)
    c["faces"] = np.array(c["faces"], dtype=np.uint32)
    return c


def read_out(name: str) -> dict:
    """Parses ``NAME.out``: status, then the task's records keyed by their first token."""
    o = {"name": name, "rows": {}}
    for t in _lines(os.path.join(LEAN_DIR, name + ".out")):
        k = t[0]
        if k == "status":
            o["status"] = t[1]
            o["error"] = t[2:] if len(t) > 2 else []
        elif k == "energy":
            o["energy"] = hexf(t[1])
        elif k in ("d", "x", "g", "h", "z"):
            o["rows"].setdefault(k, {})[int(t[1])] = [hexf(x) for x in t[2:]]
        elif k in ("L", "B", "B0", "W", "WMW", "A", "Lfr", "Lp", "S", "J"):
            o["rows"].setdefault(k, {})[int(t[1])] = [hexf(x) for x in t[2:]]
        elif k == "M":
            o["M"] = [hexf(x) for x in t[1:]]
        elif k in ("tau", "s", "eps", "pivot", "gx", "cnx", "cqx", "energy0", "slope", "xnorm", "lbar", "tau0",
                   "energy1", "energy-final"):
            o[k] = hexf(t[1])
        elif k == "trial":
            o.setdefault("trials", []).append((int(t[1]), hexf(t[2]), t[3], hexf(t[4]), hexf(t[5])))
        elif k == "result":
            o["result"] = (t[1], hexf(t[2]), int(t[3]))
        elif k == "newton":
            o["newton"] = int(t[1])
        elif k in ("phi", "phi-before", "phin"):
            o[k] = [hexf(x) for x in t[1:]]
        elif k == "df":
            o.setdefault("df", []).append((int(t[1]), int(t[2]), [hexf(x) for x in t[3:]]))
        elif k == "mesh":
            o["mesh"] = [int(x) for x in t[1:]]
        elif k == "rows":
            o["nrows"] = [int(x) for x in t[1:]]
        elif k == "row":
            o.setdefault("crows", []).append((int(t[1]), t[2], t[3], hexf(t[4])))
        elif k == "p":
            o.setdefault("positions", []).append([hexf(x) for x in t[1:4]])
        elif k == "f":
            o.setdefault("faces", []).append([int(x) for x in t[1:4]])
        elif k == "parent":
            o.setdefault("parent", []).append(int(t[2]))
        elif k == "vmap":
            o.setdefault("vmap", []).append([int(x) for x in t[2:]])
    return o


def matrix(out: dict, key: str) -> np.ndarray:
    rows = out["rows"][key]
    return np.array([rows[i] for i in range(len(rows))])


def cases(task: str) -> list[str]:
    """The fixture names of a task whose expected status is ``ok`` (MANIFEST.tsv)."""
    names = []
    for t in _lines(os.path.join(LEAN_DIR, "MANIFEST.tsv")):
        if len(t) >= 3 and t[1] == task and t[2] == "ok" and os.path.exists(os.path.join(LEAN_DIR, t[0] + ".in")):
            names.append(t[0])
    return names


def read_obj(path: str):
    pos, faces = [], []
    with open(path) as fh:
        for ln in fh:
            t = ln.split()
            if not t:
                continue
            if t[0] == "v":
                pos.append([float(x) for x in t[1:4]])
            elif t[0] == "f":
                faces.append([int(x.split("/")[0]) - 1 for x in t[1:4]])
    return np.array(pos, dtype=np.float64), np.array(faces, dtype=np.uint32)


AUTHORS_NAMES = ("tube_regular", "torus192", "torus504")


def authors(name: str) -> dict:
    """One authors'-program fixture (reference_fixtures.rs:60 ``load``)."""
    pos, faces = read_obj(os.path.join(AUTHORS_DIR, name + ".obj"))
    kv = {}
    for t in _lines(os.path.join(AUTHORS_DIR, name + ".txt")):
        kv[t[0]] = float(t[1])
    d_all = np.array([[float(x) for x in t] for t in _lines(os.path.join(AUTHORS_DIR, name + ".dE_all.txt"))])
    d_bh = np.array([[float(x) for x in t] for t in _lines(os.path.join(AUTHORS_DIR, name + ".dE_bh.txt"))])
    return {"positions": pos, "faces": faces, "kv": kv, "d_all": d_all, "d_bh": d_bh}


def authors_metric(name: str):
    """``NAME.metric.txt``: the x-channel block of the reference's dense metric (first ``rows``
    vertices) and the 5 constraint rows."""
    rows = _lines(os.path.join(AUTHORS_DIR, name + ".metric.txt"))
    nv, nr, nc = int(rows[0][1]), int(rows[1][1]), int(rows[2][1])
    block = np.array([[float(x) for x in r] for r in rows[3:3 + nr]])
    cons = np.array([[float(x) for x in r] for r in rows[3 + nr:3 + nr + nc]])
    return nv, block, cons
