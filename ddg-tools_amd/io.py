"""OBJ input and output of triangle meshes (``crates/ddg-repulsive-surfaces/src/io.rs:38-110``,
through the rules of ``crates/ddg-core/src/io.rs:166-350`` ``parse_obj`` / ``obj_to_string``).

* ``v x y z`` lines give positions (finite numbers); ``f`` lines give faces whose corners are
  1-based or negative (relative to the vertices defined so far) indices, ``v``, ``v/vt``,
  ``v//vn`` or ``v/vt/vn``; texture coordinates, normals, groups, materials and comments (``#``)
  are ignored; a line ending in ``\\`` continues on the next one.  Positive indices may refer to
  vertices defined later and are range-checked at the end.
* Only triangles are accepted: a face with another number of corners is ``PolygonFace``.
* Parse errors name the line (``Parse``); unreadable files are ``Io``.

Meshes are returned as (positions (nv × 3 float64), faces (nf × 3 uint32)); their topology and
positions are validated by ``Surface`` when a device surface is made of them (``Surface::new``,
``validate_positions``, src/surface.rs:132,513).
"""
from __future__ import annotations

import math

import numpy as np

from . import RsError

PARSE, IO, POLYGON_FACE = 9, 10, 11


def _parse_error(line: int, message: str) -> RsError:
    return RsError(PARSE, "parse_obj", f"line {line}: {message}", a=line)


def _index(token: str, defined: int, line: int) -> int:
    try:
        k = int(token)
    except ValueError:
        raise _parse_error(line, f"cannot parse '{token}' as a vertex index") from None
    if k > 0:
        return k - 1
    if k < 0:
        if -k > defined:
            raise _parse_error(line, f"relative vertex index {k} reaches before the first of the {defined} defined so far")
        return defined + k
    raise _parse_error(line, "vertex index 0 is invalid: OBJ indices start at 1")


def parse_obj(text: str):
    """(positions, faces) of OBJ text (``parse_obj``, src/io.rs:46)."""
    positions, faces, face_lines = [], [], []
    lines = text.splitlines()
    joined, start = "", None
    logical = []
    for i, raw in enumerate(lines, 1):
        raw = raw.rstrip()
        if raw.endswith("\\"):
            joined += raw[:-1] + " "
            start = start or i
            continue
        if start is not None:
            logical.append((joined + raw, start))
            joined, start = "", None
        else:
            logical.append((raw, i))
    if start is not None:
        logical.append((joined, start))
    for line, number in logical:
        tokens = line.split("#", 1)[0].split()
        if not tokens:
            continue
        if tokens[0] == "v":
            if len(tokens) < 4:
                raise _parse_error(number, f"vertex needs 3 coordinates, found {len(tokens) - 1}")
            p = []
            for t in tokens[1:4]:
                try:
                    x = float(t)
                except ValueError:
                    raise _parse_error(number, f"cannot parse '{t}' as a number") from None
                if not math.isfinite(x):
                    raise _parse_error(number, f"non-finite number '{t}'")
                p.append(x)
            positions.append(p)
        elif tokens[0] == "f":
            corners = [_index(t.split("/", 1)[0], len(positions), number) for t in tokens[1:]]
            if len(corners) < 3:
                raise _parse_error(number, f"face has {len(corners)} corners, at least 3 are required")
            faces.append(corners)
            face_lines.append(number)
    for f, (corners, number) in enumerate(zip(faces, face_lines)):
        for v in corners:
            if v >= len(positions):
                raise _parse_error(number, f"face refers to vertex {v + 1}, but the file defines {len(positions)} vertices")
        if len(corners) != 3:
            raise RsError(POLYGON_FACE, "parse_obj", f"face {f} has {len(corners)} corners; only triangles are accepted",
                          a=f, b=len(corners))
    return (np.array(positions, dtype=np.float64).reshape(-1, 3), np.array(faces, dtype=np.uint32).reshape(-1, 3))


def read_obj(path):
    """(positions, faces) of an OBJ file (``read_obj``, src/io.rs:38)."""
    try:
        with open(path) as fh:
            text = fh.read()
    except OSError as e:
        raise RsError(IO, "read_obj", f"{path}: {e}") from None
    return parse_obj(text)


def obj_to_string(positions, faces) -> str:
    """OBJ text of a triangle mesh: ``v`` and ``f`` lines, faces in the caller's vertex order,
    numbers in the shortest form that reads back exactly (``obj_to_string``, src/io.rs:57)."""
    p = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    f = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    bad = np.flatnonzero(~np.isfinite(p).all(axis=1))
    if bad.size:
        raise RsError(7, "obj_to_string", f"position of vertex {int(bad[0])} is not finite", a=int(bad[0]))
    if f.size and (f.min() < 0 or f.max() >= p.shape[0]):
        raise RsError(1, "obj_to_string", "a face refers to a vertex that does not exist")
    out = ["# written by ddg-tools_amd"]
    out += [f"v {x!r} {y!r} {z!r}" for x, y, z in p.tolist()]
    out += [f"f {a + 1} {b + 1} {c + 1}" for a, b, c in f.tolist()]
    return "\n".join(out) + "\n"


def write_obj(path, positions, faces):
    """Writes a triangle mesh as an OBJ file (``write_obj``, src/io.rs:70)."""
    text = obj_to_string(positions, faces)
    try:
        with open(path, "w") as fh:
            fh.write(text)
    except OSError as e:
        raise RsError(IO, "write_obj", f"{path}: {e}") from None
