"""Writes one-layer OpenFOAM polyMesh cases (ASCII) for the fixtures.

A structured quadrilateral mesh of (nx+1) x (ny+1) nodes (x[i, j], y[i, j])
is extruded from z = 0 to z = depth. Cells are numbered i-fastest; internal
faces follow OpenFOAM's upper-triangular order (each cell's faces to
higher-numbered neighbours, in neighbour order); the boundary faces come in
the patches bottom (wall), top (wall), inlet, outlet and frontAndBack
(empty), in that order. Every face's nodes run so that its normal points
out of its owner (checked on the computed normal).

Test infrastructure: the fixture generator writes a case with it, the
reference's solver and this project's reader load the same files.
"""
from __future__ import annotations

import os

import numpy as np

HEADER = """FoamFile
{{
    version     2.0;
    format      ascii;
    class       {cls};
    location    "constant/polyMesh";
    object      {obj};
}}
"""


def _area_normal(pts):
    c = pts.mean(axis=0)
    s = np.zeros(3)
    for j in range(len(pts)):
        s += np.cross(pts[j] - c, pts[(j + 1) % len(pts)] - c)
    return s


def _orient(nodes, face, outward):
    """The face's node list, reversed if its normal opposes outward."""
    return face if _area_normal(nodes[face]) @ outward > 0 else face[::-1]


def write_extruded(case_dir: str, x: np.ndarray, y: np.ndarray, depth: float) -> str:
    """Write constant/polyMesh under case_dir; x and y are [nx+1, ny+1] node
    coordinates. Returns the polyMesh directory."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    nx, ny = x.shape[0] - 1, x.shape[1] - 1
    per = (nx + 1) * (ny + 1)

    def pid(i, j, k):
        return k * per + j * (nx + 1) + i

    nodes = np.zeros((2 * per, 3))
    for k in range(2):
        for j in range(ny + 1):
            for i in range(nx + 1):
                nodes[pid(i, j, k)] = (x[i, j], y[i, j], depth * k)

    def cid(i, j):
        return j * nx + i

    def centre(i, j):
        return np.array([x[i:i + 2, j:j + 2].mean(), y[i:i + 2, j:j + 2].mean(), 0.5 * depth])

    faces, owner, neighbour = [], [], []

    def add(face, o, n, outward):
        faces.append(_orient(nodes, face, outward))
        owner.append(o)
        if n is not None:
            neighbour.append(n)

    # internal faces: per cell, the east neighbour (c + 1) then the north one (c + nx)
    for j in range(ny):
        for i in range(nx):
            c = cid(i, j)
            if i + 1 < nx:
                f = [pid(i + 1, j, 0), pid(i + 1, j + 1, 0), pid(i + 1, j + 1, 1), pid(i + 1, j, 1)]
                add(f, c, cid(i + 1, j), centre(i + 1, j) - centre(i, j))
            if j + 1 < ny:
                f = [pid(i, j + 1, 0), pid(i + 1, j + 1, 0), pid(i + 1, j + 1, 1), pid(i, j + 1, 1)]
                add(f, c, cid(i, j + 1), centre(i, j + 1) - centre(i, j))
    n_internal = len(faces)
    patches = []

    def patch(name, ptype, items):
        start = len(faces)
        for f, o, out in items:
            add(f, o, None, out)
        patches.append((name, ptype, len(faces) - start, start))

    patch("bottom", "wall", [([pid(i, 0, 0), pid(i + 1, 0, 0), pid(i + 1, 0, 1), pid(i, 0, 1)],
                              cid(i, 0), np.array([0.0, -1.0, 0.0])) for i in range(nx)])
    patch("top", "wall", [([pid(i, ny, 0), pid(i + 1, ny, 0), pid(i + 1, ny, 1), pid(i, ny, 1)],
                           cid(i, ny - 1), np.array([0.0, 1.0, 0.0])) for i in range(nx)])
    patch("inlet", "inlet", [([pid(0, j, 0), pid(0, j + 1, 0), pid(0, j + 1, 1), pid(0, j, 1)],
                              cid(0, j), np.array([-1.0, 0.0, 0.0])) for j in range(ny)])
    patch("outlet", "outlet", [([pid(nx, j, 0), pid(nx, j + 1, 0), pid(nx, j + 1, 1),
                                 pid(nx, j, 1)], cid(nx - 1, j), np.array([1.0, 0.0, 0.0]))
                               for j in range(ny)])
    fb = []
    for k, sgn in ((0, -1.0), (1, 1.0)):
        for j in range(ny):
            for i in range(nx):
                fb.append(([pid(i, j, k), pid(i + 1, j, k), pid(i + 1, j + 1, k), pid(i, j + 1, k)],
                           cid(i, j), np.array([0.0, 0.0, sgn])))
    patch("frontAndBack", "empty", fb)

    poly = os.path.join(case_dir, "constant", "polyMesh")
    os.makedirs(poly, exist_ok=True)

    def write(name, cls, count, body):
        with open(os.path.join(poly, name), "w") as fh:
            fh.write(HEADER.format(cls=cls, obj=name))
            fh.write(f"\n{count}\n(\n{body}\n)\n")

    write("points", "vectorField", len(nodes),
          "\n".join("(%r %r %r)" % tuple(float(v) for v in p) for p in nodes))
    write("faces", "faceList", len(faces),
          "\n".join(f"{len(f)}(" + " ".join(str(v) for v in f) + ")" for f in faces))
    write("owner", "labelList", len(owner), "\n".join(str(v) for v in owner))
    write("neighbour", "labelList", n_internal, "\n".join(str(v) for v in neighbour))
    body = "\n".join(f"    {n}\n    {{\n        type            {t};\n"
                     f"        nFaces          {nf};\n        startFace       {s};\n    }}"
                     for n, t, nf, s in patches)
    with open(os.path.join(poly, "boundary"), "w") as fh:
        fh.write(HEADER.format(cls="polyBoundaryMesh", obj="boundary"))
        fh.write(f"\n{len(patches)}\n(\n{body}\n)\n")
    return poly


def channel_nodes(nx: int, ny: int, length: float, height: float, ratio: float,
                  wave: float) -> tuple:
    """Nodes of a developing-channel mesh: x uniform, y graded towards both
    walls by a geometric ratio per cell, the interior rows displaced by
    wave * sin(2 pi x / length) * 4 eta (1 - eta) (eta = y / height), so the
    internal faces tilt and the cells are trapezoids; the walls stay flat."""
    xs = np.linspace(0.0, length, nx + 1)
    half = ny // 2
    steps = ratio ** np.arange(half)
    e = np.concatenate([[0.0], np.cumsum(steps)]) / steps.sum() * 0.5
    eta = np.concatenate([e, 1.0 - e[::-1][1:]]) if ny % 2 == 0 else None
    if eta is None:
        raise ValueError("ny must be even")
    x = np.repeat(xs[:, None], ny + 1, axis=1)
    bump = wave * np.sin(2.0 * np.pi * xs / length)
    y = height * (eta[None, :] + bump[:, None] * 4.0 * eta[None, :] * (1.0 - eta[None, :]))
    return x, y
