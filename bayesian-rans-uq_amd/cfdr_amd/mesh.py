"""2D finite-volume meshes in structure-of-arrays form.

Mirrors the reference ``Mesh`` (core/include/Mesh.hpp:98-268) for the
streamwise-periodic generators the forward solve uses, plus the MSH2 binary
format (core/src/Mesh.cpp:746-987) so any mesh can be exchanged with the
reference. Geometry is computed with the reference's scalar operation order
(``math`` for libm functions, elementwise IEEE arithmetic otherwise) so a
generated mesh is bit-identical to the reference generator's.

All meshes are one cell deep in z (dz = 1): every face normal has a zero z
component and every cell/face centre sits at z = dz/2.
"""
from __future__ import annotations

import math
import os
import re
import struct
from dataclasses import dataclass, field

import numpy as np

MESH_MAGIC_V2 = 0x4D534832
MESH_MAGIC_V1 = 0x4D534831
PATCH_TYPES = {"wall": 0, "inlet": 1, "outlet": 2, "symmetry": 3}


@dataclass
class Patch:
    name: str
    type: str
    faces: np.ndarray  # int32 global face ids


@dataclass
class Mesh:
    """Cells, faces (internal first, then boundary) and patches.

    Face arrays have length n_faces; 3-vectors are stored as (n, 3) float64.
    ``cell_face_ptr``/``cell_face_idx`` hold each cell's face list in the
    reference's order (a self-wrapped face appears twice in its cell).
    """

    n_internal: int
    owner: np.ndarray
    neighbor: np.ndarray
    face_center: np.ndarray
    face_normal: np.ndarray
    face_area: np.ndarray
    face_delta: np.ndarray
    face_d: np.ndarray
    face_weight: np.ndarray
    face_shift: np.ndarray
    cell_center: np.ndarray
    cell_volume: np.ndarray
    cell_face_ptr: np.ndarray
    cell_face_idx: np.ndarray
    patches: list = field(default_factory=list)
    nodes: np.ndarray = None
    wall_distance_: np.ndarray = None

    # ---- counts ---------------------------------------------------------
    @property
    def n_cells(self) -> int:
        return int(self.cell_volume.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def n_boundary(self) -> int:
        return self.n_faces - self.n_internal

    def n_patches(self) -> int:
        return len(self.patches)

    def patch_by_name(self, name: str) -> int:
        for i, p in enumerate(self.patches):
            if p.name == name:
                return i
        raise ValueError(f"Patch not found: {name}")

    @property
    def face_patch(self) -> np.ndarray:
        out = np.full(self.n_faces, -1, np.int32)
        for i, p in enumerate(self.patches):
            out[p.faces] = i
        return out

    def copy(self) -> "Mesh":
        return Mesh(
            self.n_internal, self.owner.copy(), self.neighbor.copy(), self.face_center.copy(),
            self.face_normal.copy(), self.face_area.copy(), self.face_delta.copy(),
            self.face_d.copy(), self.face_weight.copy(), self.face_shift.copy(),
            self.cell_center.copy(), self.cell_volume.copy(), self.cell_face_ptr.copy(),
            self.cell_face_idx.copy(),
            [Patch(p.name, p.type, p.faces.copy()) for p in self.patches],
            None if self.nodes is None else self.nodes.copy(),
            None if self.wall_distance_ is None else self.wall_distance_.copy(),
        )

    def validate_2d(self) -> None:
        """The solver path is planar: every face normal must lie in x-y."""
        if np.any(self.face_normal[:, 2] != 0.0):
            raise ValueError("mesh must be 2D: every face normal needs a zero z component")

    def fold_empty_faces(self):
        """The planar form of a one-layer extruded mesh (an OpenFOAM 2D case).

        Its front and back faces (unit normal along z) must make up the
        patches of type "empty"; every other face normal must lie in x-y
        exactly. In the reference's solve such a face adds, in the momentum,
        k and omega rows of its owner, the zero-gradient boundary terms
        diag += D and source += D phi_b with D = nuEff S / max(delta, 1e-20)
        and phi_b the owner's current value, and nothing else: its flux is
        zero and its in-plane normal components are zero, so it drops out of
        the x-y gradients, the fluxes and the pressure correction. Returns
        (planar mesh, indices of the kept patches, per-cell folded faces
        [n_cells, 4] as (S, delta) twice in face order, zeros where a cell
        has fewer) or (self, all patch indices, None) for a planar mesh."""
        nz = self.face_normal[:, 2]
        zf = np.abs(nz) > 0.5
        if not zf.any():
            self.validate_2d()
            return self, list(range(self.n_patches())), None
        if np.any(nz[~zf] != 0.0):
            raise ValueError("mesh must be 2D: every face normal needs a zero z component "
                             "(except the front and back faces of an 'empty' patch)")
        if np.any(zf[:self.n_internal]):
            raise ValueError("a z-normal internal face: only one-layer (2D) meshes are supported")
        patch_of = self.face_patch
        empty = [i for i, p in enumerate(self.patches) if p.type == "empty"]
        if np.any(zf & ~np.isin(patch_of, empty)):
            raise ValueError("front and back faces must belong to a patch of type 'empty'")
        for i in empty:
            if not np.all(zf[self.patches[i].faces]):
                raise ValueError(f"patch {self.patches[i].name!r} of type 'empty' holds a face "
                                 "whose normal is not along z")
        keep = ~zf
        new_id = np.cumsum(keep) - 1
        folded = np.zeros((self.n_cells, 4))
        ptr = np.zeros(self.n_cells + 1, np.int32)
        idx = []
        for c in range(self.n_cells):
            k = 0
            for f in self.cell_face_idx[self.cell_face_ptr[c]:self.cell_face_ptr[c + 1]]:
                if zf[f]:
                    if k == 2:
                        raise ValueError("a cell with more than two front/back faces")
                    folded[c, 2 * k] = self.face_area[f]
                    folded[c, 2 * k + 1] = max(self.face_delta[f], 1e-20)
                    k += 1
                else:
                    idx.append(int(new_id[f]))
            ptr[c + 1] = len(idx)
        kept = [i for i in range(self.n_patches()) if i not in empty]
        patches = [Patch(self.patches[i].name, self.patches[i].type,
                         new_id[self.patches[i].faces].astype(np.int32)) for i in kept]
        m = Mesh(self.n_internal, self.owner[keep].copy(), self.neighbor[keep].copy(),
                 self.face_center[keep].copy(), self.face_normal[keep].copy(),
                 self.face_area[keep].copy(), self.face_delta[keep].copy(),
                 self.face_d[keep].copy(), self.face_weight[keep].copy(),
                 self.face_shift[keep].copy(), self.cell_center.copy(), self.cell_volume.copy(),
                 ptr, np.asarray(idx, np.int32), patches,
                 None if self.nodes is None else self.nodes.copy(),
                 None if self.wall_distance_ is None else self.wall_distance_.copy())
        return m, kept, folded

    # ---- wall distance (core/src/Mesh.cpp:431-463) ----------------------
    def compute_wall_distance(self) -> None:
        pts = [self.face_center[p.faces] for p in self.patches if p.type == "wall"]
        if not pts or sum(len(p) for p in pts) == 0:
            self.wall_distance_ = np.full(self.n_cells, 1e10)
            return
        wp = np.concatenate(pts)
        out = np.empty(self.n_cells)
        cc = self.cell_center
        chunk = max(1, 4_000_000 // max(1, len(wp)))
        for s in range(0, self.n_cells, chunk):
            c = cc[s:s + chunk]
            dx = c[:, None, 0] - wp[None, :, 0]
            dy = c[:, None, 1] - wp[None, :, 1]
            dz = c[:, None, 2] - wp[None, :, 2]
            out[s:s + chunk] = np.sqrt(dx * dx + dy * dy + dz * dz).min(axis=1)
        self.wall_distance_ = out

    def wall_distance_valid(self) -> bool:
        return self.wall_distance_ is not None

    def wall_distance(self) -> np.ndarray:
        if self.wall_distance_ is None:
            raise RuntimeError("stale or absent wall distance; call compute_wall_distance()")
        return self.wall_distance_

    # ---- generators ------------------------------------------------------
    @staticmethod
    def make_periodic_channel_2d(nx: int, ny: int, Lx: float, h: float, re_tau: float,
                                 yplus_target: float = 1.0) -> "Mesh":
        """Mesh::makePeriodicChannel2D (core/src/Mesh.cpp:511-521): tanh
        clustering sized for y1+ at Re_tau, patches "bottom" and "top"."""
        _require_positive(Lx, "mesh length")
        _require_positive(2.0 * h, "mesh height")
        _require_positive(re_tau, "friction Reynolds number")
        _require_positive(yplus_target, "mesh y-plus target")
        xN = [Lx * i / nx for i in range(nx + 1)]
        yB = [0.0] * (nx + 1)
        return _build_periodic_2d(xN, yB, 2.0 * h, ny, yplus_target * h / re_tau, "bottom", "top",
                                  eta=None, min_nx=3)

    @staticmethod
    def make_curved_channel_periodic_2d(x_nodes, y_bottom, y_top: float, ny: int, re: float,
                                        yplus_target: float = 1.0) -> "Mesh":
        """Mesh::makeCurvedChannelPeriodic2D (core/src/Mesh.cpp:490-509)."""
        if not math.isfinite(re) or re < 0:
            raise ValueError("optional mesh Reynolds number must be finite and nonnegative")
        _require_positive(yplus_target, "mesh y-plus target")
        x_nodes = [float(v) for v in x_nodes]
        y_bottom = [float(v) for v in y_bottom]
        y1t = 0.0
        if re > 0 and len(x_nodes) == len(y_bottom):
            meanH = 0.0
            for yb in y_bottom:
                meanH += (y_top - yb)
            meanH /= float(len(y_bottom))
            Cf = 0.058 * math.pow(re, -0.2)
            uTau = math.sqrt(Cf / 2.0)
            nu = meanH / re
            y1t = yplus_target * nu / uTau
        return _build_periodic_2d(x_nodes, y_bottom, y_top, ny, y1t, "bottom_wall", "top_wall",
                                  eta=None, min_nx=3)

    @staticmethod
    def make_geometric_channel_2d(nx: int, ny: int, Lx: float, h: float, ratio: float) -> "Mesh":
        """Streamwise-periodic flat channel of height 2h whose wall-normal
        spacing grows geometrically by ``ratio`` from both walls (ny even,
        ny/2 cells per half). nx = 1 gives the one-column wall-normal mesh
        whose wrap faces join the column to itself; the reference admits that
        mesh through its MSH2 loader (core/src/Mesh.cpp:971-972)."""
        if ny < 2 or ny % 2:
            raise ValueError("geometric channel needs an even, positive ny")
        _require_positive(ratio - 1.0, "geometric ratio minus one")
        half = ny // 2
        first = h * (ratio - 1.0) / (math.pow(ratio, half) - 1.0)
        yh = [0.0] * (half + 1)
        for j in range(1, half + 1):
            yh[j] = yh[j - 1] + first * math.pow(ratio, j - 1)
        scale = h / yh[half]
        yh = [v * scale for v in yh]
        yh[half] = h
        ys = yh + [2.0 * h - yh[half - j] for j in range(1, half + 1)]
        eta = [y / (2.0 * h) for y in ys]
        eta[-1] = 1.0
        xN = [Lx * i / nx for i in range(nx + 1)]
        yB = [0.0] * (nx + 1)
        return _build_periodic_2d(xN, yB, 2.0 * h, ny, 0.0, "bottom", "top", eta=eta, min_nx=1)

    # ---- MSH2 binary I/O -------------------------------------------------
    def save_binary(self, path: str) -> None:
        """Mesh::saveBinary (core/src/Mesh.cpp:746-815), native little endian."""
        out = bytearray()
        nn = 0 if self.nodes is None else len(self.nodes)
        out += struct.pack("<I", MESH_MAGIC_V2)
        out += struct.pack("<5i", self.n_cells, self.n_faces, nn, len(self.patches),
                           self.n_internal)
        if nn:
            out += np.ascontiguousarray(self.nodes, "<f8").tobytes()
        fp = self.face_patch
        rec = np.zeros(self.n_faces, dtype=np.dtype([
            ("own", "<i4"), ("nbr", "<i4"), ("patch", "<i4"), ("c", "<f8", 3), ("n", "<f8", 3),
            ("a", "<f8"), ("delta", "<f8"), ("d", "<f8", 3), ("w", "<f8"), ("s", "<f8", 3)]))
        rec["own"], rec["nbr"], rec["patch"] = self.owner, self.neighbor, fp
        rec["c"], rec["n"], rec["a"] = self.face_center, self.face_normal, self.face_area
        rec["delta"], rec["d"], rec["w"] = self.face_delta, self.face_d, self.face_weight
        rec["s"] = self.face_shift
        out += rec.tobytes()
        for c in range(self.n_cells):
            out += struct.pack("<4d", *self.cell_center[c], self.cell_volume[c])
            f = self.cell_face_idx[self.cell_face_ptr[c]:self.cell_face_ptr[c + 1]]
            out += struct.pack("<i", len(f)) + np.asarray(f, "<i4").tobytes()
        for p in self.patches:
            for s in (p.name, p.type):
                b = s.encode()
                out += struct.pack("<i", len(b)) + b
            out += struct.pack("<i", len(p.faces)) + np.asarray(p.faces, "<i4").tobytes()
        with open(path, "wb") as fh:
            fh.write(out)

    @staticmethod
    def load_from_file(path: str) -> "Mesh":
        """Mesh::loadFromFile (core/src/Mesh.cpp:49-75): a directory is an
        OpenFOAM case or polyMesh directory, a .foam or extensionless path an
        OpenFOAM polyMesh, a .msh file the binary format."""
        if os.path.isdir(path):
            if os.path.exists(os.path.join(path, "points")) and \
                    os.path.exists(os.path.join(path, "faces")):
                return Mesh.load_openfoam(path)
            poly = os.path.join(path, "constant", "polyMesh")
            if os.path.exists(os.path.join(poly, "points")):
                return Mesh.load_openfoam(poly)
            raise RuntimeError("Cannot find polyMesh in: " + path)
        ext = os.path.splitext(path)[1]
        if ext in (".foam", ""):
            return Mesh.load_openfoam(path)
        if ext == ".msh":
            return Mesh.load_binary(path)
        raise RuntimeError("Unknown mesh format: " + ext)

    @staticmethod
    def load_openfoam(poly_mesh_dir: str) -> "Mesh":
        """Mesh::loadOpenFOAM (core/src/Mesh.cpp:119-320): points, faces,
        owner, neighbour and boundary of an ASCII polyMesh. Faces keep the
        file's order (internal faces first, then each patch from its
        startFace); a cell's face list is its faces in ascending order;
        geometry follows the reference's conventions, so every value is the
        one its loader computes. A 2D case is one layer of cells whose front
        and back faces form a patch of type "empty"."""
        j = os.path.join
        n_pts, blk = _foam_block(j(poly_mesh_dir, "points"), "points file")
        nodes = np.asarray(_foam_numbers(blk, n_pts, _foam_vector), dtype=float).reshape(-1, 3)
        n_f, blk = _foam_block(j(poly_mesh_dir, "faces"), "faces file")
        face_nodes = _foam_numbers(blk, n_f, _foam_face)
        n_own, blk = _foam_block(j(poly_mesh_dir, "owner"), f"owner file in {poly_mesh_dir}")
        owner = np.asarray(_foam_numbers(blk, n_own, _foam_label), dtype=np.int32)
        n_nbr, blk = _foam_block(j(poly_mesh_dir, "neighbour"),
                                 f"neighbour file in {poly_mesh_dir}")
        nbr_list = np.asarray(_foam_numbers(blk, n_nbr, _foam_label), dtype=np.int32)
        if n_own != n_f or n_nbr > n_f:
            raise RuntimeError("polyMesh owner/neighbour counts do not match the faces")
        if any(v < 0 or v >= n_pts for fn in face_nodes for v in fn) or \
                any(len(fn) < 3 for fn in face_nodes):
            raise RuntimeError("polyMesh face refers to a missing point")
        neighbor = np.full(n_f, -1, np.int32)
        neighbor[:n_nbr] = nbr_list
        if np.any(owner < 0) or np.any(nbr_list < 0):
            raise RuntimeError("polyMesh cell labels must be nonnegative")
        nc = int(max(owner.max(initial=-1), nbr_list.max(initial=-1))) + 1
        lists = [[] for _ in range(nc)]
        for fi in range(n_f):
            lists[owner[fi]].append(fi)
            if neighbor[fi] >= 0:
                lists[neighbor[fi]].append(fi)
        ptr = np.zeros(nc + 1, np.int32)
        ptr[1:] = np.cumsum([len(v) for v in lists])
        idx = np.asarray([f for v in lists for f in v], np.int32)
        patches = []
        for name, ptype, nfaces, start in _foam_boundary(j(poly_mesh_dir, "boundary")):
            if start < 0 or nfaces < 0 or start + nfaces > n_f:
                raise RuntimeError(f"patch {name} lies outside the face list")
            patches.append(Patch(name, ptype, np.arange(start, start + nfaces, dtype=np.int32)))
        fcen, fnorm, farea = _polygon_faces(nodes, face_nodes)
        shift = np.zeros((n_f, 3))
        cc, vol = _cells_from_faces(nc, owner, neighbor, fcen, fnorm, farea, shift, ptr, idx)
        d, delta, w = _face_metrics(n_nbr, owner, neighbor, fcen, fnorm, shift, cc)
        return Mesh(n_nbr, owner, neighbor, fcen, fnorm, farea, delta, d, w, shift, cc, vol,
                    ptr, idx, patches, nodes)

    @staticmethod
    def load_binary(path: str) -> "Mesh":
        """Mesh::loadBinary (core/src/Mesh.cpp:817-982), MSH2 and MSH1."""
        with open(path, "rb") as fh:
            buf = fh.read()
        pos = 0

        def take(fmt):
            nonlocal pos
            size = struct.calcsize(fmt)
            if pos + size > len(buf):
                raise ValueError("Truncated binary mesh: " + path)
            v = struct.unpack_from(fmt, buf, pos)
            pos += size
            return v

        (magic,) = take("<I")
        if magic not in (MESH_MAGIC_V1, MESH_MAGIC_V2):
            raise ValueError("Invalid binary mesh magic or byte order")
        has_shift = magic == MESH_MAGIC_V2
        nc, nf, nn, npt, nint = take("<5i")
        if min(nc, nf, nn, npt) < 0 or nint < 0 or nint > nf:
            raise ValueError("Invalid binary mesh count")
        nodes = np.frombuffer(buf, "<f8", 3 * nn, pos).reshape(nn, 3).copy()
        pos += 24 * nn
        fields = [("own", "<i4"), ("nbr", "<i4"), ("patch", "<i4"), ("c", "<f8", 3),
                  ("n", "<f8", 3), ("a", "<f8"), ("delta", "<f8"), ("d", "<f8", 3), ("w", "<f8")]
        if has_shift:
            fields.append(("s", "<f8", 3))
        dt = np.dtype(fields)
        rec = np.frombuffer(buf, dt, nf, pos).copy()
        pos += dt.itemsize * nf
        cc = np.empty((nc, 3))
        vol = np.empty(nc)
        ptr = [0]
        idx = []
        for c in range(nc):
            x, y, z, v = take("<4d")
            (cnt,) = take("<i")
            cc[c] = (x, y, z)
            vol[c] = v
            idx.extend(take(f"<{cnt}i"))
            ptr.append(len(idx))
        patches = []
        for _ in range(npt):
            (ln,) = take("<i")
            name = buf[pos:pos + ln].decode()
            pos += ln
            (ln,) = take("<i")
            typ = buf[pos:pos + ln].decode()
            pos += ln
            (cnt,) = take("<i")
            faces = np.asarray(take(f"<{cnt}i"), np.int32)
            patches.append(Patch(name, typ, faces))
        if pos != len(buf):
            raise ValueError("Unexpected trailing bytes in binary mesh")
        shift = rec["s"] if has_shift else np.zeros((nf, 3))
        m = Mesh(nint, rec["own"].astype(np.int32), rec["nbr"].astype(np.int32), rec["c"],
                 rec["n"], rec["a"], rec["delta"], rec["d"], rec["w"], shift, cc, vol,
                 np.asarray(ptr, np.int32), np.asarray(idx, np.int32), patches, nodes)
        _admit_loaded(m)
        return m


# ---------------------------------------------------------------------------
# OpenFOAM polyMesh input (core/src/Mesh.cpp:49-320 in the reference)
# ---------------------------------------------------------------------------
# Lines a count search passes over: banner and header keywords start with one
# of these characters (readFoamCount, core/src/Mesh.cpp:97-115).
_FOAM_SKIP_FIRST = set("/FOovCa{}")
_FOAM_TOKEN = re.compile(r"[()]|[^\s()]+")


def _foam_strip(line: str) -> str:
    cut = line.find("//")
    return (line[:cut] if cut >= 0 else line).lstrip(" \t\r")


def _foam_leading_int(text: str):
    """The integer a C++ std::stoi reads from the start of text, or None."""
    m = re.match(r"\s*[+-]?\d+", text)
    return int(m.group(0)) if m else None


def _foam_block(path: str, what: str):
    """(count, tokens of the data block) of one polyMesh file: the first
    line holding an integer (after the header) is the count, the data start
    after the next line that is exactly "(" (core/src/Mesh.cpp:85-115)."""
    try:
        with open(path) as fh:
            lines = fh.read().splitlines()
    except OSError:
        raise RuntimeError(f"Cannot open {what}") from None
    count, at = None, 0
    while at < len(lines):
        line = _foam_strip(lines[at]).rstrip(" \t\r")
        at += 1
        if not line or line[0] in _FOAM_SKIP_FIRST:
            continue
        count = _foam_leading_int(line)
        if count is not None:
            break
    if count is None:
        raise RuntimeError("Could not read count")
    while at < len(lines):
        line = _foam_strip(lines[at])
        at += 1
        if line == "(":
            return count, lines[at:]
    raise RuntimeError("Could not find data block")


def _foam_numbers(lines, count: int, per_item):
    """Parse count items of a data block; per_item(tokens, pos) -> (item, pos)."""
    toks = _FOAM_TOKEN.findall("\n".join(lines))
    out, pos = [], 0
    try:
        for _ in range(count):
            item, pos = per_item(toks, pos)
            out.append(item)
    except (IndexError, ValueError):
        raise RuntimeError("Malformed OpenFOAM data block") from None
    return out


def _foam_vector(toks, pos):
    if toks[pos] != "(" or toks[pos + 4] != ")":
        raise ValueError
    return (float(toks[pos + 1]), float(toks[pos + 2]), float(toks[pos + 3])), pos + 5


def _foam_face(toks, pos):
    nv = int(toks[pos])
    if toks[pos + 1] != "(" or toks[pos + 2 + nv] != ")":
        raise ValueError
    return [int(t) for t in toks[pos + 2:pos + 2 + nv]], pos + 3 + nv


def _foam_label(toks, pos):
    return int(toks[pos]), pos + 1


def _foam_boundary(path: str):
    """Patches (name, type, nFaces, startFace) of a polyMesh boundary file
    (core/src/Mesh.cpp:205-262); an absent file means no patches."""
    if not os.path.exists(path):
        return []
    n, lines = _foam_block(path, "boundary file")
    patches, at = [], 0
    for _ in range(n):
        name = None
        while at < len(lines):
            line = _foam_strip(lines[at]).rstrip(" \t\r")
            at += 1
            if line and line not in ("{", "}"):
                name = line
                break
        ptype, nfaces, start = "", 0, 0
        while at < len(lines):
            line = _foam_strip(lines[at])
            at += 1
            if "}" in line:
                break
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "type" and len(parts) > 1:
                ptype = parts[1][:-1] if parts[1].endswith(";") else parts[1]
            elif parts[0] == "nFaces" and len(parts) > 1:
                nfaces = _foam_leading_int(parts[1]) or 0
            elif parts[0] == "startFace" and len(parts) > 1:
                start = _foam_leading_int(parts[1]) or 0
        patches.append((name or "", ptype, nfaces, start))
    return patches


def _vnorm(v):
    return np.sqrt(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1] + v[..., 2] * v[..., 2])


def _vcross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _vdot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _polygon_faces(nodes: np.ndarray, face_nodes):
    """Centre (area centroid of the triangle fan about the node average),
    area and unit normal of every face (core/src/Mesh.cpp:268-303): each sum
    runs over the face's nodes in order, faces of equal node count at once."""
    nf = len(face_nodes)
    center = np.zeros((nf, 3))
    normal = np.zeros((nf, 3))
    area = np.zeros(nf)
    by_count = {}
    for fi, fn in enumerate(face_nodes):
        by_count.setdefault(len(fn), []).append(fi)
    for nv, ids in by_count.items():
        ids = np.asarray(ids)
        conn = np.asarray([face_nodes[i] for i in ids], dtype=np.int64)
        pts = nodes[conn]  # [m, nv, 3]
        ctr = np.zeros((len(ids), 3))
        for j in range(nv):
            ctr = ctr + pts[:, j]
        ctr = ctr / float(nv)
        an = np.zeros((len(ids), 3))
        for j in range(nv):
            an = an + _vcross(pts[:, j] - ctr, pts[:, (j + 1) % nv] - ctr) * 0.5
        a = _vnorm(an)
        safe = np.where(a > 1e-300, a, 1.0)
        n = np.where((a > 1e-300)[:, None], an / safe[:, None], 0.0)
        csum = np.zeros((len(ids), 3))
        asum = np.zeros(len(ids))
        for j in range(nv):
            pa, pb = pts[:, j], pts[:, (j + 1) % nv]
            at = 0.5 * _vdot(_vcross(pa - ctr, pb - ctr), n)
            csum = csum + ((ctr + pa) + pb) * (at / 3.0)[:, None]
            asum = asum + at
        ok = np.abs(asum) > 1e-300
        c = np.where(ok[:, None], csum / np.where(ok, asum, 1.0)[:, None], ctr)
        center[ids], normal[ids], area[ids] = c, n, a
    return center, normal, area


def _cells_from_faces(nc, owner, neighbor, fcen, fnorm, farea, shift, ptr, idx):
    """Exact cell volumes and centroids (core/src/Mesh.cpp:340-373): pyramids
    on the faces about the face-centre average for closed cells, extruded
    triangles for cells without z-normal faces; cells of equal face count at
    once, each sum over the cell's faces in its face-list order."""
    center = np.zeros((nc, 3))
    volume = np.zeros(nc)
    counts = np.diff(ptr)
    for k in np.unique(counts):
        cells = np.nonzero(counts == k)[0]
        faces = np.stack([idx[ptr[cells] + j] for j in range(k)], axis=1)  # [m, k]
        own = owner[faces] == cells[:, None]
        local = np.where(own[..., None], fcen[faces], fcen[faces] - shift[faces])
        c0 = np.zeros((len(cells), 3))
        for j in range(k):
            c0 = c0 + local[:, j]
        c0 = c0 / float(k)
        zopen = ~np.any(np.abs(fnorm[faces][..., 2]) > 1e-9, axis=1)
        hsum = np.zeros(len(cells))
        csum = np.zeros((len(cells), 3))
        for j in range(k):
            sign = np.where(own[:, j], 1.0, -1.0)
            cf = local[:, j]
            h = sign * _vdot(cf - c0, fnorm[faces[:, j]]) * farea[faces[:, j]]
            piece = np.where(zopen[:, None], (c0 + cf * 2.0) / 3.0, cf * 0.75 + c0 * 0.25)
            hsum = hsum + h
            csum = csum + piece * h[:, None]
        vol = hsum / np.where(zopen, 2.0, 3.0)
        ok = np.abs(hsum) > 1e-300
        cen = np.where(ok[:, None], csum / np.where(ok, hsum, 1.0)[:, None], c0)
        vol = np.where(vol < 1e-30, 1e-20, vol)
        center[cells], volume[cells] = cen, vol
    return center, volume


def _face_metrics(nint, owner, neighbor, fcen, fnorm, shift, cc):
    """Owner-to-neighbour vectors, distances and interpolation weights
    (core/src/Mesh.cpp:377-415)."""
    nf = owner.shape[0]
    d = np.zeros((nf, 3))
    delta = np.zeros(nf)
    w = np.full(nf, 0.5)
    fi = np.arange(nint)
    if nint:
        dd = (cc[neighbor[fi]] + shift[fi]) - cc[owner[fi]]
        d[fi] = dd
        delta[fi] = _vnorm(dd)
        bad = np.nonzero(_vdot(dd, fnorm[fi]) <= 0.0)[0]
        if bad.size:
            raise RuntimeError(f"Mesh::computeFaceGeometry: internal face {int(bad[0])} has its "
                               "neighbour behind it; periodic face without its shift?")
    fb = np.arange(nint, nf)
    if fb.size:
        dd = fcen[fb] - cc[owner[fb]]
        d[fb] = dd
        delta[fb] = _vnorm(dd)
    delta = np.where(delta < 1e-30, 1e-20, delta)
    if nint:
        dP = _vnorm(fcen[fi] - cc[owner[fi]])
        dN = _vnorm(fcen[fi] - (cc[neighbor[fi]] + shift[fi]))
        s = dP + dN
        w[fi] = np.where(s > 1e-30, dN / np.where(s > 1e-30, s, 1.0), 0.5)
    return d, delta, w


def _require_positive(value: float, name: str) -> None:
    if not math.isfinite(value) or value <= 0.0:
        raise ValueError(f"{name} must be finite and positive")


def _admit_loaded(m: Mesh) -> None:
    """Connectivity admission of the MSH2 loader (core/src/Mesh.cpp:905-975)."""
    nf, nc, ni = m.n_faces, m.n_cells, m.n_internal
    internal = np.arange(nf) < ni
    if np.any(m.owner < 0) or np.any(m.owner >= nc):
        raise ValueError("Invalid binary mesh face connectivity")
    if np.any(internal & ((m.neighbor < 0) | (m.neighbor >= nc))):
        raise ValueError("Invalid binary mesh face connectivity")
    if np.any(~internal & (m.neighbor != -1)):
        raise ValueError("Invalid binary mesh face connectivity")
    if np.any(m.face_area <= 0) or np.any(m.face_delta <= 0):
        raise ValueError("Invalid binary mesh face geometry")
    if np.any(m.cell_volume <= 0):
        raise ValueError("Invalid binary mesh cell geometry")


def _build_periodic_2d(xNodes, yBottom, yTop, ny, y1t, bottomName, topName, eta=None,
                       min_nx=3) -> Mesh:
    """Mesh::buildPeriodic2D (core/src/Mesh.cpp:523-741), restated. ``eta``
    overrides the tanh wall-normal distribution (used by the geometric
    channel), and ``min_nx`` admits the one-column mesh for it."""
    nx = len(xNodes) - 1
    if nx < min_nx or len(yBottom) != nx + 1:
        raise ValueError("periodic mesh needs nx+1 x nodes and matching yBottom samples")
    if nx <= 0 or ny <= 0:
        raise ValueError("mesh cell counts must be positive")
    if not math.isfinite(yTop) or not math.isfinite(y1t) or y1t < 0.0:
        raise ValueError("periodic mesh top and wall spacing must be finite")
    for i in range(nx + 1):
        if not math.isfinite(xNodes[i]) or (i > 0 and xNodes[i] <= xNodes[i - 1]):
            raise ValueError("periodic x nodes must be finite and strictly increasing")
        _require_positive(yTop - yBottom[i], "periodic mesh local height")
    if abs(yBottom[0] - yBottom[-1]) > 1e-12:
        raise ValueError("yBottom must be periodic (first == last sample)")
    Lx = xNodes[-1] - xNodes[0]
    _require_positive(Lx, "periodic mesh length")
    dz = 1.0

    if eta is None:
        meanH = 0.0
        for i in range(nx + 1):
            meanH += (yTop - yBottom[i])
        meanH /= (nx + 1)
        stretch = 2.0
        if y1t > 0:
            def fc(s):
                return 0.5 * meanH * (1.0 + math.tanh(s * (2.0 / ny - 1.0)) / math.tanh(s))
            if y1t < meanH / ny:
                sLo, sHi = 0.1, 20.0
                if y1t < fc(sHi):
                    stretch = sHi
                else:
                    for _ in range(100):
                        sm = 0.5 * (sLo + sHi)
                        if fc(sm) > y1t:
                            sLo = sm
                        else:
                            sHi = sm
                        if sHi - sLo < 1e-10:
                            break
                    stretch = 0.5 * (sLo + sHi)
        eta = []
        for j in range(ny + 1):
            e = float(j) / ny
            eta.append(0.5 * (1.0 + math.tanh(stretch * (2.0 * e - 1.0)) / math.tanh(stretch)))
    if len(eta) != ny + 1:
        raise ValueError("eta needs ny+1 entries")

    X = np.asarray(xNodes, float)
    YB = np.asarray(yBottom, float)
    ET = np.asarray(eta, float)
    # ynode(i, j) = yBottom[i] + (yTop - yBottom[i]) * eta[j]
    Y = YB[None, :] + (yTop - YB)[None, :] * ET[:, None]  # [ny+1, nx+1]

    ppp = (nx + 1) * (ny + 1)
    nodes = np.zeros((2 * ppp, 3))
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    nid = (jj * (nx + 1) + ii).ravel()
    nodes[nid, 0] = X[ii.ravel()]
    nodes[nid, 1] = Y[jj.ravel(), ii.ravel()]
    nodes[ppp + nid, 0] = X[ii.ravel()]
    nodes[ppp + nid, 1] = Y[jj.ravel(), ii.ravel()]
    nodes[ppp + nid, 2] = dz

    # cells: shoelace in the reference's corner order, accumulated sequentially
    px = [X[None, :-1], X[None, 1:], X[None, 1:], X[None, :-1]]
    py = [Y[:-1, :-1], Y[:-1, 1:], Y[1:, 1:], Y[1:, :-1]]
    A2 = np.zeros((ny, nx))
    cxA = np.zeros((ny, nx))
    cyA = np.zeros((ny, nx))
    for kv in range(4):
        kn = (kv + 1) % 4
        cross = px[kv] * py[kn] - px[kn] * py[kv]
        A2 = A2 + cross
        cxA = cxA + (px[kv] + px[kn]) * cross
        cyA = cyA + (py[kv] + py[kn]) * cross
    A = 0.5 * A2
    ccx = (cxA / (6.0 * A)).ravel()
    ccy = (cyA / (6.0 * A)).ravel()
    nC = nx * ny
    cc = np.stack([ccx, ccy, np.full(nC, 0.5 * dz)], axis=1)
    vol = (np.abs(A) * dz).ravel()

    nVint, nWrap, nHint = (nx - 1) * ny, ny, nx * (ny - 1)
    nInternal = nVint + nWrap + nHint
    nF = nInternal + 2 * nx
    own = np.empty(nF, np.int32)
    nbr = np.full(nF, -1, np.int32)
    fc_ = np.zeros((nF, 3))
    fn = np.zeros((nF, 3))
    area = np.zeros(nF)
    delta = np.zeros(nF)
    d = np.zeros((nF, 3))
    w = np.full(nF, 0.5)
    shift = np.zeros((nF, 3))

    def cid(i, j):
        return j * nx + i

    def edge(sl, x0, y0, x1, y1, plus_x):
        ex = x1 - x0
        ey = y1 - y0
        ln = np.sqrt(ex * ex + ey * ey)
        fc_[sl, 0] = 0.5 * (x0 + x1)
        fc_[sl, 1] = 0.5 * (y0 + y1)
        fc_[sl, 2] = 0.5 * dz
        area[sl] = ln * dz
        if plus_x:
            fn[sl, 0] = ey / ln
            fn[sl, 1] = -ex / ln
        else:
            fn[sl, 0] = -ey / ln
            fn[sl, 1] = ex / ln

    def set_dw(sl, cO, cN):
        dd = cN - cO
        d[sl] = dd
        delta[sl] = np.maximum(np.sqrt(dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1] + dd[:, 2] * dd[:, 2]), 1e-20)
        a = fc_[sl] - cO
        b = fc_[sl] - cN
        dP = np.sqrt(a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1] + a[:, 2] * a[:, 2])
        dN = np.sqrt(b[:, 0] * b[:, 0] + b[:, 1] * b[:, 1] + b[:, 2] * b[:, 2])
        s = dP + dN
        w[sl] = np.where(s > 1e-30, dN / np.where(s > 1e-30, s, 1.0), 0.5)

    fi = 0
    # internal vertical faces: j-major, i-minor
    if nVint:
        J, I = np.meshgrid(np.arange(ny), np.arange(nx - 1), indexing="ij")
        J, I = J.ravel(), I.ravel()
        sl = slice(fi, fi + nVint)
        own[sl] = cid(I, J)
        nbr[sl] = cid(I + 1, J)
        edge(sl, X[I + 1], Y[J, I + 1], X[I + 1], Y[J + 1, I + 1], True)
        set_dw(sl, cc[own[sl]], cc[nbr[sl]])
        fi += nVint
    # wrap faces
    J = np.arange(ny)
    sl = slice(fi, fi + nWrap)
    own[sl] = cid(nx - 1, J)
    nbr[sl] = cid(0, J)
    edge(sl, np.full(ny, X[nx]), Y[J, nx], np.full(ny, X[nx]), Y[J + 1, nx], True)
    shift[sl, 0] = Lx
    set_dw(sl, cc[own[sl]], cc[nbr[sl]] + shift[sl])
    fi += nWrap
    # internal horizontal faces
    if nHint:
        J, I = np.meshgrid(np.arange(ny - 1), np.arange(nx), indexing="ij")
        J, I = J.ravel(), I.ravel()
        sl = slice(fi, fi + nHint)
        own[sl] = cid(I, J)
        nbr[sl] = cid(I, J + 1)
        edge(sl, X[I], Y[J + 1, I], X[I + 1], Y[J + 1, I + 1], False)
        set_dw(sl, cc[own[sl]], cc[nbr[sl]])
        fi += nHint
    # bottom and top walls
    I = np.arange(nx)
    patches = []
    for name, row, plus_x, cellrow in ((bottomName, 0, True, 0), (topName, ny, False, ny - 1)):
        sl = slice(fi, fi + nx)
        own[sl] = cid(I, cellrow)
        edge(sl, X[I], Y[row, I], X[I + 1], Y[row, I + 1], plus_x)
        dd = fc_[sl] - cc[own[sl]]
        d[sl] = dd
        delta[sl] = np.maximum(np.sqrt(dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1] + dd[:, 2] * dd[:, 2]), 1e-20)
        patches.append(Patch(name, "wall", np.arange(fi, fi + nx, dtype=np.int32)))
        fi += nx

    # per-cell face lists in face order (owner then neighbour)
    lists = [[] for _ in range(nC)]
    for f in range(nF):
        lists[own[f]].append(f)
        if nbr[f] >= 0:
            lists[nbr[f]].append(f)
    ptr = np.zeros(nC + 1, np.int32)
    ptr[1:] = np.cumsum([len(v) for v in lists])
    idx = np.asarray([f for v in lists for f in v], np.int32)
    return Mesh(nInternal, own, nbr, fc_, fn, area, delta, d, w, shift, cc, vol, ptr, idx,
                patches, nodes)


def hill_bottom(x: float) -> float:
    """ERCOFTAC periodic-hill surface in units of H on the cell [0, 9]: the
    published piecewise-cubic polynomial (UFR 3-30, x in mm on [0, 54],
    H = 28 mm), one half hill at each end. Same geometry as the reference's
    golden hills case (tests/tools/record_goldens.cpp:529-556)."""
    xmm = min(x, 9.0 - x) * 28.0
    xmm = max(xmm, 0.0)
    if xmm <= 9.0:
        y = min(28.0, 28.0 + 6.775070969851e-3 * xmm * xmm - 2.124527775800e-3 * xmm * xmm * xmm)
    elif xmm <= 14.0:
        y = (2.507355893131e+01 + 9.754803562315e-01 * xmm - 1.016116352781e-01 * xmm * xmm
             + 1.889794677828e-03 * xmm * xmm * xmm)
    elif xmm <= 20.0:
        y = (2.579601052357e+01 + 8.206693007457e-01 * xmm - 9.055370274339e-02 * xmm * xmm
             + 1.626510569859e-03 * xmm * xmm * xmm)
    elif xmm <= 30.0:
        y = (4.046435022819e+01 - 1.379581654948e+00 * xmm + 1.945884504128e-02 * xmm * xmm
             - 2.070318932190e-04 * xmm * xmm * xmm)
    elif xmm <= 40.0:
        y = (1.792461334664e+01 + 8.743920332081e-01 * xmm - 5.567361123058e-02 * xmm * xmm
             + 6.277731764683e-04 * xmm * xmm * xmm)
    elif xmm <= 54.0:
        y = max(0.0, 5.639011190988e+01 - 2.010520359035e+00 * xmm + 1.644919857549e-02 * xmm * xmm
                + 2.674976141766e-05 * xmm * xmm * xmm)
    else:
        y = 0.0
    return y / 28.0


def make_hill_mesh(nx: int, ny: int, Lx: float = 9.0, y_top: float = 3.035, re: float = 5600.0,
                   yplus_target: float = 1.0) -> Mesh:
    """Periodic hill of height h = 1 on [0, Lx] x [hill(x), y_top]."""
    xN = [Lx * i / nx for i in range(nx + 1)]
    yB = [hill_bottom(x) for x in xN]
    yB[nx] = yB[0]
    return Mesh.make_curved_channel_periodic_2d(xN, yB, y_top, ny, re, yplus_target)
