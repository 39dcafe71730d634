"""Synthetic OpenFOAM polyMesh generator for PLBFS cases.

The reference ships no meshes, so cases are built here: a structured box of
hexahedra (optionally split into 6 Kuhn tetrahedra per hex), interior vertices
jittered by a seeded uniform fraction of the local spacing, cell labels
optionally permuted by a seeded random permutation, and boundary faces grouped
into patches by the box side they lie on.

The writer emits the ASCII layout the reference reader parses
(src/unstructured_mesh.cpp:236-507, 1912-1987): the ``note`` line with
``nPoints/nCells/nFaces/nInternalFaces`` on 0-based line 12 of ``owner``, list
bodies opened by a line starting with ``(`` and closed by a line starting with
``)``, points as ``(x y z)``, faces as ``n(v0 ... vn-1)``, and a ``boundary``
file whose entries give ``type X;`` and ``nFaces N;`` in order, patches being
contiguous from ``nInternalFaces``.  Points are single precision (the reference
reads them as ``float``, unstructured_mesh.h:50) and are written with 9
significant digits so that the reader recovers the exact float32 values.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

# box sides in the order patches are written
SIDES = ("x0", "x1", "y0", "y1", "z0", "z1")

# hex local vertex offsets (i, j, k) and outward-oriented faces (right-hand rule)
_HEX_V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                   [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)
_HEX_F = np.array([[0, 4, 7, 3], [1, 2, 6, 5], [0, 1, 5, 4],
                   [3, 7, 6, 2], [0, 3, 2, 1], [4, 5, 6, 7]], dtype=np.int64)
# Kuhn split: one tet per permutation of the axes, all sharing the v0-v6 diagonal
_AXIS_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


@dataclass
class PolyMesh:
    points: np.ndarray          # (P,3) float32
    face_off: np.ndarray        # (F+1,) int32
    face_verts: np.ndarray      # (sum,) int32
    owner: np.ndarray           # (F,) int32
    neighbour: np.ndarray       # (NI,) int32
    n_cells: int
    patch_names: list           # patch type tokens, in order
    patch_sizes: list           # faces per patch
    meta: dict = field(default_factory=dict)

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def n_internal(self) -> int:
        return int(self.neighbour.shape[0])

    @property
    def bface_patch(self) -> np.ndarray:
        return np.repeat(np.arange(len(self.patch_sizes), dtype=np.int32),
                         np.asarray(self.patch_sizes, dtype=np.int64))


def _lattice(nx, ny, nz):
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    # vertex id = i + (nx+1)*(j + (ny+1)*k)
    ijk = np.stack([i.ravel(order="F"), j.ravel(order="F"), k.ravel(order="F")], axis=1)
    return ijk.astype(np.int64)


def _vid(i, j, k, nx, ny):
    return i + (nx + 1) * (j + (ny + 1) * k)


def box_mesh(nx, ny, nz, lengths=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), tets=False,
             jitter=0.0, jitter_seed=0, permute_seed=None, patch_types=None) -> PolyMesh:
    """Build a box mesh of nx*ny*nz hexes (or 6*nx*ny*nz Kuhn tets).

    jitter: interior vertices move by U(-jitter, jitter) * spacing per axis (seeded).
    permute_seed: if not None, cell labels are randomly permuted (seeded).
    patch_types: dict side -> OpenFOAM patch type token; sides sharing a token are
      merged into one patch.  Default: every side its own token equal to its name.
    """
    if patch_types is None:
        patch_types = {s: s for s in SIDES}
    lx, ly, lz = (float(v) for v in lengths)
    hx, hy, hz = lx / nx, ly / ny, lz / nz
    ijk = _lattice(nx, ny, nz)
    pts = np.empty((ijk.shape[0], 3), dtype=np.float64)
    pts[:, 0] = origin[0] + ijk[:, 0] * hx
    pts[:, 1] = origin[1] + ijk[:, 1] * hy
    pts[:, 2] = origin[2] + ijk[:, 2] * hz
    if jitter > 0.0:
        rng = np.random.default_rng(jitter_seed)
        d = rng.uniform(-jitter, jitter, size=pts.shape) * np.array([hx, hy, hz])
        interior = ((ijk[:, 0] > 0) & (ijk[:, 0] < nx) & (ijk[:, 1] > 0) & (ijk[:, 1] < ny)
                    & (ijk[:, 2] > 0) & (ijk[:, 2] < nz))
        pts[interior] += d[interior]
    pts32 = pts.astype(np.float32)

    # hex cells, x fastest
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.ravel(order="F"), cj.ravel(order="F"), ck.ravel(order="F")
    corner = np.stack([_vid(ci + _HEX_V[v, 0], cj + _HEX_V[v, 1], ck + _HEX_V[v, 2], nx, ny)
                       for v in range(8)], axis=1)  # (H,8)
    if not tets:
        cell_faces = corner[:, _HEX_F]  # (C,6,4)
    else:
        # per axis permutation (a,b,c): v0 -> +e_a -> +e_b -> v6
        H = corner.shape[0]
        tet_list = []
        lat = ijk.astype(np.float64)
        for (a, b, c) in _AXIS_PERMS:
            s1 = np.zeros(3, dtype=np.int64); s1[a] = 1
            s2 = s1.copy(); s2[b] = 1
            p0 = _vid(ci, cj, ck, nx, ny)
            p1 = _vid(ci + s1[0], cj + s1[1], ck + s1[2], nx, ny)
            p2 = _vid(ci + s2[0], cj + s2[1], ck + s2[2], nx, ny)
            p3 = _vid(ci + 1, cj + 1, ck + 1, nx, ny)
            tet = np.stack([p0, p1, p2, p3], axis=1)
            # orient positively on the unjittered lattice
            e1 = lat[tet[:, 1]] - lat[tet[:, 0]]
            e2 = lat[tet[:, 2]] - lat[tet[:, 0]]
            e3 = lat[tet[:, 3]] - lat[tet[:, 0]]
            det = np.einsum("ij,ij->i", np.cross(e1, e2), e3)
            neg = det < 0
            tet[neg, 1], tet[neg, 2] = tet[neg, 2].copy(), tet[neg, 1].copy()
            tet_list.append(tet)
        tets_arr = np.stack(tet_list, axis=1).reshape(H * 6, 4)  # tets of a hex adjacent
        # outward faces of a positive tet: opp p3 (0,2,1), opp p2 (0,1,3), opp p1 (0,3,2), opp p0 (1,2,3)
        tf = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], dtype=np.int64)
        cell_faces = tets_arr[:, tf]  # (C,4,3)
    C, nfc, nvf = cell_faces.shape

    # cell labels
    if permute_seed is not None:
        rng = np.random.default_rng(permute_seed)
        label = rng.permutation(C).astype(np.int64)
    else:
        label = np.arange(C, dtype=np.int64)

    # unique faces by the three smallest vertex ids (unique for these meshes)
    inst = cell_faces.reshape(C * nfc, nvf)
    inst_cell = np.repeat(label, nfc)
    srt = np.sort(inst, axis=1)[:, :3]
    key = (srt[:, 0] << 42) | (srt[:, 1] << 21) | srt[:, 2]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    first = np.ones(ks.shape[0], dtype=bool)
    first[1:] = ks[1:] != ks[:-1]
    grp_start = np.flatnonzero(first)
    grp_size = np.diff(np.append(grp_start, ks.shape[0]))
    if np.any(grp_size > 2):
        raise ValueError("non-manifold face set")
    pair = grp_size == 2
    a_idx = order[grp_start[pair]]
    b_idx = order[grp_start[pair] + 1]
    ca, cb = inst_cell[a_idx], inst_cell[b_idx]
    own_is_a = ca < cb
    int_inst = np.where(own_is_a, a_idx, b_idx)     # owner's outward orientation
    int_owner = np.minimum(ca, cb)
    int_nb = np.maximum(ca, cb)
    srt_int = np.lexsort((int_nb, int_owner))
    int_inst, int_owner, int_nb = int_inst[srt_int], int_owner[srt_int], int_nb[srt_int]

    b_inst = order[grp_start[~pair]]
    b_owner = inst_cell[b_inst]
    bv = inst[b_inst]
    bijk = ijk[bv]  # (B,nvf,3)
    side = np.full(b_inst.shape[0], -1, dtype=np.int64)
    dims = (nx, ny, nz)
    for ax in range(3):
        lo = np.all(bijk[:, :, ax] == 0, axis=1)
        hi = np.all(bijk[:, :, ax] == dims[ax], axis=1)
        side[lo] = 2 * ax
        side[hi] = 2 * ax + 1
    if np.any(side < 0):
        raise ValueError("unclassified boundary face")
    # patches: distinct type tokens in first-appearance order over SIDES
    tokens = []
    for s in SIDES:
        t = patch_types[s]
        if t not in tokens:
            tokens.append(t)
    side_patch = np.array([tokens.index(patch_types[s]) for s in SIDES], dtype=np.int64)
    b_patch = side_patch[side]
    bsrt = np.lexsort((b_owner, b_patch))
    b_inst, b_owner, b_patch = b_inst[bsrt], b_owner[bsrt], b_patch[bsrt]
    sizes = [int(np.count_nonzero(b_patch == p)) for p in range(len(tokens))]

    all_inst = np.concatenate([int_inst, b_inst])
    fv = inst[all_inst].astype(np.int32)
    F = fv.shape[0]
    face_off = (np.arange(F + 1, dtype=np.int64) * nvf).astype(np.int32)
    owner = np.concatenate([int_owner, b_owner]).astype(np.int32)
    mesh = PolyMesh(points=pts32, face_off=face_off, face_verts=fv.ravel(), owner=owner,
                    neighbour=int_nb.astype(np.int32), n_cells=int(C), patch_names=tokens,
                    patch_sizes=sizes,
                    meta=dict(nx=nx, ny=ny, nz=nz, lengths=(lx, ly, lz), origin=tuple(origin),
                              tets=bool(tets), jitter=float(jitter), permuted=permute_seed is not None,
                              cell_label=label))
    return mesh


_BANNER = (
    "/*--------------------------------*- C++ -*----------------------------------*\\\n"
    "  =========                 |\n"
    "  \\\\      /  F ield         | OpenFOAM: The Open Source CFD Toolbox\n"
    "   \\\\    /   O peration     | Website:  https://openfoam.org\n"
    "    \\\\  /    A nd           | Version:  7\n"
    "     \\\\/     M anipulation  |\n"
    "\\*---------------------------------------------------------------------------*/\n"
)


def _header(cls, obj, note=None):
    s = _BANNER + "FoamFile\n{\n    version     2.0;\n    format      ascii;\n"
    s += "    class       %s;\n" % cls
    if note is not None:
        s += "    note        \"%s\";\n" % note
    s += "    location    \"constant/polyMesh\";\n    object      %s;\n}\n" % obj
    s += "// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //\n\n"
    return s


def _write_list(path, header, n, body_lines):
    with open(path, "w") as f:
        f.write(header)
        f.write("\n%d\n(\n" % n)
        f.write(body_lines)
        f.write(")\n\n// ************************************************************************* //\n")


def write_polymesh(mesh: PolyMesh, directory: str) -> str:
    """Write owner/neighbour/points/faces/boundary into ``directory`` (created).

    Returns the import prefix (directory path with a trailing slash), which is
    what the reference's ``import_file`` holds.
    """
    os.makedirs(directory, exist_ok=True)
    P, F, NI, C = mesh.points.shape[0], mesh.n_faces, mesh.n_internal, mesh.n_cells
    note = "nPoints:%d  nCells:%d  nFaces:%d  nInternalFaces:%d" % (P, C, F, NI)
    hdr = _header("labelList", "owner", note)
    # the note must sit on 0-based line 12 of owner
    assert hdr.split("\n")[12].lstrip().startswith("note"), "owner header layout"
    _write_list(os.path.join(directory, "owner"), hdr, F,
                "\n".join(map(str, mesh.owner.tolist())) + "\n")
    _write_list(os.path.join(directory, "neighbour"), _header("labelList", "neighbour", note), NI,
                "\n".join(map(str, mesh.neighbour.tolist())) + ("\n" if NI else ""))
    p = mesh.points.astype(np.float32)
    sx = np.char.mod("%.9g", p[:, 0].astype(np.float64))
    sy = np.char.mod("%.9g", p[:, 1].astype(np.float64))
    sz = np.char.mod("%.9g", p[:, 2].astype(np.float64))
    lines = np.char.add(np.char.add(np.char.add(np.char.add(np.char.add("(", sx), " "), sy), " "),
                        np.char.add(sz, ")"))
    _write_list(os.path.join(directory, "points"), _header("vectorField", "points"), P,
                "\n".join(lines.tolist()) + "\n")
    fo, fv = mesh.face_off, mesh.face_verts
    nv = np.diff(fo)
    if np.all(nv == nv[0]):
        arr = fv.reshape(F, int(nv[0]))
        cols = [np.char.mod("%d", arr[:, q]) for q in range(arr.shape[1])]
        body = np.char.add(str(int(nv[0])) + "(", cols[0])
        for q in range(1, len(cols)):
            body = np.char.add(np.char.add(body, " "), cols[q])
        body = np.char.add(body, ")")
        text = "\n".join(body.tolist()) + "\n"
    else:
        text = "".join("%d(%s)\n" % (nv[f], " ".join(map(str, fv[fo[f]:fo[f + 1]]))) for f in range(F))
    _write_list(os.path.join(directory, "faces"), _header("faceList", "faces"), F, text)
    with open(os.path.join(directory, "boundary"), "w") as f:
        f.write(_header("polyBoundaryMesh", "boundary"))
        f.write("%d\n(\n" % len(mesh.patch_names))
        start = NI
        for name, n in zip(mesh.patch_names, mesh.patch_sizes):
            f.write("    %s\n    {\n        type %s;\n        nFaces %d;\n        startFace %d;\n    }\n"
                    % (name, name, n, start))
            start += n
        f.write(")\n\n// ************************************************************************* //\n")
    return directory.rstrip("/") + "/"
