"""Mesh and partition files of the reference's asset pipeline.

* gmsh 2.2 ASCII tetrahedral meshes: ``$Nodes`` (1-based id, x, y, z) and
  ``$Elements`` (id, type 4, 0 tags, four 1-based vertex ids), the layout
  ``TetMesh::export_mesh`` writes and ``load_tetrahedraMesh`` reads
  (MeshProcess/metis_partition/src/mesh.cpp:114-137,
  StiffGIPC/load_mesh.cpp);
* Wavefront OBJ triangle meshes (``v x y z`` / ``f a b c``, 1-based),
  ``TriMesh::export_mesh`` (mesh.cpp:139-156);
* ``.part`` files: one part id per vertex and line, in vertex order
  (``NodeEdgeModel::export_partition``, node_edge_model.cpp:29-37; read by
  ``tetrahedra_obj::load_parts``, StiffGIPC/load_mesh.cpp:447-478, which adds
  the running part offset of earlier meshes and rejects a file whose length
  differs from the vertex count);
* the "_sorted" relabelling of MeshProcess (metis_sort.cpp:180-210,
  mesh.cpp:65-88,159-200): vertices ordered by part id, element indices
  renamed.  The reference sorts with an unstable ``std::sort``; here the order
  inside a part is the original vertex order.
"""
from __future__ import annotations

import numpy as np


def read_msh(path: str):
    """(x (n, 3) float64, tets (m, 4) int32, 0-based) of a gmsh 2.2 ASCII file."""
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    x = tets = None
    while i < len(lines):
        tag = lines[i].strip()
        if tag == "$Nodes":
            n = int(lines[i + 1])
            a = np.array([ln.split() for ln in lines[i + 2:i + 2 + n]], dtype=np.float64)
            ids = a[:, 0].astype(np.int64)
            if not np.array_equal(ids, np.arange(1, n + 1)):
                raise ValueError(f"{path}: node ids are not 1..{n}")
            x = np.ascontiguousarray(a[:, 1:4])
            i += 2 + n
        elif tag == "$Elements":
            m = int(lines[i + 1])
            rows = [ln.split() for ln in lines[i + 2:i + 2 + m]]
            t = []
            for r in rows:
                ntag = int(r[2])
                if int(r[1]) == 4:
                    t.append([int(v) for v in r[3 + ntag:7 + ntag]])
            tets = np.asarray(t, dtype=np.int64) - 1
            i += 2 + m
        else:
            i += 1
    if x is None or tets is None:
        raise ValueError(f"{path}: no $Nodes / $Elements section")
    if tets.size and (tets.min() < 0 or tets.max() >= x.shape[0]):
        raise ValueError(f"{path}: element vertex out of range")
    return x, tets.astype(np.int32)


def read_obj(path: str):
    """(x (n, 3) float64, faces (m, 3) int32, 0-based) of an OBJ triangle mesh."""
    v, f = [], []
    with open(path) as fh:
        for ln in fh:
            p = ln.split()
            if not p:
                continue
            if p[0] == "v":
                v.append([float(c) for c in p[1:4]])
            elif p[0] == "f":
                if len(p) != 4:
                    raise ValueError(f"{path}: only triangles are supported")
                f.append([int(c.split("/")[0]) for c in p[1:4]])
    x = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    faces = np.asarray(f, dtype=np.int64).reshape(-1, 3) - 1
    if faces.size and (faces.min() < 0 or faces.max() >= x.shape[0]):
        raise ValueError(f"{path}: face vertex out of range")
    return x, faces.astype(np.int32)


def read_part(path: str, n_verts: int | None = None, part_offset: int = 0) -> np.ndarray:
    """Part id per vertex (+ part_offset); raises when the count differs from n_verts."""
    with open(path) as f:
        ids = np.array(f.read().split(), dtype=np.int64)
    if n_verts is not None and ids.size != n_verts:
        raise ValueError(f"{path}: {ids.size} part ids for {n_verts} vertices")
    if ids.size and ids.min() < 0:
        raise ValueError(f"{path}: negative part id")
    return (ids + part_offset).astype(np.int32)


def next_part_offset(part: np.ndarray) -> int:
    """The offset the next mesh's part ids start at (load_mesh.cpp:465): last id + 1."""
    return int(part[-1]) + 1 if part.size else 0


def write_part(path: str, part: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write("".join(f"{int(p)}\n" for p in part))


def sort_by_part(part: np.ndarray, *elements: np.ndarray):
    """Relabel vertices in part order: returns (order, sorted part, renamed
    elements...), order[new] = old vertex."""
    part = np.asarray(part)
    order = np.argsort(part, kind="stable")
    new_of_old = np.empty_like(order)
    new_of_old[order] = np.arange(order.size)
    out = [new_of_old[e].astype(np.int32) for e in elements]
    return (order, part[order].astype(np.int32), *out)
