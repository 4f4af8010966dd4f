"""Writes tests/golden/sorted16_meshes.npz from the reference's METIS-sorted
assets (Assets/sorted_mesh/*_sorted.16.{msh,obj,part}): the vertex
positions, elements and per-vertex part ids of

* cube  (gmsh tets, 8 vertices, one part),
* cloth (OBJ triangles, 4225 vertices, 282 parts),
* bunny (gmsh tets, 19193 vertices, 1280 parts),

as plain arrays (data only; the part files are METIS k-way outputs that this
build cannot recompute, which is why they are kept).  Run in a checkout that
has the reference next to it:
    python tests/golden/make_sorted16_fixtures.py [/path/to/reference]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main(ref="/root/reference"):
    import importlib
    mio = importlib.import_module("stiff-gipc_amd.mesh_io")
    d = os.path.join(ref, "Assets", "sorted_mesh")
    out = {}
    for name, fn, kind in (("cube", "cube_sorted.16", "msh"), ("cloth", "cloth_high_sorted.16", "obj"),
                           ("bunny", "bunny2_sorted.16", "msh")):
        if kind == "msh":
            x, el = mio.read_msh(os.path.join(d, fn + ".msh"))
        else:
            x, el = mio.read_obj(os.path.join(d, fn + ".obj"))
        part = mio.read_part(os.path.join(d, fn + ".part"), x.shape[0])
        out[f"{name}_x"] = x
        out[f"{name}_el"] = el.astype(np.uint16 if x.shape[0] < 65536 else np.int32)
        out[f"{name}_part"] = part.astype(np.uint16)
    np.savez_compressed(os.path.join(HERE, "sorted16_meshes.npz"), **out)


if __name__ == "__main__":
    main(*sys.argv[1:])
