"""CEMAS partitions: slot maps of a vertex partition and the partitioner
(host functions of the C ABI; no device work).

``mas_partition_maps`` is ``set_mas_partition``
(app/common/sim_bootstrap.cpp:90-113); ``partition_vertices`` replaces the
METIS k-way step of MeshProcess/metis_partition (metis_sort.cpp:71-110).  A
scene carries its partition as ``part_id`` (per vertex); the linear system
then builds its MAS over the partition's slots.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib


def mas_partition_maps(part_id, bank: int = 16):
    """(partId_map_real, real_map_partId): the vertex of each slot (-1 empty)
    and the slot of each vertex; part p owns slots [bank p, bank p + bank)."""
    part = np.ascontiguousarray(part_id, dtype=np.int32)
    n = part.size
    size = ctypes.c_int(0)
    check(lib().sgipc_mas_partition_maps(part.ctypes.data, n, bank, ctypes.byref(size), None, None))
    p2r = np.empty(size.value, dtype=np.int32)
    r2p = np.empty(n, dtype=np.int32)
    check(lib().sgipc_mas_partition_maps(part.ctypes.data, n, bank, ctypes.byref(size), p2r.ctypes.data,
                                         r2p.ctypes.data))
    return p2r, r2p


def partition_vertices(nbr_start, nbr_num, nbr_list, bank: int = 16) -> np.ndarray:
    """Parts of 1..``bank`` vertices of the neighbour graph (symmetric CSR)
    by recursive multilevel bisection; part ids in bisection-tree order."""
    st = np.ascontiguousarray(nbr_start, dtype=np.uint32)
    nn = np.ascontiguousarray(nbr_num, dtype=np.uint32)
    nl = np.ascontiguousarray(nbr_list if np.size(nbr_list) else np.zeros(1), dtype=np.uint32)
    n = st.size
    part = np.empty(n, dtype=np.int32)
    npart = ctypes.c_int(0)
    check(lib().sgipc_partition_vertices(n, st.ctypes.data, nn.ctypes.data, nl.ctypes.data, bank, part.ctypes.data,
                                         ctypes.byref(npart)))
    return part


def edge_cut(part, nbr_start, nbr_num, nbr_list) -> int:
    """Undirected neighbour edges whose ends lie in different parts."""
    part = np.asarray(part)
    src = np.repeat(np.arange(part.size), np.asarray(nbr_num, dtype=np.int64))
    return int((part[src] != part[np.asarray(nbr_list, dtype=np.int64)]).sum() // 2)
