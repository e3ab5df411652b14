"""On-disk problem files (include/emagddm_problem.h) through the C ABI.

``write_problem`` stores a ``workloads.Problem`` (mesh, materials, partition, PEC
edges, J at the quadrature points); ``read_problem`` loads and validates one.  The
reference's reader interface has no implementation (fe_em::MeshReader::read,
include/fe_em/mesh_reader.hpp:32-56); this is the format a reader for the FETI-DP
path consumes.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from .workloads import Problem


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def write_problem(path: str, p: Problem) -> None:
    lib = _lib.load()
    xyz = np.ascontiguousarray(p.xyz, dtype=np.float64)
    t2n = np.ascontiguousarray(p.tet2node, dtype=np.int32)
    mat = np.ascontiguousarray(p.tet_material, dtype=np.int32)
    sub = np.ascontiguousarray(p.tet_subdomain, dtype=np.int32)
    nu = np.ascontiguousarray(p.nu, dtype=np.float64)
    ka = np.ascontiguousarray(p.kappa, dtype=np.float64)
    pec = np.ascontiguousarray(p.pec_pairs, dtype=np.int32).reshape(-1, 2)
    jq = None if p.jq is None else np.ascontiguousarray(p.jq, dtype=np.float64)
    s = _lib.Problem(n_nodes=xyz.shape[0], n_tets=t2n.shape[0], n_pec=pec.shape[0],
                     n_mat=nu.size, n_sub=int(sub.max()) + 1,
                     xyz=_ptr(xyz, C.c_double), tet2node=_ptr(t2n, C.c_int32),
                     tet_material=_ptr(mat, C.c_int32), tet_subdomain=_ptr(sub, C.c_int32),
                     nu=_ptr(nu, C.c_double), kappa=_ptr(ka, C.c_double),
                     pec_edge_nodes=_ptr(pec, C.c_int32),
                     jq=None if jq is None else _ptr(jq, C.c_double))
    if lib.emag_problem_write(os.fsencode(path), C.byref(s)) != 0:
        raise RuntimeError(lib.emag_problem_last_error().decode())


def read_problem(path: str, name: str | None = None) -> Problem:
    lib = _lib.load()
    s = _lib.Problem()
    if lib.emag_problem_read(os.fsencode(path), C.byref(s)) != 0:
        raise RuntimeError(lib.emag_problem_last_error().decode())
    try:
        nn, nt, npc, nm = s.n_nodes, s.n_tets, s.n_pec, s.n_mat
        arr = lambda ptr, n: np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else \
            np.zeros(0, dtype=np.ctypeslib.as_array(ptr, shape=(1,)).dtype)
        jq = arr(s.jq, 12 * nt).reshape(nt, 4, 3) if s.jq else None
        return Problem(name=name or os.path.splitext(os.path.basename(path))[0], n=0,
                       xyz=arr(s.xyz, 3 * nn).reshape(nn, 3),
                       tet2node=arr(s.tet2node, 4 * nt).reshape(nt, 4),
                       tet_material=arr(s.tet_material, nt), nu=arr(s.nu, nm),
                       kappa=arr(s.kappa, nm), tet_subdomain=arr(s.tet_subdomain, nt),
                       n_sub=int(s.n_sub),
                       pec_pairs=(arr(s.pec_edge_nodes, 2 * npc).reshape(npc, 2) if npc
                                  else np.zeros((0, 2), np.int32)),
                       jq=jq)
    finally:
        lib.emag_problem_free(C.byref(s))
