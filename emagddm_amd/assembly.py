"""GPU global Nedelec assembly and FP64 residual operator (include/emagddm_assembly.h).

``GlobalAssembly`` mirrors EMAssembly::assemble and its getters
(include/numeric/em_assembly.hpp:236-326): the symbolic part (edge ids, free numbering,
CSR pattern) runs on the host when the object is made; ``assemble`` fills the stiffness,
mass, system (K + M) values and the source vector on the GPU.  ``csr_residual`` is the
FP64 residual ||f - A x|| / ||f|| of a device CSR.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class GlobalAssembly:
    def __init__(self, n_nodes: int, tet2node, pec_pairs):
        self._lib = _lib.load()
        t2n = np.ascontiguousarray(tet2node, dtype=np.int32)
        pec = np.ascontiguousarray(pec_pairs, dtype=np.int32).reshape(-1, 2)
        self.n_tets = t2n.shape[0]
        self._h = self._lib.emag_assembly_create(
            int(n_nodes), self.n_tets, t2n.ctypes.data_as(C.POINTER(C.c_int32)), pec.shape[0],
            pec.ctypes.data_as(C.POINTER(C.c_int32)))
        if not self._h:
            raise RuntimeError("emag_assembly_create failed (bad mesh or PEC input)")
        nf, nnz = C.c_int64(), C.c_int64()
        self._lib.emag_assembly_sizes(self._h, C.byref(nf), C.byref(nnz))
        self.n_free, self.nnz = nf.value, nnz.value

    def pattern(self):
        """(row_ptr, col, tet_free (n_tets, 6), tet_sign (n_tets, 6)) host arrays."""
        rp = np.zeros(self.n_free + 1, np.int32)
        col = np.zeros(max(self.nnz, 1), np.int32)
        tf = np.zeros(6 * self.n_tets, np.int32)
        ts = np.zeros(6 * self.n_tets, np.int8)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        self._lib.emag_assembly_pattern(self._h, ip(rp), ip(col), ip(tf),
                                        ts.ctypes.data_as(C.POINTER(C.c_int8)))
        return rp, col[:self.nnz], tf.reshape(-1, 6), ts.reshape(-1, 6)

    def assemble(self, xyz, tet_material, nu, kappa, jq=None, outputs=("A", "f")):
        """Device (torch, cuda) tensors for the requested outputs among K, M, A, f."""
        import torch
        dev = torch.device("cuda")
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a, dtype=dt), device=dev)
        x, m = t(xyz, np.float64), t(tet_material, np.int32)
        n_, k_ = t(nu, np.float64), t(kappa, np.float64)
        j = t(jq, np.float64) if jq is not None else None
        out = {k: torch.empty(self.nnz if k != "f" else self.n_free, dtype=torch.float64,
                              device=dev) for k in outputs}
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = self._lib.emag_assemble(self._h, _ptr(x), _ptr(m), _ptr(n_), _ptr(k_), _ptr(j),
                                     _ptr(out.get("K")), _ptr(out.get("M")), _ptr(out.get("A")),
                                     _ptr(out.get("f")), stream)
        if rc != 0:
            raise RuntimeError(f"emag_assemble: {self._lib.emag_assembly_last_error(self._h).decode()}")
        torch.cuda.current_stream().synchronize()
        return out

    def close(self):
        if self._h:
            self._lib.emag_assembly_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def csr_residual(row_ptr, col, val, x, f, want_r=False, no_alloc=False):
    """||f - A x|| / ||f|| of device tensors (int32 row_ptr / col, float64 val, x, f);
    with want_r also the residual vector.  no_alloc uses the entry point that allocates
    nothing (emag_csr_residual_noalloc)."""
    import torch
    lib = _lib.load()
    n = row_ptr.numel() - 1
    r = torch.empty(n, dtype=torch.float64, device=f.device) if want_r else None
    rel = C.c_double()
    fn = lib.emag_csr_residual_noalloc if no_alloc else lib.emag_csr_residual
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(n, _ptr(row_ptr), _ptr(col), _ptr(val), _ptr(x), _ptr(f), _ptr(r), C.byref(rel), stream)
    if rc != 0:
        raise RuntimeError(f"csr_residual: status {rc}")
    return (rel.value, r) if want_r else rel.value


def tree_gauge_pairs(n_nodes: int, tet2node, pec_pairs) -> np.ndarray:
    """Tree edges of the tree-cotree gauge (emag_tree_gauge) as (n_tree, 2) node pairs;
    pass them with the PEC pairs to fix A = 0 on them for kappa = 0 problems."""
    lib = _lib.load()
    t2n = np.ascontiguousarray(tet2node, dtype=np.int32)
    pec = np.ascontiguousarray(pec_pairs, dtype=np.int32).reshape(-1, 2)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    n = lib.emag_tree_gauge(int(n_nodes), t2n.shape[0], ip(t2n), pec.shape[0], ip(pec), None)
    if n < 0:
        raise RuntimeError("emag_tree_gauge: bad mesh or PEC input")
    out = np.zeros((max(n, 1), 2), dtype=np.int32)
    lib.emag_tree_gauge(int(n_nodes), t2n.shape[0], ip(t2n), pec.shape[0], ip(pec), ip(out))
    return out[:n]
