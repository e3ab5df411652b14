"""Python mirror of the FETI-DP solver interface (over the C ABI).

Names follow the reference's solver contract: ``SolverStatus`` and
``SolverResult{x, status, iterations, residual_norm, solve_time_ms, error_msg}``
(include/numeric/em_linear_solver.h:26-61); ``get_solver_name`` / ``clear``
(:200-235).  Errors are reported through ``status`` / ``error_msg``, never
raised from ``solve`` (src/numeric/unified_direct_solver.cpp:122-136).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import enum

import numpy as np

from . import _lib


class SolverStatus(enum.IntEnum):
    SUCCESS = 0
    NUMERICAL_ERROR = 1
    DIVERGED = 2
    MAX_ITER_REACHED = 3
    INVALID_INPUT = 4


@dataclasses.dataclass
class SolverResult:
    x: np.ndarray
    status: SolverStatus
    iterations: int = 0
    residual_norm: float = 0.0
    solve_time_ms: float = 0.0
    error_msg: str = ""
    refinements: int = 0
    dual_residual: float = 0.0
    initial_dual_residual: float = 0.0
    factor_time_ms: float = 0.0
    symbolic_time_ms: float = 0.0
    fell_back_to_fp64: bool = False
    guess_time_ms: float = 0.0
    learned_guess: bool = False
    precond_fp64: bool = False
    deluxe: bool = False


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class FETIDPSolver:
    """MI355X FETI-DP solver for curl(nu curl A) + kappa A = J on Nedelec tets."""

    def __init__(self, precision="fp64", tolerance=1e-8, dual_tolerance=1e-11,
                 max_iterations=1000, max_refinements=12, profile=False,
                 inner_tolerance=1e-2, graphs=True, precond_precision="auto", scaling="auto",
                 coarse_solve="auto", coarse_group=0):
        self._lib = _lib.load()
        cfg = _lib.Config()
        self._lib.fetidp_config_default(C.byref(cfg))
        cfg.precision = {"fp64": _lib.FP64, "fp32": _lib.FP32_REFINE}[precision]
        cfg.tolerance = tolerance
        cfg.dual_tolerance = dual_tolerance
        cfg.max_iterations = max_iterations
        cfg.max_refinements = max_refinements
        cfg.verbose = 2 if profile else 0
        cfg.inner_tolerance = inner_tolerance
        cfg.graphs = 1 if graphs else 0
        cfg.precond_precision = {"auto": 0, "fp32": 1, "fp64": 2}[precond_precision]
        cfg.scaling = {"auto": 0, "stiffness": 1, "deluxe": 2}[scaling]
        cfg.coarse_solve = {"auto": 0, "dense": 1, "grouped": 2}[coarse_solve]
        cfg.coarse_group = int(coarse_group)
        self.precision = precision
        self._h = self._lib.fetidp_create(C.byref(cfg))
        if not self._h:
            raise RuntimeError("fetidp_create failed")
        self._keep = []
        self.sizes = None
        self._model = None

    # --------------------------------------------------------------- configuration
    def get_config(self):
        cfg = _lib.Config()
        self._check(self._lib.fetidp_get_config(self._h, C.byref(cfg)), "fetidp_get_config")
        return cfg

    def autotune(self, repetitions=2):
        """Time the launch variants (where the diagonal factorizations run, graph replay,
        tile-kernel waves) and, for a grouped coarse solve, the coarse group widths on the
        current problem, keep the fastest; returns [(side_stream, graphs, waves,
        coarse_group, ms)] (coarse_group 0: dense coarse inverse)."""
        buf = np.zeros(5 * 32)
        n = self._lib.fetidp_autotune(self._h, repetitions, _dp(buf), buf.size)
        if n < 0:
            raise RuntimeError(f"fetidp_autotune: {SolverStatus(-n).name}: {self.last_error()}")
        rows = buf[:5 * n].reshape(n, 5)
        return [(int(a), int(b), int(c), int(g), float(d)) for a, b, c, g, d in rows]

    def set_config_values(self, **fields):
        """Change configuration fields (names of fetidp_config) on the live handle."""
        cfg = self.get_config()
        for k, v in fields.items():
            setattr(cfg, k, v)
        self._check(self._lib.fetidp_set_config(self._h, C.byref(cfg)), "fetidp_set_config")

    def set_profile(self, on: bool):
        """Switch per-kernel-class event timing (disables graph replay while on)."""
        cfg = self.get_config()
        cfg.verbose = (cfg.verbose | 2) if on else (cfg.verbose & ~2)
        self._check(self._lib.fetidp_set_config(self._h, C.byref(cfg)), "fetidp_set_config")

    # --------------------------------------------------------------- plumbing
    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {SolverStatus(rc).name}: {self.last_error()}")

    def last_error(self) -> str:
        return self._lib.fetidp_last_error(self._h).decode()

    def get_solver_name(self) -> str:
        return f"FETIDP_MI355X_{self.precision.upper()}"

    def clear(self):
        if self._h:
            self._lib.fetidp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.clear()
        except Exception:
            pass

    # --------------------------------------------------------------- setup
    def setup(self, xyz, tet2node, tet_material, nu, kappa, tet_subdomain, pec_pairs):
        xyz, t2n, mat = _f64(xyz), _i32(tet2node), _i32(tet_material)
        nu, kappa, sub = _f64(nu), _f64(kappa), _i32(tet_subdomain)
        pec = _i32(pec_pairs).reshape(-1, 2)
        rc = self._lib.fetidp_setup(self._h, xyz.shape[0], _dp(xyz), t2n.shape[0], _ip(t2n),
                                    _ip(mat), nu.size, _dp(nu), _dp(kappa), _ip(sub),
                                    pec.shape[0], _ip(pec))
        self._check(rc, "fetidp_setup")
        out = np.zeros(10, dtype=np.int64)
        self._check(self._lib.fetidp_get_sizes(self._h, out.ctypes.data_as(C.POINTER(C.c_int64))),
                    "fetidp_get_sizes")
        keys = ["n_edges", "n_free", "n_lambda", "n_primal", "n_sub", "n_tets", "pool_elems",
                "op_elems", "coarse_ncp", "max_local"]
        self.sizes = dict(zip(keys, (int(v) for v in out)))
        return self

    def setup_problem(self, prob):
        """Symbolic setup from a workloads.Problem; also sets its J source."""
        self.setup(prob.xyz, prob.tet2node, prob.tet_material, prob.nu, prob.kappa,
                   prob.tet_subdomain, prob.pec_pairs)
        return self

    def choices(self):
        """Preconditioner choices resolved from the configuration and the setup (host
        only): {"deluxe": bool, "precond_fp64": bool, "box_overlap": float}."""
        dl, p64, ov = C.c_int32(), C.c_int32(), C.c_double()
        self._check(self._lib.fetidp_get_choices(self._h, C.byref(dl), C.byref(p64), C.byref(ov)),
                    "fetidp_get_choices")
        return {"deluxe": bool(dl.value), "precond_fp64": bool(p64.value),
                "box_overlap": float(ov.value)}

    def set_materials(self, nu, kappa):
        nu, kappa = _f64(nu), _f64(kappa)
        self._check(self._lib.fetidp_set_materials(self._h, nu.size, _dp(nu), _dp(kappa)),
                    "fetidp_set_materials")

    def set_rhs_tet_J(self, jq):
        jq = _f64(jq)
        self._check(self._lib.fetidp_set_rhs_tet_J(self._h, _dp(jq)), "fetidp_set_rhs_tet_J")

    def set_rhs_free(self, f):
        f = _f64(f)
        self._check(self._lib.fetidp_set_rhs_free(self._h, _dp(f)), "fetidp_set_rhs_free")

    def factor(self):
        self._check(self._lib.fetidp_factor(self._h), "fetidp_factor")

    # --------------------------------------------------------------- solve
    def solve(self, lambda0=None, copy_x=True) -> SolverResult:
        """PCPG solve.  copy_x=False leaves the solution in device memory (x is None):
        no PCIe transfer of the n_free-long solution."""
        n = self.sizes["n_free"]
        x = np.zeros(n) if copy_x else None
        st = _lib.Stats()
        lam0 = None if lambda0 is None else _f64(lambda0)
        rc = self._lib.fetidp_solve(self._h, _dp(lam0) if lam0 is not None else None,
                                    _dp(x) if x is not None else None, C.byref(st))
        return SolverResult(x=x, status=SolverStatus(rc), iterations=st.iterations,
                            residual_norm=st.residual_norm, solve_time_ms=st.solve_ms,
                            error_msg="" if rc == 0 else self.last_error(),
                            refinements=st.refinements, dual_residual=st.dual_residual,
                            initial_dual_residual=st.initial_dual_residual,
                            factor_time_ms=st.factor_ms, symbolic_time_ms=st.symbolic_ms,
                            fell_back_to_fp64=bool(st.fell_back), guess_time_ms=st.guess_ms,
                            learned_guess=bool(st.learned_guess),
                            precond_fp64=bool(st.precond_fp64), deluxe=bool(st.deluxe))

    # --------------------------------------------------------------- learned lambda0
    def set_lambda0_model(self, model):
        """Attach a LambdaMPNN (or None): solve() then starts PCPG from its lambda0."""
        self._check(self._lib.fetidp_set_lambda0_model(self._h, model.handle if model else None),
                    "fetidp_set_lambda0_model")
        self._model = model  # keep the model alive while attached

    def dual_graph(self):
        """(row_ptr, col, coef) of the multiplier interface graph."""
        nnz = self._lib.fetidp_get_dual_graph(self._h, None, None, None)
        if nnz < 0:
            raise RuntimeError(f"fetidp_get_dual_graph: {self.last_error()}")
        rp = np.zeros(self.sizes["n_lambda"] + 1, dtype=np.int32)
        col = np.zeros(max(nnz, 1), dtype=np.int32)
        coef = np.zeros(max(nnz, 1), dtype=np.float32)
        self._lib.fetidp_get_dual_graph(self._h, _ip(rp), _ip(col),
                                        coef.ctypes.data_as(C.POINTER(C.c_float)))
        return rp, col[:nnz], coef[:nnz]

    def lambda_features(self):
        """(features (n_lambda, 16) float32, (rms d, rms z)) for the current factor and RHS."""
        X = np.zeros((self.sizes["n_lambda"], 16), dtype=np.float32)
        sc = np.zeros(2)
        self._check(self._lib.fetidp_get_lambda_features(
            self._h, X.ctypes.data_as(C.POINTER(C.c_float)), _dp(sc)), "fetidp_get_lambda_features")
        return X, sc

    # --------------------------------------------------------------- introspection
    def topology(self):
        s = self.sizes
        tf = np.zeros(6 * s["n_tets"], dtype=np.int32)
        en = np.zeros(2 * s["n_edges"], dtype=np.int32)
        lf = np.zeros(max(s["n_lambda"], 1), dtype=np.int32)
        pf = np.zeros(max(s["n_primal"], 1), dtype=np.int32)
        self._check(self._lib.fetidp_get_topology(self._h, _ip(tf), _ip(en), _ip(lf), _ip(pf)),
                    "fetidp_get_topology")
        return dict(tet_free=tf.reshape(-1, 6), edge_nodes=en.reshape(-1, 2),
                    lam_free=lf[:s["n_lambda"]], primal_free=pf[:s["n_primal"]])

    def global_system(self):
        n = self.sizes["n_free"]
        nnz = int(self._lib.fetidp_get_global_nnz(self._h))
        rp = np.zeros(n + 1, dtype=np.int32)
        col = np.zeros(nnz, dtype=np.int32)
        val = np.zeros(nnz)
        f = np.zeros(n)
        self._check(self._lib.fetidp_get_global_system(self._h, _ip(rp), _ip(col), _dp(val),
                                                       _dp(f)), "fetidp_get_global_system")
        return rp, col, val, f

    def apply_dual_operator(self, p):
        p = _f64(p)
        q = np.zeros_like(p)
        self._check(self._lib.fetidp_apply_dual_operator(self._h, _dp(p), _dp(q)), "apply_F")
        return q

    def apply_preconditioner(self, r):
        r = _f64(r)
        z = np.zeros_like(r)
        self._check(self._lib.fetidp_apply_preconditioner(self._h, _dp(r), _dp(z)), "apply_M")
        return z

    def dual_rhs(self):
        d = np.zeros(self.sizes["n_lambda"])
        self._check(self._lib.fetidp_get_dual_rhs(self._h, _dp(d)), "get_dual_rhs")
        return d

    def get_lambda(self):
        lam = np.zeros(self.sizes["n_lambda"])
        self._check(self._lib.fetidp_get_lambda(self._h, _dp(lam)), "get_lambda")
        return lam

    def history(self):
        n = self._lib.fetidp_get_history(self._h, None, 0)
        h = np.zeros(max(n, 1))
        self._lib.fetidp_get_history(self._h, _dp(h), n)
        return h[:n]

    PROFILE_SLOTS = ["integrate", "pool_assembly", "tile_schur", "tile_column", "tile_diag",
                     "extract", "coarse_schur", "coarse_column", "coarse_diag", "dual_gemv",
                     "dual_rest", "precond_gemv", "forward_trsv", "interior_trsv"]

    def profile(self):
        """{kernel class: (ms, launches)} plus tile flops, from event timing."""
        buf = np.zeros(64)
        n = self._lib.fetidp_get_profile(self._h, _dp(buf), 64)
        out = {name: (float(buf[2 * i]), int(buf[2 * i + 1]))
               for i, name in enumerate(self.PROFILE_SLOTS)}
        names = ["tile_schur", "tile_column", "tile_diag", "coarse_schur", "coarse_column",
                 "coarse_diag"]
        fl = buf[32:38] if n >= 38 else np.zeros(6)
        uf = buf[38:44] if n >= 44 else np.zeros(6)
        out["flops"] = dict(zip(names, (float(v) for v in fl)))
        out["useful_flops"] = dict(zip(names, (float(v) for v in uf)))
        return out

    def reset_profile(self):
        self._lib.fetidp_reset_profile(self._h)

    def refinement_history(self):
        n = self._lib.fetidp_get_refinement_history(self._h, None, 0)
        h = np.zeros(max(n, 1))
        self._lib.fetidp_get_refinement_history(self._h, _dp(h), n)
        return h[:n]

    def coarse_matrix(self):
        nc = self.sizes["n_primal"]
        S = np.zeros((nc, nc))
        self._check(self._lib.fetidp_get_coarse_matrix(self._h, _dp(S)), "get_coarse_matrix")
        return S


def integrate_elements(xyz, tet2node, tet_material, nu, kappa, jq=None, device="cuda"):
    """Batched Nedelec element kernel on device tensors; returns (Ke, Me, Fe) as
    (n_tets, 21) / (n_tets, 6) float64 torch tensors on `device`."""
    import torch
    lib = _lib.load()
    dev = torch.device(device)
    t_xyz = torch.as_tensor(_f64(xyz), device=dev)
    t_t2n = torch.as_tensor(_i32(tet2node), device=dev)
    t_mat = torch.as_tensor(_i32(tet_material), device=dev)
    t_nu = torch.as_tensor(_f64(nu), device=dev)
    t_ka = torch.as_tensor(_f64(kappa), device=dev)
    nt = t_t2n.shape[0]
    Ke = torch.empty((21, nt), dtype=torch.float64, device=dev)
    Me = torch.empty((21, nt), dtype=torch.float64, device=dev)
    Fe = torch.empty((6, nt), dtype=torch.float64, device=dev)
    t_jq = torch.as_tensor(_f64(jq), device=dev) if jq is not None else None
    stream = torch.cuda.current_stream(dev).cuda_stream
    rc = lib.emag_nedelec_integrate(nt, t_xyz.data_ptr(), t_t2n.data_ptr(), t_mat.data_ptr(),
                                    t_nu.data_ptr(), t_ka.data_ptr(),
                                    t_jq.data_ptr() if t_jq is not None else None,
                                    Ke.data_ptr(), Me.data_ptr(), Fe.data_ptr(), stream)
    if rc != 0:
        raise RuntimeError(f"emag_nedelec_integrate failed: {rc}")
    torch.cuda.synchronize(dev)
    return Ke.T, Me.T, (Fe.T if jq is not None else None)
