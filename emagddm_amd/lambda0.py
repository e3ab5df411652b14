"""Learned initial multiplier lambda0 for FETI-DP: model binding, data, training, A/B.

The network (include/emagddm_mpnn.h) runs entirely on the GPU through the C ABI;
this module only orchestrates: it draws a family of problems on one fixed mesh
and partition (materials and current loops vary with the seed), solves each with
the FP64 solver to collect (features, lambda*) pairs, trains the network with
the library's Adam step, and compares PCPG from zero with PCPG from the learned
lambda0 on held-out seeds.  The reference starts its Krylov solvers from zero
(src/numeric/iterative_solvers.cpp:424-503); that stays the default.

    python -m emagddm_amd.lambda0 --workload cube32l --train 2048 --heldout 256 \\
        --epochs 30 --out emagddm_amd/weights/lambda0_cube32l.npy
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib, workloads
from .fetidp import FETIDPSolver, SolverStatus

N_FEATURES = 16
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class LambdaMPNN:
    """Handle to one network on the GPU (bf16 MFMA operands, FP32 master weights)."""

    def __init__(self, seed: int = 0, params=None):
        self._lib = _lib.load()
        self._h = self._lib.emag_mpnn_create(seed)
        if not self._h:
            raise RuntimeError("emag_mpnn_create failed")
        if params is not None:
            self.set_params(params)

    @property
    def handle(self):
        return self._h

    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.emag_mpnn_last_error(self._h).decode()
            raise RuntimeError(f"{what}: status {rc}: {msg}")

    @staticmethod
    def num_params() -> int:
        return int(_lib.load().emag_mpnn_num_params())

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), dtype=np.float32)
        self._check(self._lib.emag_mpnn_get_params(self._h, _fp(out)), "get_params")
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != self.num_params():
            raise ValueError(f"expected {self.num_params()} parameters, got {p.size}")
        self._check(self._lib.emag_mpnn_set_params(self._h, _fp(p)), "set_params")

    def save(self, path):
        np.save(path, self.get_params())

    @classmethod
    def load(cls, path):
        return cls(params=np.load(path, allow_pickle=False))

    def set_graph(self, row_ptr, col, coef):
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        cl = np.ascontiguousarray(col, dtype=np.int32)
        cf = np.ascontiguousarray(coef, dtype=np.float32)
        self.n_nodes = rp.size - 1
        self._check(self._lib.emag_mpnn_set_graph(
            self._h, self.n_nodes, rp.ctypes.data_as(C.POINTER(C.c_int32)),
            cl.ctypes.data_as(C.POINTER(C.c_int32)), _fp(cf)), "set_graph")

    @staticmethod
    def _stream():
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def forward(self, x):
        """x: float32 CUDA tensor (batch*n, 16) -> y (batch*n,)."""
        import torch
        x = x.contiguous()
        batch = x.shape[0] // self.n_nodes
        y = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        self._check(self._lib.emag_mpnn_forward(self._h, batch, x.data_ptr(), y.data_ptr(),
                                                self._stream()), "forward")
        return y

    def gradient(self, x, target, omega):
        """(loss, gradient) of L = (1/B) sum_u omega_b (y_u - target_u)^2."""
        x, target, omega = x.contiguous(), target.contiguous(), omega.contiguous()
        g = np.zeros(self.num_params(), dtype=np.float32)
        loss = C.c_double()
        self._check(self._lib.emag_mpnn_gradient(self._h, omega.numel(), x.data_ptr(),
                                                 target.data_ptr(), omega.data_ptr(), _fp(g),
                                                 C.byref(loss), self._stream()), "gradient")
        return loss.value, g

    def train_step(self, x, target, omega, lr):
        x, target, omega = x.contiguous(), target.contiguous(), omega.contiguous()
        loss = C.c_double()
        self._check(self._lib.emag_mpnn_train_step(self._h, omega.numel(), x.data_ptr(),
                                                   target.data_ptr(), omega.data_ptr(),
                                                   C.c_float(lr), C.byref(loss), self._stream()),
                    "train_step")
        return loss.value

    def close(self):
        if self._h:
            self._lib.emag_mpnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# workloads served by a network trained on a smaller family of the same kind (the
# network is local, so its weights do not depend on the mesh size)
TRAINED_ON = {"cube96j": "cube48c"}


def default_weights(workload: str):
    path = os.path.join(WEIGHTS_DIR, f"lambda0_{TRAINED_ON.get(workload, workload)}.npy")
    return path if os.path.exists(path) else None


# ----------------------------------------------------------------------------- data
def loop_source_torch(points, loops):
    """workloads.loop_source evaluated with torch on the points' device."""
    import torch
    out = torch.zeros_like(points)
    for lp in loops:
        c = torch.as_tensor(lp.center, dtype=points.dtype, device=points.device)
        nrm = torch.as_tensor(lp.normal, dtype=points.dtype, device=points.device)
        d = points - c
        z = d @ nrm
        rv = d - z[..., None] * nrm
        rho = torch.linalg.norm(rv, dim=-1)
        ephi = torch.cross(nrm.expand_as(rv), rv, dim=-1)
        safe = torch.where(rho > 1e-300, rho, torch.ones_like(rho))
        ephi = ephi / safe[..., None]
        amp = lp.current * torch.exp(-((rho - lp.radius) ** 2 + z ** 2) / (2 * lp.width ** 2))
        out = out + amp[..., None] * ephi
    return out


class Family:
    """Problems of one named workload on the mesh/partition/material map of seed 0."""

    def __init__(self, name: str, precision: str = "fp64", device: str = "cuda"):
        import torch
        self.name = name
        self.prob = workloads.make_workload(name, 0)
        self.solver = FETIDPSolver(precision=precision)
        self.solver.setup_problem(self.prob)
        pts = workloads.quadrature_points(self.prob.xyz, self.prob.tet2node)
        self.points = torch.as_tensor(pts, device=device)
        self._mat_key = None

    def load(self, seed: int):
        """Set materials and source of `seed`; refactor only when materials change."""
        nu, kappa, loops = workloads.family_parameters(self.name, seed)
        key = (tuple(np.asarray(nu).tolist()), tuple(np.asarray(kappa).tolist()))
        jq = loop_source_torch(self.points, loops).cpu().numpy()
        if key != self._mat_key:
            self.solver.set_materials(nu, kappa)
            self.solver.set_rhs_tet_J(jq)
            self.solver.factor()
            self._mat_key = key
        else:
            self.solver.set_rhs_tet_J(jq)


def build_dataset(fam: Family, seeds, log=None):
    """Solve every seed with lambda0 = 0; return features, targets and weights on the GPU."""
    import torch
    nl = fam.solver.sizes["n_lambda"]
    S = len(seeds)
    X = torch.empty((S, nl, N_FEATURES), dtype=torch.float32, device="cuda")
    T = torch.empty((S, nl), dtype=torch.float32, device="cuda")
    Om = torch.empty((S,), dtype=torch.float32, device="cuda")
    its = []
    fam.solver.set_lambda0_model(None)
    t0 = time.time()
    for k, seed in enumerate(seeds):
        fam.load(seed)
        feats, sc = fam.solver.lambda_features()
        r = fam.solver.solve()
        if r.status != SolverStatus.SUCCESS:
            raise RuntimeError(f"seed {seed}: {r.status.name} {r.error_msg}")
        lam = fam.solver.get_lambda()
        sz = float(sc[1])
        X[k] = torch.from_numpy(feats).cuda()
        T[k] = torch.from_numpy((lam / sz).astype(np.float32)).cuda()
        Om[k] = float(sz * sz / max(float(lam @ lam), 1e-300))
        its.append(r.iterations)
        if log and (k + 1) % 64 == 0:
            log(f"[lambda0] data {k + 1}/{S} seeds, {time.time() - t0:.1f} s, "
                f"mean PCPG its {np.mean(its):.2f}")
    return dict(X=X, T=T, Om=Om, its=np.array(its))


def rel_error(model, data):
    """Per-sample relative L2 error ||s_z y - lambda*|| / ||lambda*||."""
    import torch
    S, nl, _ = data["X"].shape
    errs = []
    for k in range(S):
        y = model.forward(data["X"][k].reshape(nl, N_FEATURES))
        e = torch.sum((y - data["T"][k]) ** 2) * data["Om"][k]
        errs.append(float(torch.sqrt(e)))
    return np.array(errs)


def train(model: LambdaMPNN, data, epochs=30, batch=8, lr=1e-3, seed=0, log=None):
    """Adam with cosine learning-rate decay over shuffled mini-batches."""
    S = data["X"].shape[0]
    nl = data["X"].shape[1]
    rng = np.random.default_rng(seed)
    steps_per_epoch = max(1, S // batch)
    total = epochs * steps_per_epoch
    step = 0
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(S)
        losses = []
        for b in range(steps_per_epoch):
            idx = perm[b * batch:(b + 1) * batch]
            import torch
            ii = torch.as_tensor(idx, device="cuda")
            x = data["X"].index_select(0, ii).reshape(-1, N_FEATURES)
            t = data["T"].index_select(0, ii).reshape(-1)
            om = data["Om"].index_select(0, ii)
            cur = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
            losses.append(model.train_step(x, t, om, cur))
            step += 1
        hist.append(float(np.mean(losses)))
        if log:
            log(f"[lambda0] epoch {ep + 1}/{epochs} loss {hist[-1]:.4e} (rel err^2, batch mean)")
    del nl
    return hist


def evaluate(fam: Family, model: LambdaMPNN, seeds, log=None):
    """A/B on `seeds`: PCPG from zero vs from the learned lambda0 (same factor)."""
    rows = []
    for k, seed in enumerate(seeds):
        fam.load(seed)
        fam.solver.set_lambda0_model(None)
        r0 = fam.solver.solve()
        fam.solver.set_lambda0_model(model)
        r1 = fam.solver.solve()
        fam.solver.set_lambda0_model(None)
        x0n = float(np.linalg.norm(r0.x))
        rows.append(dict(seed=int(seed), its_zero=r0.iterations, its_learned=r1.iterations,
                         ms_zero=r0.solve_time_ms, ms_learned=r1.solve_time_ms,
                         guess_ms=r1.guess_time_ms, res_zero=r0.residual_norm,
                         res_learned=r1.residual_norm,
                         init_dual_zero=r0.initial_dual_residual,
                         init_dual_learned=r1.initial_dual_residual,
                         ok=bool(r0.status == SolverStatus.SUCCESS and
                                 r1.status == SolverStatus.SUCCESS),
                         x_rel_diff=float(np.linalg.norm(r1.x - r0.x) / max(x0n, 1e-300))))
        if log and (k + 1) % 32 == 0:
            log(f"[lambda0] eval {k + 1}/{len(seeds)}")
    return rows


def summarize(rows):
    a = lambda key: np.array([r[key] for r in rows], dtype=float)
    return dict(n=len(rows), all_ok=bool(all(r["ok"] for r in rows)),
                its_zero_mean=float(a("its_zero").mean()),
                its_learned_mean=float(a("its_learned").mean()),
                ms_zero_mean=float(a("ms_zero").mean()),
                ms_learned_mean=float(a("ms_learned").mean()),
                guess_ms_mean=float(a("guess_ms").mean()),
                init_dual_zero_mean=float(a("init_dual_zero").mean()),
                init_dual_learned_mean=float(a("init_dual_learned").mean()),
                max_res_learned=float(a("res_learned").max()),
                max_x_rel_diff=float(a("x_rel_diff").max()))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="cube32l")
    ap.add_argument("--train", type=int, default=2048)
    ap.add_argument("--heldout", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--out", default=None)
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    fam = Family(a.workload)
    log(f"[lambda0] {a.workload}: n_lambda={fam.solver.sizes['n_lambda']}")
    train_seeds = list(range(1, a.train + 1))
    held = list(range(1_000_000, 1_000_000 + a.heldout))
    data = build_dataset(fam, train_seeds, log)
    model = LambdaMPNN(seed=a.init_seed)
    rp, col, coef = fam.solver.dual_graph()
    model.set_graph(rp, col, coef)
    t0 = time.time()
    hist = train(model, data, a.epochs, a.batch, a.lr, log=log)
    train_s = time.time() - t0
    err_train = rel_error(model, data)
    del data
    rows = evaluate(fam, model, held, log)
    summ = summarize(rows)
    summ.update(workload=a.workload, n_train=a.train, epochs=a.epochs, batch=a.batch, lr=a.lr,
                train_seconds=round(train_s, 1), final_loss=hist[-1],
                train_rel_err_mean=float(err_train.mean()))
    log(json.dumps(summ))
    if a.out:
        model.save(a.out)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(dict(summary=summ, loss_history=hist, heldout=rows), f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
