"""GPU global assembly (include/emagddm_assembly.h) against the FP64 oracle, and the
FP64 residual operator.  The stiffness / mass / system split mirrors EMAssembly's
getStiffnessMatrix / getMassMatrix / getSourceVector
(include/numeric/em_assembly.hpp:236-326)."""
import numpy as np
import pytest
import scipy.sparse as sp

from emagddm_amd import FETIDPSolver, workloads as W
from emagddm_amd.assembly import GlobalAssembly, csr_residual
from oracle import assembly, solvers

pytestmark = pytest.mark.gpu


def _oracle(p, nu=None, kappa=None):
    nu = p.nu if nu is None else nu
    kappa = p.kappa if kappa is None else kappa
    return assembly.assemble(p.xyz, p.tet2node, p.tet_material, nu, kappa, p.pec_pairs, p.jq)


@pytest.mark.parametrize("kind", ["cube16", "contrast"])
def test_assembly_matches_oracle(gpu_available, kind):
    p = (W.make_workload("cube16") if kind == "cube16"
         else W.make_problem(12, 3, jitter=(3, 5), materials="contrast", seed=11))
    print(f"[assembly] {kind} tets={p.n_tets}", flush=True)
    g = GlobalAssembly(p.n_nodes, p.tet2node, p.pec_pairs)
    out = g.assemble(p.xyz, p.tet_material, p.nu, p.kappa, p.jq, outputs=("K", "M", "A", "f"))
    o = _oracle(p)
    ok = _oracle(p, kappa=np.zeros_like(p.kappa))   # curl-curl part alone
    om = _oracle(p, nu=np.zeros_like(p.nu))         # kappa-mass part alone
    rp, col, _, _ = g.pattern()
    assert np.array_equal(rp, o["A"].indptr) and np.array_equal(col, o["A"].indices)
    for key, ref in (("A", o["A"]), ("K", ok["A"]), ("M", om["A"])):
        # the oracle drops no structural entry: same pattern in all three
        assert np.array_equal(ref.indices, col)
        got = out[key].cpu().numpy()
        assert np.abs(got - ref.data).max() <= 1e-13 * np.abs(ref.data).max(), key
    f = out["f"].cpu().numpy()
    assert np.abs(f - o["f"]).max() <= 1e-13 * np.abs(o["f"]).max()
    # K + M equals A up to one rounding per entry
    KM = (out["K"] + out["M"]).cpu().numpy()
    assert np.abs(KM - out["A"].cpu().numpy()).max() <= 4e-16 * np.abs(KM).max()


def test_assembly_equals_fetidp_global_system(gpu_available):
    """The same plan and kernels as the FETI-DP handle's residual system: bitwise equal."""
    p = W.make_workload("cube16")
    g = GlobalAssembly(p.n_nodes, p.tet2node, p.pec_pairs)
    out = g.assemble(p.xyz, p.tet_material, p.nu, p.kappa, p.jq)
    s = FETIDPSolver().setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    s.factor()
    rp, col, val, f = s.global_system()
    assert np.array_equal(out["A"].cpu().numpy(), val)
    assert np.array_equal(out["f"].cpu().numpy(), f)


def test_csr_residual(gpu_available):
    import torch
    p = W.make_workload("cube16")
    g = GlobalAssembly(p.n_nodes, p.tet2node, p.pec_pairs)
    out = g.assemble(p.xyz, p.tet_material, p.nu, p.kappa, p.jq)
    rp, col, _, _ = g.pattern()
    A = sp.csr_matrix((out["A"].cpu().numpy(), col, rp), shape=(g.n_free, g.n_free))
    fh = out["f"].cpu().numpy()
    xd = solvers.direct(A, fh)
    rng = np.random.default_rng(3)
    dev = torch.device("cuda")
    trp = torch.as_tensor(rp, device=dev)
    tcol = torch.as_tensor(col, device=dev)
    for x in (xd, xd + 1e-6 * rng.standard_normal(g.n_free), np.zeros(g.n_free)):
        rel, r = csr_residual(trp, tcol, out["A"], torch.as_tensor(x, device=dev), out["f"],
                              want_r=True)
        rh = fh - A @ x
        ref = np.linalg.norm(rh) / np.linalg.norm(fh)
        assert np.abs(r.cpu().numpy() - rh).max() <= 1e-12 * np.abs(fh).max()
        assert abs(rel - ref) <= 1e-10 * ref + 1e-15
        assert rel == csr_residual(trp, tcol, out["A"], torch.as_tensor(x, device=dev), out["f"])
    # f = 0 -> 0
    z = torch.zeros(g.n_free, dtype=torch.float64, device=dev)
    assert csr_residual(trp, tcol, out["A"], z, z) == 0.0


def test_gauged_kappa0_assembly(gpu_available):
    """kappa = 0 with the tree-cotree gauge (tree edges fixed like PEC): the GPU assembles
    the gauged system as the oracle does, and it is SPD (its CG converges from zero)."""
    from emagddm_amd.assembly import tree_gauge_pairs
    p = W.make_problem(12, 2, seed=6, name="gauge12")
    zero = np.zeros_like(p.kappa)
    t = tree_gauge_pairs(p.n_nodes, p.tet2node, p.pec_pairs)
    pairs = np.vstack([p.pec_pairs, t])
    g = GlobalAssembly(p.n_nodes, p.tet2node, pairs)
    out = g.assemble(p.xyz, p.tet_material, p.nu, zero, p.jq, outputs=("A",))
    o = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, zero, pairs, p.jq)
    rp, col, _, _ = g.pattern()
    assert np.array_equal(rp, o["A"].indptr) and np.array_equal(col, o["A"].indices)
    a = out["A"].cpu().numpy()
    assert np.abs(a - o["A"].data).max() <= 1e-13 * np.abs(o["A"].data).max()
    A = sp.csr_matrix((a, col, rp), shape=(g.n_free, g.n_free))
    y = np.random.default_rng(1).standard_normal(g.n_free)
    x, st, its, res = solvers.pcg_jacobi(A, A @ y, 1e-12, 20000)
    assert st == solvers.SUCCESS and np.linalg.norm(x - y) <= 1e-8 * np.linalg.norm(y)
