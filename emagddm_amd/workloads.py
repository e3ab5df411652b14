"""Seeded synthetic workloads for the FETI-DP Nedelec solver.

The reference ships no mesh files and no generator (``mesh/README.md`` says the
meshes are "to be created"; ``include/fe_em/mesh_reader.hpp`` is abstract), so
the unit-cube problems are produced here, deterministically from a seed:

* ``kuhn_cube(n)``: n^3 hexes, each split into the 6 Kuhn tetrahedra around
  the (0,0,0)->(1,1,1) diagonal.  Node id = i + (n+1)*(j + (n+1)*k).  Every
  tet lists its nodes in ascending global id (the path v000 -> v000+e_a ->
  v000+e_a+e_b -> v111 is monotone in id).
* ``box_partition``: subdomain id per tet from per-axis cut planes, either
  even or jittered (widths in [wmin, wmax] summing to n).
* ``pec_edges``: every boundary edge as an explicit node pair, the same form
  as ``EMBoundaryMarker{DIRICHLET, VECTOR_EDGE_ONLY, edge node pairs, 0}``
  (``include/fe_em/em_mesh_data.hpp:166-176``).
* ``loop_source``: a sum of Gaussian current loops, azimuthal about each
  loop axis, so div J = 0 analytically; sampled at the 4 quadrature points of
  each tet (point q has barycentric weight b on node q and a on the others,
  a=(5-sqrt5)/20, b=(5+3sqrt5)/20).

``make_workload(name, seed)`` builds the five named problem classes
(16^3 .. 96^3 hexes).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

QA = (5.0 - math.sqrt(5.0)) / 20.0
QB = (5.0 + 3.0 * math.sqrt(5.0)) / 20.0

# Axis permutations; tet p of a hex follows v000 -> +e_a -> +e_b -> v111.
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


@dataclasses.dataclass
class Problem:
    """One synthetic curl(nu curl A) + kappa A = J problem on a Kuhn cube."""
    name: str
    n: int
    xyz: np.ndarray           # (n_nodes, 3) float64
    tet2node: np.ndarray      # (n_tets, 4) int32, ascending node ids
    tet_material: np.ndarray  # (n_tets,) int32
    nu: np.ndarray            # (n_mat,) float64, reluctivity 1/mu
    kappa: np.ndarray         # (n_mat,) float64
    tet_subdomain: np.ndarray  # (n_tets,) int32
    n_sub: int
    pec_pairs: np.ndarray     # (n_pec, 2) int32
    jq: np.ndarray            # (n_tets, 4, 3) float64: J at the 4 quadrature points
    cuts: tuple = ()

    @property
    def n_nodes(self):
        return self.xyz.shape[0]

    @property
    def n_tets(self):
        return self.tet2node.shape[0]


def kuhn_cube(n: int):
    """Nodes and tets of the unit cube split into n^3 hexes x 6 Kuhn tets."""
    m = n + 1
    g = np.arange(m, dtype=np.float64) / n
    k3, j3, i3 = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    xyz = np.stack([g[i3.ravel()], g[j3.ravel()], g[k3.ravel()]], axis=1)
    hk, hj, hi = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = (hi + m * (hj + m * hk)).ravel().astype(np.int64)
    step = np.array([1, m, m * m], dtype=np.int64)
    tets = np.empty((base.size, 6, 4), dtype=np.int64)
    for p, (a, b, _c) in enumerate(_PERMS):
        tets[:, p, 0] = base
        tets[:, p, 1] = base + step[a]
        tets[:, p, 2] = base + step[a] + step[b]
        tets[:, p, 3] = base + step.sum()
    return xyz, tets.reshape(-1, 4).astype(np.int32)


def jittered_widths(n: int, parts: int, wmin: int, wmax: int, rng) -> list:
    """Random composition of n into `parts` widths, each in [wmin, wmax]."""
    if not (parts * wmin <= n <= parts * wmax):
        raise ValueError("infeasible jitter bounds")
    w = [n // parts] * parts
    for i in range(n - sum(w)):
        w[i % parts] += 1
    for _ in range(50 * parts):
        i, j = rng.integers(0, parts, size=2)
        d = int(rng.integers(1, 3))
        if i != j and w[i] + d <= wmax and w[j] - d >= wmin:
            w[i] += d
            w[j] -= d
    return w


def box_partition(n: int, widths_xyz) -> np.ndarray:
    """Subdomain id per tet for box cuts given by per-axis width lists."""
    owner = []
    for widths in widths_xyz:
        o = np.repeat(np.arange(len(widths)), widths)
        assert o.size == n
        owner.append(o)
    px, py = len(widths_xyz[0]), len(widths_xyz[1])
    hk, hj, hi = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    sub = owner[0][hi.ravel()] + px * (owner[1][hj.ravel()] + py * owner[2][hk.ravel()])
    return np.repeat(sub.astype(np.int32), 6)


_GEOM_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]


def pec_edges(n: int, tet2node: np.ndarray) -> np.ndarray:
    """All boundary edges (both end nodes on one face of the cube) as pairs."""
    m = n + 1
    a = np.concatenate([tet2node[:, [p, q]] for p, q in _GEOM_EDGES]).astype(np.int64)
    a.sort(axis=1)
    key = np.unique(a[:, 0] * (m ** 3) + a[:, 1])
    lo, hi = key // (m ** 3), key % (m ** 3)

    def ijk(v):
        return v % m, (v // m) % m, v // (m * m)

    ilo, ihi = ijk(lo), ijk(hi)
    on = np.zeros(lo.size, dtype=bool)
    for ax in range(3):
        for side in (0, n):
            on |= (ilo[ax] == side) & (ihi[ax] == side)
    return np.stack([lo[on], hi[on]], axis=1).astype(np.int32)


def quadrature_points(xyz: np.ndarray, tet2node: np.ndarray) -> np.ndarray:
    """Physical 4-point rule nodes, shape (n_tets, 4, 3); point q leans to node q."""
    X = xyz[tet2node]                     # (t, 4, 3)
    s = X.sum(axis=1, keepdims=True)      # (t, 1, 3)
    return QA * s + (QB - QA) * X


@dataclasses.dataclass
class Loop:
    center: np.ndarray
    normal: np.ndarray
    radius: float
    width: float
    current: float


def random_loops(rng, count: int, rmin=0.1, rmax=0.2, width=0.05) -> list:
    loops = []
    for _ in range(count):
        r = float(rng.uniform(rmin, rmax))
        margin = r + 3 * width
        c = rng.uniform(min(margin, 0.45), max(1 - margin, 0.55), size=3)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        loops.append(Loop(c, v, r, width, float(rng.uniform(0.5, 1.5))))
    return loops


def loop_source(points: np.ndarray, loops) -> np.ndarray:
    """Sum of divergence-free Gaussian current loops evaluated at `points`."""
    shp = points.shape
    p = points.reshape(-1, 3)
    out = np.zeros_like(p)
    for lp in loops:
        d = p - lp.center
        z = d @ lp.normal
        rv = d - z[:, None] * lp.normal[None, :]
        rho = np.linalg.norm(rv, axis=1)
        ephi = np.cross(np.broadcast_to(lp.normal, rv.shape), rv)
        safe = np.where(rho > 1e-300, rho, 1.0)
        ephi /= safe[:, None]
        amp = lp.current * np.exp(-((rho - lp.radius) ** 2 + z ** 2) / (2 * lp.width ** 2))
        out += amp[:, None] * ephi
    return out.reshape(shp)


# name, n, parts per axis, jitter (wmin, wmax) or None, material model, loops
WORKLOADS = {
    # 16^3 hexes, 2x2x2 even boxes, nu=1, kappa=1e-2, 2 loops (CPU-runnable)
    "cube16": dict(n=16, parts=2, jitter=None, materials="uniform", loops=2),
    # 48^3 hexes, 6x6x6 jittered boxes 6..10 wide, nu=1, kappa=1e-2
    "cube48j": dict(n=48, parts=6, jitter=(6, 10), materials="uniform", loops=2),
    # 64^3 hexes, 8x8x8 even boxes, log-uniform nu/kappa per box + nu=1e-3 sphere
    "cube64c": dict(n=64, parts=8, jitter=None, materials="contrast", loops=2),
    # 32^3 hexes, 4x4x4 boxes, nu per box, 1..4 random loops (training family)
    "cube32l": dict(n=32, parts=4, jitter=None, materials="nu_per_sub", loops=(1, 4)),
    # 96^3 hexes, 12x12x12 jittered boxes, contrast materials
    "cube96j": dict(n=96, parts=12, jitter=(6, 10), materials="contrast", loops=2),
    # 48^3 hexes, 6x6x6 jittered boxes, contrast materials (training family of the
    # lambda0 network used at 96^3)
    "cube48c": dict(n=48, parts=6, jitter=(6, 10), materials="contrast", loops=2),
}


def _draw_materials(rng, materials: str, n_sub: int):
    if materials == "uniform":
        return np.array([1.0]), np.array([1e-2])
    if materials == "nu_per_sub":
        return 10.0 ** rng.uniform(-3, 0, size=n_sub), np.full(n_sub, 1e-2)
    raise ValueError(materials)


def _draw_loops(rng, loops):
    if isinstance(loops, tuple):
        count = int(rng.integers(loops[0], loops[1] + 1))
        return random_loops(rng, count, 0.05, 0.2, 0.04)
    return random_loops(rng, loops)


def family_parameters(name: str, seed: int):
    """(nu, kappa, loops) that make_workload(name, seed) draws, for reuse on the
    mesh, partition and material map of seed 0.  For "contrast" families the
    material map (which tets lie in the nu = 1e-3 sphere) stays that of seed 0;
    the per-subdomain nu and kappa and the loops are the seed's."""
    c = WORKLOADS[name]
    n, parts = c["n"], c["parts"]
    rng = np.random.default_rng(np.random.SeedSequence([seed, n, parts]))
    if c["jitter"] is not None:  # consume the cut draws exactly as make_problem does
        for _ in range(3):
            jittered_widths(n, parts, c["jitter"][0], c["jitter"][1], rng)
    n_sub = parts ** 3
    if c["materials"] == "contrast":
        nu_s = 10.0 ** rng.uniform(-3, 0, size=n_sub)
        ka_s = 10.0 ** rng.uniform(-4, 0, size=n_sub)
        rng.uniform(0.3, 0.7, size=3)  # the sphere's centre and radius (map of seed 0)
        rng.uniform(0.1, 0.2)
        nu = np.concatenate([nu_s, np.full(n_sub, 1e-3)])
        kappa = np.concatenate([ka_s, ka_s])
    else:
        nu, kappa = _draw_materials(rng, c["materials"], n_sub)
    return nu, kappa, _draw_loops(rng, c["loops"])


def make_problem(n: int, parts: int, jitter=None, materials="uniform", loops=2,
                 seed: int = 0, name: str = "custom") -> Problem:
    """Build one problem.  `seed` fixes cuts, materials and sources."""
    rng = np.random.default_rng(np.random.SeedSequence([seed, n, parts]))
    xyz, tets = kuhn_cube(n)
    if jitter is None:
        if n % parts:
            raise ValueError("even partition needs parts | n")
        widths = [[n // parts] * parts] * 3
    else:
        widths = [jittered_widths(n, parts, jitter[0], jitter[1], rng) for _ in range(3)]
    sub = box_partition(n, widths)
    n_sub = parts ** 3
    pec = pec_edges(n, tets)
    if materials in ("uniform", "nu_per_sub"):
        mat = np.zeros(tets.shape[0], dtype=np.int32) if materials == "uniform" else sub.copy()
        nu, kappa = _draw_materials(rng, materials, n_sub)
    elif materials == "contrast":
        nu_s = 10.0 ** rng.uniform(-3, 0, size=n_sub)
        ka_s = 10.0 ** rng.uniform(-4, 0, size=n_sub)
        c = rng.uniform(0.3, 0.7, size=3)
        r = float(rng.uniform(0.1, 0.2))
        cent = xyz[tets].mean(axis=1)
        inside = np.linalg.norm(cent - c, axis=1) < r
        mat = np.where(inside, sub + n_sub, sub).astype(np.int32)
        nu = np.concatenate([nu_s, np.full(n_sub, 1e-3)])
        kappa = np.concatenate([ka_s, ka_s])
    else:
        raise ValueError(materials)
    lps = _draw_loops(rng, loops)
    jq = loop_source(quadrature_points(xyz, tets), lps)
    return Problem(name=name, n=n, xyz=xyz, tet2node=tets, tet_material=mat, nu=nu,
                   kappa=kappa, tet_subdomain=sub, n_sub=n_sub, pec_pairs=pec, jq=jq,
                   cuts=tuple(tuple(w) for w in widths))


def irregular_problem(n: int, n_seeds: int, seed: int = 0, perturb: float = 0.25,
                      name: str = "irregular") -> Problem:
    """General geometry and partition: interior nodes of the Kuhn cube moved by up to
    `perturb`·h per coordinate (non-Kuhn element shapes), and each tet assigned to the
    nearest of `n_seeds` random points (Voronoi cells: staircase faces, subdomain
    edges along arbitrary lines, corners shared by three or more subdomains)."""
    base = make_problem(n, 1, seed=seed, name=name)
    rng = np.random.default_rng(np.random.SeedSequence([seed, n, n_seeds, 7]))
    xyz = base.xyz.copy()
    inside = np.all((xyz > 1e-9) & (xyz < 1 - 1e-9), axis=1)
    xyz[inside] += rng.uniform(-perturb, perturb, size=(int(inside.sum()), 3)) / n
    cent = xyz[base.tet2node].mean(axis=1)
    pts = rng.uniform(0, 1, size=(n_seeds, 3))
    near = np.argmin(((cent[:, None, :] - pts[None]) ** 2).sum(-1), axis=1)
    _, sub = np.unique(near, return_inverse=True)
    sub = sub.astype(np.int32)
    n_sub = int(sub.max()) + 1
    nu = base.nu
    return Problem(name=name, n=n, xyz=xyz, tet2node=base.tet2node,
                   tet_material=base.tet_material, nu=nu, kappa=base.kappa,
                   tet_subdomain=sub, n_sub=n_sub, pec_pairs=base.pec_pairs, jq=base.jq,
                   cuts=())


def make_workload(name: str, seed: int = 0) -> Problem:
    c = WORKLOADS[name]
    return make_problem(c["n"], c["parts"], c["jitter"], c["materials"], c["loops"],
                        seed=seed, name=name)
