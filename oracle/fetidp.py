"""Dense CPU FETI-DP in FP64 -- the oracle of the GPU solver's algorithm.

The reference has no FETI-DP solver (src/fetidp/fetidp_context.cpp:19-35 only
zero-fills SubdomainData records), so this module states the method the GPU
solver implements, in the plainest dense form, for small meshes:

* Edge classes by subdomain multiplicity: 1 interior (i), 2 dual (Delta, one
  multiplier lambda, B = +1 on the lower subdomain id, -1 on the higher), >=3
  primal (c, kept continuous through the coarse problem).  Lambdas and
  primals are numbered in ascending free-dof order; local dofs are ordered
  (i, Delta, c), each block ascending in free-dof index.
* Local K_s = sum over the subdomain's tets of s_a s_b (K_e + M_e)(a, b);
  local f_s = global f divided by the edge's subdomain multiplicity (any split
  with sum_s f_s = f gives the same u; this one fixes lambda).
* Z_s = [K_rr^-1]_DD, Phi_s = [K_rr^-1 K_rc]_D, S~_s = K_cc - K_cr K_rr^-1 K_rc,
  Dirichlet Schur P_s = K_DD - K_Di K_ii^-1 K_iD.
* F = sum B Z B^T + sum B Phi R S~^-1 R^T Phi^T B^T,
  d = sum B a_s - sum B Phi R S~^-1 f~_c,  a_s = [K_rr^-1 f_r]_D,
  f~_c = sum R^T (f_c - K_cr K_rr^-1 f_r).
* Preconditioner M^-1 = sum B_D P B_D^T with stiffness scaling (or, with
  scaling="deluxe", the face matrices D_F = (P_s0,FF + P_s1,FF)^-1 P_s1,FF): the row of
  subdomain s for a multiplier shared with t is weighted rho_t/(rho_s+rho_t),
  rho = diagonal of the local matrix at that edge.
* PCPG = PCG on F lambda = d from lambda0 (default 0), stop at
  ||r||/||d|| < tol, with the CGSolver breakdown / divergence rules
  (src/numeric/iterative_solvers.cpp:449-491).
* Recovery: u_c = S~^-1 (f~_c + sum R^T Phi^T B^T lambda),
  u_D,s = a_s - Z_s B^T lambda - Phi_s R u_c on each side; the global vector
  takes u_c on primal edges and on each dual edge the mean of the two sides
  (stiffness scaling) or their deluxe average u_s0 + D_F (u_s1 - u_s0) per face
  (deluxe scaling), then every interior block is re-solved against those values,
  u_i = K_ii^-1 (f_i - K_ib u_b), so that only interface rows keep a residual.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from . import assembly, topology
from .solvers import SUCCESS, NUMERICAL_ERROR, DIVERGED, MAX_ITER_REACHED


class DenseFETIDP:
    def __init__(self, xyz, tet2node, tet_material, nu, kappa, tet_subdomain,
                 pec_pairs, jq, scaling="stiffness"):
        self.scaling = scaling
        asm = assembly.assemble(xyz, tet2node, tet_material, nu, kappa, pec_pairs, jq)
        self.A, self.f = asm["A"], asm["f"]
        l2g, sgn = asm["l2g"], asm["signs"]
        n = self.A.shape[0]
        self.n_free = n
        Ae = (asm["Ke"] + asm["Me"]) * sgn[:, :, None] * sgn[:, None, :]
        Fe = asm["Fe"] * sgn
        tsub = np.asarray(tet_subdomain, dtype=np.int64)
        n_sub = int(tsub.max()) + 1
        self.n_sub = n_sub
        mult, dofs, subs = topology.classify(l2g, tsub, n)
        self.mult = mult
        dual = np.flatnonzero(mult == 2)
        prim = np.flatnonzero(mult >= 3)
        self.lam_of = -np.ones(n, dtype=np.int64)
        self.lam_of[dual] = np.arange(dual.size)
        self.c_of = -np.ones(n, dtype=np.int64)
        self.c_of[prim] = np.arange(prim.size)
        self.n_lam, self.n_c = dual.size, prim.size
        self.subs = []
        for s in range(n_sub):
            mine = np.sort(dofs[subs == s])
            ii = mine[mult[mine] == 1]
            dd = mine[mult[mine] == 2]
            cc = mine[mult[mine] >= 3]
            order = np.concatenate([ii, dd, cc])
            loc = -np.ones(n, dtype=np.int64)
            loc[order] = np.arange(order.size)
            tets = np.flatnonzero(tsub == s)
            L = l2g[tets]
            ok = L >= 0
            ml = np.where(ok, loc[np.where(ok, L, 0)], -1)
            nl = order.size
            K = np.zeros((nl, nl))
            rr = np.repeat(ml, 6, axis=1)
            cc2 = np.tile(ml, (1, 6))
            m2 = (rr >= 0) & (cc2 >= 0)
            np.add.at(K, (rr[m2], cc2[m2]), Ae[tets].reshape(-1, 36)[m2])
            # right-hand side split: f_s = f / multiplicity on every shared edge
            fl = self.f[order] / mult[order]
            ni, nd = ii.size, dd.size
            self.subs.append(dict(order=order, ni=ni, nd=nd, nc=cc.size, K=K, f=fl,
                                  lam=self.lam_of[dd], cidx=self.c_of[cc], dofs_d=dd))
        # multiplier signs and stiffness scaling
        owners = {}
        for s, S in enumerate(self.subs):
            for k, l in enumerate(S["lam"]):
                owners.setdefault(int(l), []).append((s, k))
        for S in self.subs:
            S["sgn"] = np.zeros(S["nd"])
            S["D"] = np.zeros(S["nd"])
        for l, lst in owners.items():
            assert len(lst) == 2
            (s0, k0), (s1, k1) = sorted(lst)
            r0 = self.subs[s0]["K"][self.subs[s0]["ni"] + k0, self.subs[s0]["ni"] + k0]
            r1 = self.subs[s1]["K"][self.subs[s1]["ni"] + k1, self.subs[s1]["ni"] + k1]
            self.subs[s0]["sgn"][k0] = 1.0
            self.subs[s1]["sgn"][k1] = -1.0
            self.subs[s0]["D"][k0] = r1 / (r0 + r1)
            self.subs[s1]["D"][k1] = r0 / (r0 + r1)
        self._factor()
        if scaling == "deluxe":
            self._deluxe_setup()

    def _deluxe_setup(self):
        """Deluxe weights per interface face F = (s0 < s1): D_F = (S0 + S1)^-1 S1 with
        S_k = P_k restricted to F (Dohrmann & Widlund 2016, "A BDDC algorithm with
        deluxe scaling for three-dimensional H(curl) problems").  In the dual form the
        scaled jump of side s0 applies D_F and that of side s1 applies -(I - D_F), so
        sum_s B_D,s B_s^T = I; with diagonal face blocks this is stiffness scaling
        with rho = diag(P).  No reference counterpart (the reference has no FETI-DP)."""
        owners = {}
        for s, S in enumerate(self.subs):
            for k, l in enumerate(S["lam"]):
                owners.setdefault(int(l), []).append((s, k))
        pairs = {}
        for l, lst in sorted(owners.items()):
            (s0, k0), (s1, k1) = sorted(lst)
            pairs.setdefault((s0, s1), []).append((l, k0, k1))
        self.faces = []
        for (s0, s1), items in sorted(pairs.items()):
            lams = np.array([i[0] for i in items])
            k0 = np.array([i[1] for i in items])
            k1 = np.array([i[2] for i in items])
            S0 = self.subs[s0]["P"][np.ix_(k0, k0)]
            S1 = self.subs[s1]["P"][np.ix_(k1, k1)]
            D0 = np.linalg.solve(S0 + S1, S1)
            self.faces.append((s0, s1, lams, k0, k1, D0))

    def _apply_prec_deluxe(self, r):
        # x_s0 = D r_F, x_s1 = -(I - D) r_F;  z_F = D^T y_s0 - (I - D)^T y_s1
        xs = [np.zeros(S["nd"]) for S in self.subs]
        for (s0, s1, lams, k0, k1, D0) in self.faces:
            g = D0 @ r[lams]
            xs[s0][k0] += g
            xs[s1][k1] += g - r[lams]
        ys = [S["P"] @ x for S, x in zip(self.subs, xs)]
        z = np.zeros(self.n_lam)
        for (s0, s1, lams, k0, k1, D0) in self.faces:
            z[lams] += D0.T @ (ys[s0][k0] + ys[s1][k1]) - ys[s1][k1]
        return z

    def _factor(self):
        Sc = np.zeros((self.n_c, self.n_c))
        fc = np.zeros(self.n_c)
        for S in self.subs:
            ni, nd, nc = S["ni"], S["nd"], S["nc"]
            nr = ni + nd
            K, f = S["K"], S["f"]
            Krr, Krc, Kcc = K[:nr, :nr], K[:nr, nr:], K[nr:, nr:]
            cf = sla.cho_factor(Krr, lower=True)
            X = sla.cho_solve(cf, np.concatenate([Krc, f[:nr, None]], axis=1))
            S["Z"] = sla.cho_solve(cf, np.eye(nr)[:, ni:])[ni:, :]
            S["Phi"] = X[ni:, :nc]
            S["Stil"] = Kcc - Krc.T @ X[:, :nc]
            S["a"] = X[ni:, nc]
            S["b"] = Krc.T @ X[:, nc]
            S["cf"] = cf
            if ni:
                cfi = sla.cho_factor(K[:ni, :ni], lower=True)
                S["P"] = K[ni:nr, ni:nr] - K[ni:nr, :ni] @ sla.cho_solve(cfi, K[:ni, ni:nr])
            else:
                S["P"] = K[ni:nr, ni:nr].copy()
            ci = S["cidx"]
            Sc[np.ix_(ci, ci)] += S["Stil"]
            np.add.at(fc, ci, f[nr:] - S["b"])
        self.Sc = Sc
        self.fc = fc
        self.Sc_cf = sla.cho_factor(Sc, lower=True) if self.n_c else None
        self.d = self.apply_F_parts(None)

    def _coarse_solve(self, g):
        return sla.cho_solve(self.Sc_cf, g) if self.n_c else g

    def apply_F(self, p):
        y = np.zeros(self.n_lam)
        g = np.zeros(self.n_c)
        ys = []
        for S in self.subs:
            x = S["sgn"] * p[S["lam"]]
            ys.append(S["Z"] @ x)
            np.add.at(g, S["cidx"], S["Phi"].T @ x)
        uc = self._coarse_solve(g)
        for S, yy in zip(self.subs, ys):
            w = yy + S["Phi"] @ uc[S["cidx"]]
            np.add.at(y, S["lam"], S["sgn"] * w)
        return y

    def apply_F_parts(self, _):
        """Dual right-hand side d."""
        d = np.zeros(self.n_lam)
        uc = self._coarse_solve(self.fc)
        for S in self.subs:
            w = S["a"] - S["Phi"] @ uc[S["cidx"]]
            np.add.at(d, S["lam"], S["sgn"] * w)
        return d

    def apply_prec(self, r):
        if self.scaling == "deluxe":
            return self._apply_prec_deluxe(r)
        z = np.zeros(self.n_lam)
        for S in self.subs:
            x = S["D"] * S["sgn"] * r[S["lam"]]
            np.add.at(z, S["lam"], S["sgn"] * S["D"] * (S["P"] @ x))
        return z

    def dense_F(self):
        return np.stack([self.apply_F(e) for e in np.eye(self.n_lam)], axis=1)

    def pcpg(self, tol=1e-8, max_iter=1000, lam0=None, precondition=True):
        d = self.d
        dn = np.linalg.norm(d)
        lam = np.zeros(self.n_lam) if lam0 is None else np.array(lam0, dtype=np.float64)
        hist = []
        if dn < 1e-300:
            return lam, SUCCESS, 0, hist
        r = d - self.apply_F(lam)
        res = np.linalg.norm(r) / dn
        hist.append(res)
        if res < tol:
            return lam, SUCCESS, 0, hist
        z = self.apply_prec(r) if precondition else r.copy()
        p = z.copy()
        rz = r @ z
        prev = res
        status = MAX_ITER_REACHED
        it = 0
        for it in range(1, max_iter + 1):
            q = self.apply_F(p)
            pq = p @ q
            if abs(pq) < 1e-300:
                status = NUMERICAL_ERROR
                break
            alpha = rz / pq
            lam += alpha * p
            r -= alpha * q
            res = np.linalg.norm(r) / dn
            hist.append(res)
            if res < tol:
                status = SUCCESS
                break
            if it > 10 and res > prev * 1000.0:
                status = DIVERGED
                break
            z = self.apply_prec(r) if precondition else r.copy()
            rz_new = r @ z
            p = z + (rz_new / rz) * p
            rz = rz_new
            prev = res
        return lam, status, min(it, max_iter), hist

    def recover(self, lam):
        """Dirichlet recovery: averaged interface values (mean, or deluxe average with
        deluxe scaling), then local interior solves."""
        g = self.fc.copy()
        for S in self.subs:
            x = S["sgn"] * lam[S["lam"]]
            np.add.at(g, S["cidx"], S["Phi"].T @ x)
        uc = self._coarse_solve(g)
        xg = np.zeros(self.n_free)
        cnt = np.zeros(self.n_free)
        uds = []
        for S in self.subs:
            ni, nd = S["ni"], S["nd"]
            ud = S["a"] - S["Z"] @ (S["sgn"] * lam[S["lam"]]) - S["Phi"] @ uc[S["cidx"]]
            uds.append(ud)
            o = S["order"]
            xg[o[ni:ni + nd]] += ud
            cnt[o[ni:ni + nd]] += 1
        xg = np.where(cnt > 0, xg / np.maximum(cnt, 1), xg)
        if self.scaling == "deluxe":
            # (S0 + S1)^-1 (S0 u0 + S1 u1) on every face
            for (s0, s1, lams, k0, k1, D0) in self.faces:
                S = self.subs[s0]
                u0, u1 = uds[s0][k0], uds[s1][k1]
                xg[S["order"][S["ni"] + k0]] = u0 + D0 @ (u1 - u0)
        prim = np.flatnonzero(self.c_of >= 0)
        xg[prim] = uc[self.c_of[prim]]
        for S in self.subs:
            ni = S["ni"]
            if not ni:
                continue
            K, o = S["K"], S["order"]
            rhs = S["f"][:ni] - K[:ni, ni:] @ xg[o[ni:]]
            xg[o[:ni]] = sla.cho_solve(sla.cho_factor(K[:ni, :ni], lower=True), rhs)
        return xg

    def solve(self, tol=1e-8, max_iter=1000, lam0=None):
        lam, status, it, hist = self.pcpg(tol, max_iter, lam0)
        x = self.recover(lam)
        res = np.linalg.norm(self.f - self.A @ x) / np.linalg.norm(self.f)
        return dict(x=x, lam=lam, status=status, iterations=it, history=hist,
                    residual=float(res))
