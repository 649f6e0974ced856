"""CPU multifrontal LU + Krylov-Schur (TEST INFRASTRUCTURE / CPU BASELINE).

A restatement on the host of the algorithm the GPU path runs, built from
LAPACK and BLAS through scipy (zgetrf on each pivot block, ztrsm, zgemm),
so that the CPU baseline is the same method at CPU speed rather than a
general-purpose sparse LU:

* geometric nested dissection of the FD-q block graph with one-sided
  vertex separators (minimum separator of the two coordinate splits, leaves
  of at most 64 nodes);
* multifrontal factorization of K = A_bc - sigma B_bc with partial pivoting
  inside each front's fully summed block;
* shift-invert Krylov-Schur from oracle/eigen.py on top of its solves.

The reference's own path here is SLEPc + MUMPS (EigenSolver.cpp:50-75),
which is not available in this image.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

import oracle
from oracle import eigen


def blas_threads() -> int:
    try:
        from threadpoolctl import threadpool_info
        info = threadpool_info()
        return max((d.get("num_threads", 1) for d in info if d.get("user_api") == "blas"),
                   default=1)
    except Exception:
        return 1


def stencil_graph(case):
    ny, nz = case.ny, case.nz
    xoff = case.xi_rule.stencil_offsets
    eoff = case.eta_rule.stencil_offsets
    n = ny * nz
    nb = [set() for _ in range(n)]
    for j in range(nz):
        jc = (j + eoff[j]) % nz
        for i in range(ny):
            u = j * ny + i
            cols = (jc[:, None] * ny + (i + xoff[i])[None, :]).ravel()
            nb[u].update(cols.tolist())
            for v in cols.tolist():
                nb[v].add(u)
    return [np.array(sorted(s), dtype=np.int64) for s in nb]


def nested_dissection(case, adj, leaf: int = 64):
    """Supernodes in postorder: list of (sorted node array, children)."""
    ny, nz = case.ny, case.nz
    n = ny * nz
    tree = []
    mark = np.zeros(n, dtype=bool)

    def rec(nodes, ring):
        if len(nodes) == 0:
            return None
        if len(nodes) <= leaf:
            tree.append((np.sort(nodes), []))
            return len(tree) - 1
        best = None
        for d in (0, 1):
            coord = nodes % ny if d == 0 else nodes // ny
            if d == 1 and ring:
                mid = nz // 2
            else:
                if d == 1:
                    cnt = len(np.unique(coord))
                    if coord.max() - coord.min() + 1 > cnt:
                        coord = np.where(coord < nz // 2, coord + nz, coord)
                lo, hi = coord.min(), coord.max()
                if hi - lo < 2:
                    continue
                mid = (lo + hi + 1) // 2
            L, R = nodes[coord < mid], nodes[coord >= mid]
            if len(L) == 0 or len(R) == 0:
                continue
            mark[R] = True
            adjL = np.array([mark[adj[u]].any() for u in L], dtype=bool)
            mark[R] = False
            sep, left = L[adjL], L[~adjL]
            if best is None or len(sep) < len(best[0]):
                best = (sep, left, R, d == 0)
        if best is None:
            tree.append((np.sort(nodes), []))
            return len(tree) - 1
        sep, left, right, xsplit = best
        c1 = rec(left, ring and xsplit)
        c2 = rec(right, ring and xsplit)
        tree.append((np.sort(sep), [c for c in (c1, c2) if c is not None]))
        return len(tree) - 1

    rec(np.arange(n, dtype=np.int64), True)
    return tree


class MultifrontalLU:
    def __init__(self, case, rowptr, colidx, kvals, adj=None):
        ny, nz = case.ny, case.nz
        n = ny * nz
        self.n = n
        adj = adj if adj is not None else stencil_graph(case)
        tree = nested_dissection(case, adj)
        pos = np.empty(n, dtype=np.int64)
        starts = [0]
        order = []
        for nodes, _ in tree:
            pos[nodes] = np.arange(len(order), len(order) + len(nodes))
            order.extend(nodes.tolist())
            starts.append(len(order))
        self.pos, self.tree = pos, tree
        parent = [-1] * len(tree)
        for s, (_, ch) in enumerate(tree):
            for c in ch:
                parent[c] = s
        structs = []
        for s, (nodes, ch) in enumerate(tree):
            last = starts[s + 1] - 1
            cand = set()
            for u in nodes:
                a = adj[u]
                cand.update(a[pos[a] > last].tolist())
            for c in ch:
                cand.update(v for v in structs[c] if pos[v] > last)
            structs.append(np.array(sorted(cand, key=lambda v: pos[v]), dtype=np.int64))
        self.structs, self.parent = structs, parent
        # front index maps and block assignment
        per = len(colidx) // n
        u_of = np.repeat(np.arange(n), per)
        first = np.minimum(pos[u_of], pos[colidx])
        sn_of_pos = np.empty(n, dtype=np.int64)
        for s in range(len(tree)):
            sn_of_pos[starts[s]:starts[s + 1]] = s
        fblk = sn_of_pos[first]
        order_blocks = np.argsort(fblk, kind="stable")
        bounds = np.searchsorted(fblk[order_blocks], np.arange(len(tree) + 1))
        local = np.full(n, -1, dtype=np.int64)
        self.factors = [None] * len(tree)
        upd = [None] * len(tree)
        for s, (nodes, ch) in enumerate(tree):
            fnodes = np.concatenate([nodes, structs[s]])
            local[fnodes] = np.arange(len(fnodes))
            m = 5 * len(fnodes)
            p = 5 * len(nodes)
            F = np.zeros((m, m), dtype=np.complex128)
            ks = order_blocks[bounds[s]:bounds[s + 1]]
            lu_ = local[u_of[ks]]
            lv_ = local[colidx[ks]]
            rr = (5 * lu_)[:, None, None] + np.arange(5)[None, :, None]
            cc = (5 * lv_)[:, None, None] + np.arange(5)[None, None, :]
            F[rr, cc] = kvals[ks]
            for c in ch:
                cn = structs[c]
                rel = (5 * local[cn])[:, None] + np.arange(5)[None, :]
                rel = rel.ravel()
                F[np.ix_(rel, rel)] += upd[c]
                upd[c] = None
            lu, piv = sla.lu_factor(F[:p, :p], check_finite=False)
            F12 = F[:p, p:].copy()
            for i in range(p):
                if piv[i] != i:
                    F12[[i, piv[i]]] = F12[[piv[i], i]]
            U12 = sla.solve_triangular(lu, F12, lower=True, unit_diagonal=True, check_finite=False)
            L21 = sla.solve_triangular(lu, F[p:, :p].T, trans="T", lower=False,
                                       check_finite=False).T
            if m > p:
                upd[s] = F[p:, p:] - L21 @ U12
            self.factors[s] = (lu, piv, L21, U12, 5 * starts[s], p)
            local[fnodes] = -1
        self.perm = np.repeat(5 * np.asarray(order), 5) + np.tile(np.arange(5), n)
        self.local_struct = [None] * len(tree)
        for s in range(len(tree)):
            self.local_struct[s] = (np.repeat(5 * pos[structs[s]], 5)
                                    + np.tile(np.arange(5), len(structs[s])))

    def solve(self, b):
        """x = K^-1 b (natural order)."""
        be = b[self.perm].astype(np.complex128)
        nf = len(self.tree)
        y = np.zeros_like(be)
        u = [None] * nf
        for s in range(nf):
            lu, piv, L21, U12, d0, p = self.factors[s]
            m = p + L21.shape[0]
            v = np.zeros(m, dtype=np.complex128)
            v[:p] = be[d0:d0 + p]
            for c in self.tree[s][1]:
                v[self._rel(c, s)] += u[c]
                u[c] = None
            vp = v[:p].copy()
            for i in range(p):
                if piv[i] != i:
                    vp[i], vp[piv[i]] = vp[piv[i]], vp[i]
            ys = sla.solve_triangular(lu, vp, lower=True, unit_diagonal=True, check_finite=False)
            y[d0:d0 + p] = ys
            u[s] = v[p:] - L21 @ ys
        x = np.zeros_like(be)
        for s in range(nf - 1, -1, -1):
            lu, piv, L21, U12, d0, p = self.factors[s]
            rhs = y[d0:d0 + p] - U12 @ x[self.local_struct[s]]
            x[d0:d0 + p] = sla.solve_triangular(lu, rhs, lower=False, check_finite=False)
        out = np.empty_like(x)
        out[self.perm] = x
        return out

    def _rel(self, child, parent):
        key = (child, parent)
        cache = getattr(self, "_relcache", None)
        if cache is None:
            cache = self._relcache = {}
        if key not in cache:
            nodes, _ = self.tree[parent]
            fnodes = np.concatenate([nodes, self.structs[parent]])
            where = {int(v): k for k, v in enumerate(fnodes)}
            idx = np.array([where[int(v)] for v in self.structs[child]], dtype=np.int64)
            cache[key] = (5 * idx[:, None] + np.arange(5)[None, :]).ravel()
        return cache[key]


def solve_case(case, nev: int = 8, tol: float = 1e-10, maxit: int = 1000, ncv: int | None = None):
    """Assemble on the CPU, factor, Krylov-Schur; returns lambdas/residuals."""
    a = oracle.assemble(case)
    n = case.ny * case.nz
    sigma = eigen.target_lambda(case.omega_shift)
    per = len(a.colidx) // n
    kvals = a.A.copy()
    diag = a.colidx == np.repeat(np.arange(n), per)
    kvals[diag] -= sigma * a.B
    lu = MultifrontalLU(case, a.rowptr, a.colidx, kvals)
    A = oracle.to_scipy(a.rowptr, a.colidx, a.A, n)
    B = oracle.block_diag_to_scipy(a.B)
    lams, vecs, res, restarts = eigen.krylov_schur(A, B, sigma, nev=nev, tol=tol, maxit=maxit,
                                                   solver=lu.solve, true_residual=True, ncv=ncv)
    return {"lambdas": lams, "residuals": res, "restarts": restarts}
