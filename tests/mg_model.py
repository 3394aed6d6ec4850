"""CPU model of the library's multigrid preconditioner (IGA_PC_FEM_MG), in
scipy sparse matrices, written from its definition rather than its code:

  A_0      the Q1 matrix K with the couplings to Dirichlet (boundary) nodes
           removed and identity rows on them: M K M + (I - M), M the
           interior mask;
  P_l      per direction with more than 17 nodes, coarse nodes = the even
           nodes and the last one, a fine node between two coarse nodes
           takes 1/2 of each; P = Py (x) Px with the fine and coarse
           boundary rows / columns removed;
  A_{l+1}  Galerkin P^T A_l P (identity rows on the coarse boundary);
  smoother Chebyshev-Jacobi of degree 2 on [lmax/4, lmax], lmax the largest
           interior row sum of |D^-1 A|, pre and post;
  coarsest both directions with at most 17 nodes: exact solve.
Used by the GPU tests as the reference of one V-cycle.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _interp1d(n, coarsen):
    if not coarsen:
        return sp.identity(n, format="csr"), n
    C = list(range(0, n, 2))
    if C[-1] != n - 1:
        C.append(n - 1)
    idx = {c: k for k, c in enumerate(C)}
    r, c, v = [], [], []
    for i in range(n):
        if i in idx:
            r.append(i); c.append(idx[i]); v.append(1.0)
        else:
            r += [i, i]; c += [idx[i - 1], idx[i + 1]]; v += [0.5, 0.5]
    return sp.csr_matrix((v, (r, c)), shape=(n, len(C))), len(C)


def _ring_mask(nx, ny):
    m = np.ones((ny, nx))
    m[0, :] = m[-1, :] = m[:, 0] = m[:, -1] = 0
    return m.reshape(-1)


class VCycle:
    def __init__(self, K, nx, ny):
        mask = _ring_mask(nx, ny)
        M = sp.diags(mask)
        A = (M @ K @ M + sp.diags(1.0 - mask)).tocsr()
        self.levels = []
        while True:
            self.levels.append(dict(A=A, nx=nx, ny=ny, mask=mask))
            if nx <= 17 and ny <= 17:
                break
            Px, cnx = _interp1d(nx, nx > 17)
            Py, cny = _interp1d(ny, ny > 17)
            cm = _ring_mask(cnx, cny)
            P = (sp.diags(mask) @ sp.kron(Py, Px) @ sp.diags(cm)).tocsr()
            A = (sp.diags(cm) @ (P.T @ A @ P) @ sp.diags(cm) + sp.diags(1.0 - cm)).tocsr()
            self.levels[-1]["P"] = P
            nx, ny, mask = cnx, cny, cm
        for L in self.levels:
            D = L["A"].diagonal()
            rows = np.abs(L["A"]).sum(axis=1).A1 / np.abs(D)
            lmax = rows[L["mask"] > 0].max()
            lmin = lmax / 4
            th, de = (lmax + lmin) / 2, (lmax - lmin) / 2
            sig = th / de
            rho0 = 1 / sig
            rho1 = 1 / (2 * sig - rho0)
            L.update(Dinv=1 / D, th=th, c1=rho1 * rho0, c2=2 * rho1 / de)
        last = self.levels[-1]
        self.lu = spla.splu(last["A"].tocsc())

    def _smooth(self, L, x, b):
        A, Dinv, msk = L["A"], L["Dinv"], L["mask"]
        d = msk * Dinv * (b - A @ x) / L["th"]
        x = x + d
        d = L["c1"] * d + L["c2"] * msk * Dinv * (b - A @ x)
        return x + d

    def apply(self, b, l=0):
        L = self.levels[l]
        b = b * L["mask"]
        if l == len(self.levels) - 1:
            return self.lu.solve(b) * L["mask"]
        x = self._smooth(L, np.zeros_like(b), b)
        r = (b - L["A"] @ x) * L["mask"]
        x = x + L["P"] @ self.apply(L["P"].T @ r, l + 1)
        return self._smooth(L, x, b)
