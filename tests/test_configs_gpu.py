"""GPU parity at the full sizes of the benchmark workloads.

C3  quarter annulus p = 4, 512 x 512 elements, 5 x 5 Gauss, k = 32 right-hand
    sides f_j = sum_{m,n<=8} a_mn sin(m pi xi) sin(n pi eta), a_mn ~ N(0,1)
    from seed j: the block apply against the oracle's FP64 apply on two
    columns, and the load vectors of f_j against the oracle's.
C5  unit square p = 5, 2048 x 2048: the FEM-preconditioner assembly
    (37,843,009 nonzeros) bit-exact in every index array against the
    oracle's restatement of the reference's elem2coo / dir2coo, values to
    4 eps (L / h) = 1.8e-12 of max|val| (see q1_tol).
"""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import igamf  # noqa: E402

# FP32 block apply against FP64, relative to max|y| (the smooth C3 columns
# measure 4e-6: the element inputs are taken relative to each element's first
# value, so their FP32 rounding scales with h |grad x|, not |x|)
TOL_SMOOTH = 2e-5
TOL_ROUGH = 2e-5
# Q1 values in FP64: a Q1 element's Jacobian is formed from node differences
# of size h, so the nodes' rounding (eps * L) is relative eps * L / h in every
# value; GPU and CPU round differently (fused multiply-adds).  Relative to
# max|val|: 4 eps L / h.
def q1_tol(ne):
    return 4 * np.finfo(np.float64).eps * (ne + 2)


def _oracle(m):
    a = m.arrays()
    return O.Mesh(a["p"], a["q"], a["nex"], a["ney"], a["CP"], a["W"], a["IEN"], a["ID"],
                  a["Dir"], a["C1"], a["C2"], a["h1"], a["h2"])


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def c3_coeffs(j):
    return igamf.rng_normal(j, 64).reshape(8, 8)  # a[m-1][n-1]


def c3_greville_columns(p, ne, k):
    """f_j at the Greville abscissae of the open uniform knots (row eta,
    column xi): the block apply's input columns."""
    nf = ne + p
    knots = np.concatenate([np.zeros(p), np.linspace(0, 1, ne + 1), np.ones(p)])
    gr = np.array([knots[i + 1:i + p + 1].mean() for i in range(nf)])
    S = np.sin(np.pi * np.outer(gr, np.arange(1, 9)))
    return np.stack([(S @ c3_coeffs(j) @ S.T).T.reshape(-1) for j in range(k)])


def c3_qp_values(p, ne, j):
    """f_j at the quadrature points' parametric coordinates, [E][nqp] in the
    reference's element (x-fastest) and point (y-outer) order."""
    nq = p + 1
    g = (np.polynomial.legendre.leggauss(nq)[0] + 1) / 2
    xi = ((np.arange(ne)[:, None] + g[None, :]) / ne).reshape(-1)  # (e, i)
    S = np.sin(np.pi * np.outer(xi, np.arange(1, 9)))
    T = S @ c3_coeffs(j) @ S.T  # [(ex,i)][(ey,jj)]
    return T.reshape(ne, nq, ne, nq).transpose(2, 0, 3, 1).reshape(-1)


@pytest.fixture(scope="module")
def c3():
    p, ne, k = 4, 512, 32
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, p, ne, r0=1.0, r1=2.0)
    return m, igamf.Operator(m), _oracle(m), p, ne, k


def test_c3_block_apply_matches_oracle(c3):
    m, op, om, p, ne, k = c3
    X = torch.from_numpy(c3_greville_columns(p, ne, k)).cuda().float()
    Y = op.apply_block(X).double().cpu().numpy()
    for j in (0, k - 1):
        ref, _ = O.apply_mf_colored(om, X[j].double().cpu().numpy(), 0)
        err = _rel(Y[j], ref)
        print(f"C3 column {j}: FP32 block vs FP64 oracle {err:.2e}")
        assert err < TOL_SMOOTH, (j, err)
    # a rough column through the same launch shape
    Xr = X.clone()
    Xr[1] = torch.from_numpy(igamf.rng_normal(77, op.nfunc)).cuda().float()
    Yr = op.apply_block(Xr)[1].double().cpu().numpy()
    ref, _ = O.apply_mf_colored(om, Xr[1].double().cpu().numpy(), 0)
    assert _rel(Yr, ref) < TOL_ROUGH, _rel(Yr, ref)


def test_c3_load_vectors_match_oracle(c3):
    m, op, om, p, ne, k = c3
    for j in (0, k - 1):
        fq = c3_qp_values(p, ne, j)
        F = op.load_qp(torch.from_numpy(fq).cuda()).cpu().numpy()
        ref = O.load_qp(om, fq)
        assert _rel(F, ref) < 1e-12, (j, _rel(F, ref))
        assert np.all(F[m.arrays()["Dir"]] == 0.0)


def test_c5_fem_assembly_bit_exact():
    p, ne = 5, 2048
    nf = ne + p
    nodes = igamf.fem_reference_nodes(p, p, ne, ne)
    assert np.array_equal(nodes, O.fem_nodes(p, p, ne, ne))
    f = igamf.FemPC(nf, nf, nodes)
    rows, cols, e2c, d2c, val = O.fem_assemble(nf, nf, nodes)
    assert f.view.nnz == len(rows) == 37_843_009
    t = f.to_torch()
    assert np.array_equal(t["coo_row"].cpu().numpy(), rows)
    assert np.array_equal(t["col"].cpu().numpy(), cols)
    rowptr = np.searchsorted(rows, np.arange(nf * nf + 1)).astype(np.int32)
    assert np.array_equal(t["rowptr"].cpu().numpy(), rowptr)
    assert np.array_equal(t["elem2coo"].cpu().numpy().reshape(-1), e2c)
    assert np.array_equal(t["dir2coo"].cpu().numpy(), d2c)
    v = t["val"].cpu().numpy()
    err = float(np.abs(v - val).max()) / float(np.abs(val).max())
    assert err <= q1_tol(ne), (err, q1_tol(ne))
