"""ctypes front end of the CPU oracle (liboracle.so) and of the reference's own
CPU units (_ref/libref_mf.so, _ref/libref_gpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product package never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class _Mesh(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("p", "q", "nex", "ney", "nq1", "nq2", "nfunc")] + [
        ("CP", C.c_void_p), ("W", C.c_void_p), ("IEN", C.c_void_p), ("ID", C.c_void_p),
        ("Dir", C.c_void_p), ("ndir", C.c_int), ("C1", C.c_void_p), ("C2", C.c_void_p),
        ("h1", C.c_void_p), ("h2", C.c_void_p), ("kappa", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _lib = C.CDLL(path)
        L = _lib
        L.orc_rng_normal.argtypes = [C.c_uint64, C.c_int64, _F64P]
        L.orc_rng_uniform.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double, _F64P]
        L.orc_open_uniform_knots.argtypes = [C.c_int, C.c_int, C.c_double, _F64P]
        L.orc_bernstein.argtypes = [C.c_int, C.c_double, _F64P, _F64P]
        L.orc_gauss_legendre01.argtypes = [C.c_int, _F64P, _F64P]
        L.orc_bspline_basis.argtypes = [C.c_int, _F64P, C.c_int, C.c_double, _F64P,
                                        C.POINTER(C.c_int)]
        L.orc_extraction_1d.argtypes = [C.c_int, _F64P, C.c_int, _F64P]
        L.orc_greville.argtypes = [C.c_int, _F64P, C.c_int, _F64P]
        L.orc_ien_2d.argtypes = [C.c_int] * 4 + [_I32P]
        L.orc_id_2d.argtypes = [C.c_int, C.c_int, _I32P, _I32P]
        L.orc_geometry_unit_square.argtypes = [C.c_int] * 4 + [C.c_double] * 2 + [_F64P, _F64P]
        L.orc_geometry_quarter_annulus.argtypes = [C.c_int] * 3 + [C.c_double] * 2 + [_F64P, _F64P]
        L.orc_element_qp.argtypes = [C.c_int] * 4 + [_F64P, _F64P, C.c_double, C.c_double,
                                                     _F64P, C.c_void_p, _F64P, _F64P, _F64P,
                                                     _F64P, C.c_void_p]
        L.orc_apply_mf.argtypes = [C.POINTER(_Mesh), _F64P, _F64P]
        L.orc_apply_mf_colored.argtypes = [C.POINTER(_Mesh), _F64P, _F64P, C.c_int]
        L.orc_apply_sf_ref.argtypes = [C.POINTER(_Mesh), _F64P, _F64P]
        L.orc_load.argtypes = [C.POINTER(_Mesh), C.c_int, C.c_double, _F64P]
        L.orc_load_qp.argtypes = [C.POINTER(_Mesh), _F64P, _F64P]
        L.orc_kappa_grad.argtypes = [C.POINTER(_Mesh), _F64P, _F64P, _F64P]
        L.orc_element_matrices.argtypes = [C.POINTER(_Mesh), C.c_int64, _F64P]
        L.orc_element_matrices.restype = C.c_int64
        L.orc_fem_nodes.argtypes = [C.c_int] * 4 + [C.c_double] * 2 + [_F64P]
        L.orc_fem_pattern.argtypes = [C.c_int, C.c_int, C.c_int64, _I32P, _I32P]
        L.orc_fem_pattern.restype = C.c_int64
        L.orc_fem_elem2coo.argtypes = [C.c_int, C.c_int, _I32P, _I32P, C.c_int64, _I32P]
        L.orc_fem_dir2coo.argtypes = [C.c_int, C.c_int, _I32P, C.c_int64, _I32P]
        L.orc_fem_values.argtypes = [C.c_int, C.c_int, _F64P, _I32P, _I32P, C.c_int,
                                     C.c_int64, _F64P]
        L.orc_invlm.argtypes = [C.c_int, C.c_int, C.c_int, _I32P, _I32P, _I32P, _I32P, _I32P]
        L.orc_invlm.restype = C.c_int64
    return _lib


# --------------------------------------------------------------------------- inputs
def rng_normal(seed, n):
    out = np.empty(n, np.float64)
    lib().orc_rng_normal(seed, n, out)
    return out


def rng_uniform(seed, n, lo, hi):
    out = np.empty(n, np.float64)
    lib().orc_rng_uniform(seed, n, lo, hi, out)
    return out


def open_knots(p, nelem, L=1.0):
    S = np.empty(nelem + 2 * p + 1)
    lib().orc_open_uniform_knots(p, nelem, L, S)
    return S


def bernstein(p, xi):
    B = np.empty(p + 1); dB = np.empty(p + 1)
    lib().orc_bernstein(p, xi, B, dB)
    return B, dB


def gauss01(n):
    x = np.empty(n); w = np.empty(n)
    if lib().orc_gauss_legendre01(n, x, w) != 0:
        raise ValueError("n out of range")
    return x, w


def bspline_basis(p, S, u):
    N = np.empty(p + 1)
    span = C.c_int()
    lib().orc_bspline_basis(p, np.ascontiguousarray(S, np.float64), len(S), u, N,
                            C.byref(span))
    return N, span.value


def extraction(p, S):
    S = np.ascontiguousarray(S, np.float64)
    nelem = len(S) - 2 * p - 1
    Cm = np.empty(nelem * (p + 1) ** 2)
    lib().orc_extraction_1d(p, S, len(S), Cm)
    return Cm.reshape(nelem, p + 1, p + 1)


def greville(p, S):
    S = np.ascontiguousarray(S, np.float64)
    g = np.empty(len(S) - p - 1)
    lib().orc_greville(p, S, len(S), g)
    return g


class Mesh:
    """Arrays of one tensor-product patch in the reference's layout."""

    def __init__(self, p, q, nex, ney, CP, W, IEN, ID, Dir, C1, C2, h1, h2, nq1=None,
                 nq2=None, kappa=None):
        self.p, self.q, self.nex, self.ney = p, q, nex, ney
        self.nq1 = nq1 or p + 1
        self.nq2 = nq2 or q + 1
        self.nfx, self.nfy = nex + p, ney + q
        self.nfunc = self.nfx * self.nfy
        self.CP = np.ascontiguousarray(CP, np.float64).reshape(-1)
        self.W = None if W is None else np.ascontiguousarray(W, np.float64)
        self.IEN = np.ascontiguousarray(IEN, np.int32).reshape(-1)
        self.ID = np.ascontiguousarray(ID, np.int32)
        self.Dir = np.ascontiguousarray(Dir, np.int32)
        self.C1 = np.ascontiguousarray(C1, np.float64).reshape(-1)
        self.C2 = np.ascontiguousarray(C2, np.float64).reshape(-1)
        self.h1 = np.ascontiguousarray(h1, np.float64)
        self.h2 = np.ascontiguousarray(h2, np.float64)
        self.kappa = None if kappa is None else np.ascontiguousarray(kappa, np.float64)

    @property
    def nelem(self):
        return self.nex * self.ney

    @property
    def nloc(self):
        return (self.p + 1) * (self.q + 1)

    def with_kappa(self, kappa):
        m = Mesh(self.p, self.q, self.nex, self.ney, self.CP, self.W, self.IEN, self.ID,
                 self.Dir, self.C1, self.C2, self.h1, self.h2, self.nq1, self.nq2, kappa)
        return m

    def struct(self):
        def ptr(a):
            return None if a is None else a.ctypes.data
        s = _Mesh(self.p, self.q, self.nex, self.ney, self.nq1, self.nq2, self.nfunc,
                  ptr(self.CP), ptr(self.W), ptr(self.IEN), ptr(self.ID), ptr(self.Dir),
                  len(self.Dir), ptr(self.C1), ptr(self.C2), ptr(self.h1), ptr(self.h2),
                  ptr(self.kappa))
        s._keep = self  # keep arrays alive
        return s


def topology(p, q, nex, ney):
    nfx, nfy = nex + p, ney + q
    IEN = np.empty(nex * ney * (p + 1) * (q + 1), np.int32)
    lib().orc_ien_2d(p, q, nex, ney, IEN)
    ID = np.empty(nfx * nfy, np.int32)
    Dir = np.empty(nfx * nfy, np.int32)
    nd = lib().orc_id_2d(nfx, nfy, ID, Dir)
    return IEN, ID, Dir[:nd].copy()


def mesh_unit_square(p, nex, ney=None, L=1.0, q=None):
    q = p if q is None else q
    ney = nex if ney is None else ney
    nfx, nfy = nex + p, ney + q
    CP = np.empty(2 * nfx * nfy)
    W = np.empty(nfx * nfy)
    lib().orc_geometry_unit_square(p, q, nex, ney, L, L, CP, W)
    IEN, ID, Dir = topology(p, q, nex, ney)
    S, T = open_knots(p, nex, L), open_knots(q, ney, L)
    C1, C2 = extraction(p, S), extraction(q, T)
    h1 = np.full(nex, L / nex)
    h2 = np.full(ney, L / ney)
    return Mesh(p, q, nex, ney, CP, None, IEN, ID, Dir, C1, C2, h1, h2)


def mesh_quarter_annulus(p, nex, ney=None, r0=1.0, r1=2.0):
    ney = nex if ney is None else ney
    nfx, nfy = nex + p, ney + p
    CP = np.empty(2 * nfx * nfy)
    W = np.empty(nfx * nfy)
    rc = lib().orc_geometry_quarter_annulus(p, nex, ney, r0, r1, CP, W)
    if rc != 0:
        raise RuntimeError("annulus construction failed: %d" % rc)
    IEN, ID, Dir = topology(p, p, nex, ney)
    S, T = open_knots(p, nex), open_knots(p, ney)
    C1, C2 = extraction(p, S), extraction(p, T)
    return Mesh(p, p, nex, ney, CP, W, IEN, ID, Dir, C1, C2, np.full(nex, 1.0 / nex),
                np.full(ney, 1.0 / ney))


# --------------------------------------------------------------------------- operators
def element_qp(m, ex, ey):
    nloc = m.nloc
    nqp = m.nq1 * m.nq2
    e = ey * m.nex + ex
    ien = m.IEN[e * nloc:(e + 1) * nloc]
    eCP = np.ascontiguousarray(m.CP.reshape(-1, 2)[ien].reshape(-1))
    eW = None if m.W is None else np.ascontiguousarray(m.W[ien])
    R = np.empty(nqp * nloc); dx = np.empty(nqp * nloc); dy = np.empty(nqp * nloc)
    JxW = np.empty(nqp); xq = np.empty(2 * nqp)
    C1e = np.ascontiguousarray(m.C1.reshape(m.nex, -1)[ex])
    C2e = np.ascontiguousarray(m.C2.reshape(m.ney, -1)[ey])
    lib().orc_element_qp(m.p, m.q, m.nq1, m.nq2, C1e, C2e, float(m.h1[ex]), float(m.h2[ey]),
                         eCP, None if eW is None else eW.ctypes.data, R, dx, dy, JxW,
                         xq.ctypes.data)
    return (R.reshape(nqp, nloc), dx.reshape(nqp, nloc), dy.reshape(nqp, nloc), JxW,
            xq.reshape(nqp, 2))


def apply_mf(m, x):
    y = np.empty(m.nfunc)
    s = m.struct()
    lib().orc_apply_mf(C.byref(s), np.ascontiguousarray(x, np.float64), y)
    return y


def apply_mf_colored(m, x, nthreads=0):
    y = np.empty(m.nfunc)
    s = m.struct()
    used = lib().orc_apply_mf_colored(C.byref(s), np.ascontiguousarray(x, np.float64), y,
                                      nthreads)
    return y, used


def apply_sf_ref(m, x):
    y = np.empty(m.nfunc)
    s = m.struct()
    if lib().orc_apply_sf_ref(C.byref(s), np.ascontiguousarray(x, np.float64), y) != 0:
        raise ValueError("mult_sf semantics need p == q and nq1 == nq2")
    return y


LOAD_REF_POLY, LOAD_CONST, LOAD_SINSIN = 0, 1, 2


def load(m, kind=LOAD_REF_POLY, value=0.0):
    F = np.empty(m.nfunc)
    s = m.struct()
    lib().orc_load(C.byref(s), kind, value, F)
    return F


def load_qp(m, fq):
    """Load vector of f given at the quadrature points ([E][nqp], the
    order of element_qp)."""
    F = np.empty(m.nfunc)
    s = m.struct()
    lib().orc_load_qp(C.byref(s), np.ascontiguousarray(fq, np.float64).reshape(-1), F)
    return F


def kappa_grad(m, u, g):
    out = np.empty(m.nelem * m.nq1 * m.nq2)
    s = m.struct()
    lib().orc_kappa_grad(C.byref(s), np.ascontiguousarray(u, np.float64),
                         np.ascontiguousarray(g, np.float64), out)
    return out


def element_matrices(m):
    n = m.nelem * m.nloc * m.nloc
    Ke = np.empty(n)
    s = m.struct()
    lib().orc_element_matrices(C.byref(s), n, Ke)
    return Ke.reshape(m.nelem, m.nloc, m.nloc)


def assembled_matrix(m):
    """Assembled A = P_int * K~ (rows of Dirichlet functions empty) as scipy CSR."""
    import scipy.sparse as sp
    Ke = element_matrices(m)
    ien = m.IEN.reshape(m.nelem, m.nloc)
    rows = np.repeat(ien, m.nloc, axis=1).reshape(-1)
    cols = np.tile(ien, (1, m.nloc)).reshape(-1)
    rid = m.ID[rows]
    keep = rid >= 0
    A = sp.coo_matrix((Ke.reshape(-1)[keep], (rid[keep], cols[keep])),
                      shape=(m.nfunc, m.nfunc)).tocsr()
    A.sum_duplicates()
    return A


# --------------------------------------------------------------------------- FEM-PC
def fem_nodes(p, q, nex, ney, Lx=1.0, Ly=1.0):
    nfx, nfy = nex + p, ney + q
    CPf = np.empty(2 * nfx * nfy)
    if lib().orc_fem_nodes(p, q, nex, ney, Lx, Ly, CPf) < 0:
        raise ValueError("FEM preconditioner needs p, q >= 2")
    return CPf


def fem_assemble(nfx, nfy, CPf):
    cap = (3 * nfx) * (3 * nfy) + nfx * nfy
    rows = np.empty(cap, np.int32); cols = np.empty(cap, np.int32)
    nnz = lib().orc_fem_pattern(nfx, nfy, cap, rows, cols)
    rows, cols = rows[:nnz].copy(), cols[:nnz].copy()
    e2c = np.empty((nfx - 1) * (nfy - 1) * 16, np.int32)
    lib().orc_fem_elem2coo(nfx, nfy, rows, cols, nnz, e2c)
    d2c = np.empty(nfx * nfy, np.int32)
    nd = lib().orc_fem_dir2coo(nfx, nfy, rows, nnz, d2c)
    d2c = d2c[:nd].copy()
    val = np.empty(nnz)
    lib().orc_fem_values(nfx, nfy, np.ascontiguousarray(CPf, np.float64), e2c, d2c, nd,
                         nnz, val)
    return rows, cols, e2c, d2c, val


def invlm(nloc, nelem, nfunc, IEN):
    en = np.empty(nfunc, np.int32); off = np.empty(nfunc, np.int32)
    tot = nelem * nloc
    ei = np.empty(tot, np.int32); bi = np.empty(tot, np.int32)
    lib().orc_invlm(nloc, nelem, nfunc, np.ascontiguousarray(IEN, np.int32), en, off, ei, bi)
    return en, off, ei, bi


# --------------------------------------------------------------------------- ILU(0)
def ilu0(rowptr, col, val):
    """ILU(0) in natural ordering, the factorisation of PETSc's PCILU with 0
    levels of fill (PETSc 3.18.6, matrix_free/CMakeLists.txt:55; selected by
    the plain-method runs' -pc_type ilu): for each row i and each k < i in its
    pattern, l_ik = a_ik / u_kk, then a_ij -= l_ik u_kj for every j > k that
    both rows hold.  Pure Python loops: small matrices only.  Returns the
    factor values in the matrix's pattern (unit L below the diagonal, U on and
    above it) and the diagonal positions."""
    rowptr = np.asarray(rowptr); col = np.asarray(col)
    a = np.array(val, dtype=np.float64)
    n = len(rowptr) - 1
    diag = np.full(n, -1, np.int64)
    pos = []
    for i in range(n):
        d = {int(col[e]): e for e in range(rowptr[i], rowptr[i + 1])}
        pos.append(d)
        diag[i] = d[i]
    for i in range(n):
        for e in range(rowptr[i], diag[i]):
            k = int(col[e])
            a[e] /= a[diag[k]]
            for f in range(diag[k] + 1, rowptr[k + 1]):
                g = pos[i].get(int(col[f]))
                if g is not None:
                    a[g] -= a[e] * a[f]
    return a, diag


def ilu0_solve(rowptr, col, lu, diag, r):
    """z = (L U)^-1 r for the factors of ilu0 (sequential triangular solves)."""
    n = len(rowptr) - 1
    z = np.array(r, dtype=np.float64)
    for i in range(n):
        z[i] -= sum(lu[e] * z[col[e]] for e in range(rowptr[i], diag[i]))
    for i in range(n - 1, -1, -1):
        z[i] = (z[i] - sum(lu[e] * z[col[e]] for e in range(diag[i] + 1, rowptr[i + 1]))) \
            / lu[diag[i]]
    return z


# --------------------------------------------------------------------------- reference units
_ref_mf = None
_ref_gpu = None


def ref_available():
    return (os.path.exists(os.path.join(HERE, "_ref", "libref_mf.so")) and
            os.path.exists(os.path.join(HERE, "_ref", "libref_gpu.so")))


def ref_mf():
    global _ref_mf
    if _ref_mf is None:
        _ref_mf = C.CDLL(os.path.join(HERE, "_ref", "libref_mf.so"))
        L = _ref_mf
        L.ref_quadrature.argtypes = [C.c_int, _F64P, _F64P]
        L.ref_bernstein_table.argtypes = [C.c_int, C.c_int, _F64P, _F64P]
        L.ref_open_knots.argtypes = [C.c_int, C.c_int, C.c_double, _F64P]
        L.ref_extraction.argtypes = [C.c_int, C.c_int, C.c_double, _F64P]
        L.ref_bspline_basis.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _F64P]
        L.ref_greville.argtypes = [C.c_int, C.c_int, C.c_double, _F64P]
        L.ref_topology.argtypes = [C.c_int] * 4 + [_I32P] * 3
        L.ref_element_qp.argtypes = [C.c_int, C.c_int, _F64P, _F64P, C.c_double, C.c_double,
                                     _F64P, _F64P, _F64P, _F64P, _F64P]
        mesh_args = [C.c_int] * 5 + [_F64P, _I32P, _I32P, _I32P, C.c_int] + [_F64P] * 4
        L.ref_apply_mf.argtypes = mesh_args + [_F64P, _F64P]
        L.ref_apply_sf.argtypes = mesh_args + [_F64P, _F64P]
        L.ref_load.argtypes = mesh_args + [_F64P]
        L.ref_write_partition.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _F64P, _F64P,
                                          _F64P, _I32P, _I32P, C.c_int, _I32P, C.c_int, _F64P,
                                          C.c_int, _F64P, C.c_int]
        L.ref_read_partition.argtypes = [C.c_char_p, _I32P, _F64P, _F64P, _F64P, _I32P, _I32P,
                                         _I32P, _F64P, _F64P, C.c_int]
    return _ref_mf


def ref_gpu():
    global _ref_gpu
    if _ref_gpu is None:
        _ref_gpu = C.CDLL(os.path.join(HERE, "_ref", "libref_gpu.so"))
        L = _ref_gpu
        L.ref_fem_nodes.argtypes = [C.c_int] * 4 + [C.c_double] * 2 + [_F64P]
        L.ref_fem_assemble.argtypes = [C.c_int] * 4 + [C.c_double] * 2 + [
            C.c_long, _I32P, _I32P, _I32P, _I32P, _F64P, C.POINTER(C.c_int)]
        L.ref_fem_assemble.restype = C.c_long
        L.ref_invlm.argtypes = [C.c_int] * 3 + [_I32P] * 5
        L.ref_invlm.restype = C.c_long
        L.ref_fem_write_partitions.argtypes = [C.c_char_p] + [C.c_int] * 5 + [_I32P]
        L.ref_fem_read_partition.argtypes = [C.c_char_p, C.c_long, _I32P, _I32P, _I32P,
                                             _I32P, _F64P]
        L.ref_write_nonzero.argtypes = [C.c_char_p, C.c_int, _I32P, _I32P]
        L.ref_write_nonzero.restype = None
        L.ref_read_nonzero.argtypes = [C.c_char_p, C.c_long, _I32P, _I32P]
        L.ref_read_nonzero.restype = C.c_long
        L.ref_write_new_to_old.argtypes = [C.c_char_p, C.c_int, _I32P]
        L.ref_write_new_to_old.restype = None
        L.ref_read_new_to_old.argtypes = [C.c_char_p, C.c_long, _I32P]
        L.ref_read_new_to_old.restype = C.c_long
    return _ref_gpu


def _mesh_args(m):
    return (m.p, m.q, m.nex, m.ney, m.nfunc, m.CP, m.IEN, m.ID, m.Dir, len(m.Dir), m.C1,
            m.C2, m.h1, m.h2)


def ref_apply_mf(m, x):
    y = np.empty(m.nfunc)
    ref_mf().ref_apply_mf(*_mesh_args(m), np.ascontiguousarray(x, np.float64), y)
    return y


def ref_apply_sf(m, x):
    y = np.empty(m.nfunc)
    ref_mf().ref_apply_sf(*_mesh_args(m), np.ascontiguousarray(x, np.float64), y)
    return y


def ref_load(m):
    F = np.empty(m.nfunc)
    ref_mf().ref_load(*_mesh_args(m), F)
    return F


def ref_write_partition(fn, a):
    """The reference's FileManager::WritePartition (full form) of mesh arrays a
    (as igamf.Mesh.arrays())."""
    L = ref_mf()
    nf = len(a["ID"])
    i32 = lambda v: np.ascontiguousarray(v, np.int32)
    f64 = lambda v: np.ascontiguousarray(v, np.float64)
    L.ref_write_partition(os.fsencode(fn), nf, a["nex"], a["ney"], f64(a["h1"]), f64(a["h2"]),
                          f64(a["CP"]), i32(a["ID"]), i32(a["Dir"]), len(a["Dir"]),
                          i32(a["IEN"]), len(a["IEN"]), f64(a["C1"]), len(a["C1"]),
                          f64(a["C2"]), len(a["C2"]))


def ref_read_partition(fn, cap):
    """The reference's FileManager::ReadPartition (full form) -> dict of arrays."""
    L = ref_mf()
    cnt = np.zeros(11, np.int32)
    bufs = {k: np.zeros(cap, np.float64) for k in ("h1", "h2", "CP", "C1", "C2")}
    ibufs = {k: np.zeros(cap, np.int32) for k in ("ID", "Dir", "IEN")}
    rc = L.ref_read_partition(os.fsencode(fn), cnt, bufs["h1"], bufs["h2"], bufs["CP"],
                              ibufs["ID"], ibufs["Dir"], ibufs["IEN"], bufs["C1"], bufs["C2"],
                              cap)
    if rc != 0:
        raise ValueError("capacity too small")
    names = ("h1", "h2", "CP", "ID", "Dir", "IEN", "C1", "C2")
    out = {"nfunc": int(cnt[0]), "nex": int(cnt[1]), "ney": int(cnt[2])}
    for k, n in zip(names, cnt[3:]):
        out[k] = (bufs.get(k) if k in bufs else ibufs[k])[:n].copy()
    return out


# --------------------------------------------------------------------------- PETSc KSP semantics
# The reference solves with PETSc 3.18.6 (matrix_free/CMakeLists.txt:55), a
# dependency absent here.  Its FEM-preconditioned drivers
# (matrix_free/iteration_number_fempc_mf.cpp:190-221,
# matrix_free_gpu/matrixfreegpu.cu:264-301) call KSPSetFromOptions on the outer
# KSP and never set its type, so it is PETSc's default KSPGMRES; the PCSHELL
# runs an inner KSPCG + PCJACOBI on the Q1 matrix; both get rtol = abstol =
# 1e-10.  The two functions below restate the published algorithms with the
# defaults of that release: zero initial guess, KSP_NORM_PRECONDITIONED, and
# KSPConvergedDefault (stop when the norm is <= max(rtol * norm_0, abstol));
# GMRES restarts every 30 iterations, left preconditioning, classical
# Gram-Schmidt without refinement, the residual norm estimated from the
# Givens-rotated Hessenberg system.

def petsc_cg(matvec, b, pc, rtol, atol, maxit=10000):
    """KSPCG with a preconditioner pc(r): (x, iterations)."""
    x = np.zeros_like(b)
    r = b.copy()
    z = pc(r)
    n0 = np.linalg.norm(z)
    if n0 <= max(rtol * n0, atol):
        return x, 0
    p = z.copy()
    rz = r @ z
    for it in range(1, maxit + 1):
        q = matvec(p)
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        z = pc(r)
        if np.linalg.norm(z) <= max(rtol * n0, atol):
            return x, it
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, maxit


def petsc_gmres(matvec, b, pc, rtol, atol, restart=30, maxit=10000):
    """KSPGMRES(restart), left preconditioning: (x, iterations)."""
    x = np.zeros_like(b)
    its, norm0 = 0, None
    while True:
        r = pc(b - matvec(x))
        beta = np.linalg.norm(r)
        if norm0 is None:
            norm0 = beta
        if beta <= max(rtol * norm0, atol) or its >= maxit:
            return x, its
        V = np.zeros((restart + 1, len(b)))
        V[0] = r / beta
        H = np.zeros((restart + 1, restart))
        g = np.zeros(restart + 1)
        g[0] = beta
        cs, sn = np.zeros(restart), np.zeros(restart)
        k = 0
        for j in range(restart):
            w = pc(matvec(V[j]))
            h = V[:j + 1] @ w              # classical Gram-Schmidt
            w = w - V[:j + 1].T @ h
            H[:j + 1, j] = h
            H[j + 1, j] = np.linalg.norm(w)
            for i in range(j):             # earlier rotations
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            d = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = H[j, j] / d, H[j + 1, j] / d
            H[j, j], H[j + 1, j] = d, 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] *= cs[j]
            its += 1
            k = j + 1
            if H[j, j] == 0.0 or abs(g[j + 1]) <= max(rtol * norm0, atol) or its >= maxit:
                break
            V[j + 1] = w / np.linalg.norm(w)
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
        x = x + V[:k].T @ y
        if abs(g[k]) <= max(rtol * norm0, atol) or its >= maxit:
            return x, its


def fempc_reference_iterations(p, n, exact_pc=False):
    """The reference's FEM-preconditioned solve of the unit square (f =
    x(1-x)y(1-y), matrix_free/include/LocalAssemblyMF.hpp:52-55) restated on
    the oracle's matrices: outer GMRES on the interior IGA matrix, PCSHELL =
    inner CG + Jacobi on the interior Q1 matrix of the reference's node grid,
    all tolerances 1e-10.  Returns the outer iteration count (exact_pc: the
    Q1 matrix factorised instead of the inner CG)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    m = mesh_unit_square(p, n)
    idx = np.where(m.ID >= 0)[0]
    A = assembled_matrix(m)[idx][:, idx].tocsr()
    F = load(m, LOAD_REF_POLY)[idx]
    nf = n + p
    rows, cols, _, _, val = fem_assemble(nf, nf, fem_nodes(p, p, n, n))
    K = sp.coo_matrix((val, (rows, cols)), shape=(nf * nf, nf * nf)).tocsr()[idx][:, idx].tocsr()
    if exact_pc:
        lu = spla.splu(K.tocsc())
        pc = lu.solve
    else:
        dinv = 1.0 / K.diagonal()
        pc = lambda v: petsc_cg(lambda q: K @ q, v, lambda r: dinv * r, 1e-10, 1e-10)[0]
    _, its = petsc_gmres(lambda v: A @ v, F, pc, 1e-10, 1e-10)
    return its
