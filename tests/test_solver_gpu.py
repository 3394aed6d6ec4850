"""GPU tests of the FEM-preconditioner assembly and of the solver.

FEM assembly: indices (COO rows/cols, CSR, elem2coo, dir2coo) bit-exact
against the oracle and the reference golden vectors, values within 1e-13.
Solver: converged solutions against a direct solve of the oracle's assembled
FP64 matrix; FP32 + FP64 refinement against plain FP64; graph replay equal to
eager launches; iteration-limit and invalid-option paths.
"""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import igamf  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def fem_arrays(f):
    t = f.to_torch()
    return {k: v.cpu().numpy() for k, v in t.items()}


@pytest.mark.parametrize("p,n", [(3, 16), (4, 12), (2, 20), (3, 37)])
def test_fem_assembly_matches_oracle(p, n):
    nf = n + p
    nodes = igamf.fem_reference_nodes(p, p, n, n)
    assert np.abs(nodes - O.fem_nodes(p, p, n, n)).max() < 1e-15
    f = igamf.FemPC(nf, nf, nodes)
    a = fem_arrays(f)
    rows, cols, e2c, d2c, val = O.fem_assemble(nf, nf, nodes)
    assert f.view.nnz == len(rows)
    assert np.array_equal(a["coo_row"], rows)
    assert np.array_equal(a["col"], cols)
    assert np.array_equal(a["elem2coo"], e2c)
    assert np.array_equal(a["dir2coo"], d2c)
    assert np.abs(a["val"] - val).max() < 1e-13 * np.abs(val).max()
    rp = a["rowptr"]
    assert rp[0] == 0 and rp[-1] == len(rows)
    assert np.array_equal(np.repeat(np.arange(nf * nf), np.diff(rp)), rows)


@pytest.mark.parametrize("name", ["fem_p3_n16.npz", "fem_p4_n12.npz"])
def test_fem_assembly_matches_reference_golden(name):
    g = np.load(os.path.join(GOLD, name))
    p, n = int(g["p"]), int(g["nelem"])
    f = igamf.FemPC(n + p, n + p, g["nodes"])
    a = fem_arrays(f)
    assert np.array_equal(a["coo_row"], g["rows"])
    assert np.array_equal(a["col"], g["cols"])
    assert np.array_equal(a["elem2coo"], g["elem2coo"])
    assert np.array_equal(a["dir2coo"], g["dir2coo"])
    assert np.abs(a["val"] - g["val"]).max() < 1e-13


def test_fem_assembly_is_deterministic():
    p, n = 3, 64
    nodes = igamf.fem_reference_nodes(p, p, n, n)
    v0 = fem_arrays(igamf.FemPC(n + p, n + p, nodes))["val"]
    v1 = fem_arrays(igamf.FemPC(n + p, n + p, nodes))["val"]
    assert np.array_equal(v0, v1)


def _direct(m, om, F):
    A = O.assembled_matrix(om)
    I = np.where(om.ID >= 0)[0]
    u = np.zeros(om.nfunc)
    u[I] = spla.spsolve(A[I][:, I].tocsc(), F[I])
    return u


def _oracle_mesh(m):
    a = m.arrays()
    return O.Mesh(a["p"], a["q"], a["nex"], a["ney"], a["CP"], a["W"], a["IEN"], a["ID"],
                  a["Dir"], a["C1"], a["C2"], a["h1"], a["h2"])


@pytest.mark.parametrize("geom", ["square", "annulus", "rational"])
def test_pcg_solves(geom):
    if geom == "square":
        m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 2, 32)
        kind, val = igamf.LOAD_SINSIN, 2 * np.pi ** 2
    elif geom == "rational":
        m = igamf.Mesh.build(igamf.GEOM_PERTURBED_SQUARE, 3, 20, perturb=0.1, w_lo=0.8,
                             w_hi=1.2, seed=9)
        kind, val = igamf.LOAD_SINSIN, 2 * np.pi ** 2
    else:
        m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 24)
        kind, val = igamf.LOAD_CONST, 1.0
    om = _oracle_mesh(m)
    op = igamf.Operator(m)
    F = op.load(kind, val)
    u_ref = _direct(m, om, F.cpu().numpy())
    for skind in (igamf.SOLVE_PCG_F64, igamf.SOLVE_PCG_F32_REFINE):
        for asm in (igamf.INVLM_GATHER, igamf.IEN_ATOMIC):
            s = igamf.Solver(op, None, kind=skind, pc=igamf.PC_JACOBI, rtol=1e-10, assembly=asm)
            u, st = s.solve(F)
            assert st.status == 0 and st.rel_residual <= 1e-10 * 1.0001, (skind, st.rel_residual)
            err = np.abs(u.cpu().numpy() - u_ref).max() / np.abs(u_ref).max()
            assert err < 1e-7, (skind, asm, err)
    if geom == "square":
        # reference sign: u_h ~ -sin(pi x) sin(pi y)
        nf = 34
        assert u_ref[(nf // 2) * nf + nf // 2] < -0.9


def test_graph_replay_equals_eager():
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 40)
    op = igamf.Operator(m)
    F = op.load(igamf.LOAD_CONST, 1.0)
    for skind, fused in [(k, f) for k in (igamf.SOLVE_PCG_F64, igamf.SOLVE_PCG_F32_REFINE)
                         for f in (0, 1, 2)]:
        outs = []
        for g in (0, 1, 1):
            s = igamf.Solver(op, None, kind=skind, pc=igamf.PC_JACOBI, rtol=1e-9, use_graph=g,
                             fused=fused)
            u, st = s.solve(F)
            outs.append((u.clone(), st.iterations))
            u2, st2 = s.solve(F)  # replay the same solver
            assert torch.equal(u, u2)
        assert all(torch.equal(outs[0][0], o[0]) and outs[0][1] == o[1] for o in outs[1:])


@pytest.mark.parametrize("geom", ["annulus", "rational"])
def test_fused_iteration_matches_classic(geom):
    """p.Ap accumulated inside the apply (fused=1, the default) and the
    single-reduction iteration with u.Au inside the apply (fused=2) against
    the iteration with a separate p.Ap kernel (fused=0):
    same converged solution, iteration counts within 5 % (FP64) / 12 % (FP32
    inner recursion; 35 % for its single-reduction form), for both solve
    kinds, both assemblies and both local
    preconditioners; an operator whose apply variant cannot fuse falls back
    to the separate kernel."""
    if geom == "rational":
        m = igamf.Mesh.build(igamf.GEOM_PERTURBED_SQUARE, 3, 33, 27, perturb=0.1, w_lo=0.8,
                             w_hi=1.2, seed=11)
    else:
        m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 45, 38)
    op = igamf.Operator(m)
    F = op.load(igamf.LOAD_CONST, 1.0)
    for skind in (igamf.SOLVE_PCG_F64, igamf.SOLVE_PCG_F32_REFINE):
        for pc in (igamf.PC_JACOBI, igamf.PC_NONE):
            for asm in (igamf.INVLM_GATHER, igamf.IEN_ATOMIC):
                res = {}
                for fused in (0, 1, 2):
                    s = igamf.Solver(op, None, kind=skind, pc=pc, rtol=1e-10, assembly=asm,
                                     fused=fused)
                    u, st = s.solve(F)
                    assert st.status == 0 and st.rel_residual <= 1.0001e-10, (fused, st)
                    res[fused] = (u, st.iterations)
                u0, i0 = res[0]
                for fused in (1, 2):
                    u1, i1 = res[fused]
                    d = ((u1 - u0).abs().max() / u0.abs().max()).item()
                    assert d < 1e-8, (skind, pc, asm, fused, d)
                    # FP32 recurrences (and the atomic assembly's summation
                    # order) move the count by a few percent either way; the
                    # single-reduction recurrence in FP32 needs up to about a
                    # third more (its s = A p drifts: 2,205 against 1,698 on
                    # 1024^2)
                    slack = 0.05 if skind == igamf.SOLVE_PCG_F64 else (0.12 if fused == 1 else 0.35)
                    assert abs(i1 - i0) <= max(3, slack * i0), (skind, pc, asm, fused, i0, i1)
    if geom == "annulus":
        op1 = igamf.Operator(m, variant_f64=1, variant_f32=1)
        s = igamf.Solver(op1, None, kind=igamf.SOLVE_PCG_F32_REFINE, rtol=1e-10)
        u, st = s.solve(F)
        assert st.status == 0 and st.rel_residual <= 1.0001e-10


def test_matrix_core_inner_solve():
    """FP32 inner solves on the matrix-core apply (variant 4, degree 3), on
    FP32 and FP16-stored G_q, with both assemblies and the fused iteration:
    the same converged solution as the default operator's FP64 PCG, and
    iteration counts within 12 % of the default FP32 kernel's."""
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 45, 38)
    op0 = igamf.Operator(m, f16_geometry=True)
    op4 = igamf.Operator(m, variant_f32=4, f16_geometry=True)
    F = op0.load(igamf.LOAD_CONST, 1.0)
    s = igamf.Solver(op0, None, kind=igamf.SOLVE_PCG_F64, rtol=1e-10)
    u64, st = s.solve(F)
    assert st.status == 0
    for g16 in (0, 1):
        for asm in (igamf.INVLM_GATHER, igamf.IEN_ATOMIC):
            its = []
            for op in (op0, op4):
                s = igamf.Solver(op, None, kind=igamf.SOLVE_PCG_F32_REFINE, rtol=1e-10,
                                 assembly=asm, inner_f16_geometry=g16)
                u, st = s.solve(F)
                assert st.status == 0 and st.rel_residual <= 1.0001e-10, (g16, asm, st)
                d = ((u - u64).abs().max() / u64.abs().max()).item()
                assert d < 1e-8, (g16, asm, d)
                its.append(st.iterations)
            assert abs(its[1] - its[0]) <= max(3, 0.12 * its[0]), (g16, asm, its)


def test_lds_staged_inner_solve():
    """FP32 inner solves on the LDS-staged persistent apply (variant 5), on
    FP32 and FP16-stored G_q, both assemblies, every iteration form: the FP64
    solution to 1e-8, and variant 0's iteration count within 3 % with the
    gather (the same element arithmetic and summation order) and 12 % with
    the scatter-add (whose summation order varies from run to run)."""
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 45, 38)
    op0 = igamf.Operator(m, variant_f32=0, f16_geometry=True)  # not the degree-3 default
    op5 = igamf.Operator(m, variant_f32=5, f16_geometry=True)
    F = op0.load(igamf.LOAD_CONST, 1.0)
    u64, st = igamf.Solver(op0, None, kind=igamf.SOLVE_PCG_F64, rtol=1e-10).solve(F)
    assert st.status == 0
    for g16 in (0, 1):
        for asm in (igamf.INVLM_GATHER, igamf.IEN_ATOMIC):
            for fused in (0, 1, 2):
                its = []
                for op in (op0, op5):
                    s = igamf.Solver(op, None, kind=igamf.SOLVE_PCG_F32_REFINE, rtol=1e-10,
                                     assembly=asm, inner_f16_geometry=g16, fused=fused)
                    u, st = s.solve(F)
                    assert st.status == 0 and st.rel_residual <= 1.0001e-10, (g16, asm, fused)
                    d = ((u - u64).abs().max() / u64.abs().max()).item()
                    assert d < 1e-8, (g16, asm, fused, d)
                    its.append(st.iterations)
                tol = 0.03 if asm == igamf.INVLM_GATHER else 0.12
                assert abs(its[1] - its[0]) <= max(3, tol * its[0]), (g16, asm, fused, its)


@pytest.mark.parametrize("asm", [igamf.INVLM_GATHER, igamf.IEN_ATOMIC])
def test_f16_geometry_inner_solve(asm):
    """FP32 inner PCG on the FP16-geometry operator under FP64 reliable
    updates: the FP64 residual removes the operator's 2^-11 perturbation, so
    the solve reaches rtol 1e-10 and matches the FP64 solve; the inner
    iteration count stays near the FP32-geometry one."""
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 64, 50)
    op = igamf.Operator(m, f16_geometry=True)
    F = op.load(igamf.LOAD_CONST, 1.0)
    s64 = igamf.Solver(op, None, kind=igamf.SOLVE_PCG_F64, rtol=1e-10, assembly=asm)
    u64, st64 = s64.solve(F)
    its = {}
    for g16 in (0, 1):
        for fused in (0, 1, 2):
            s = igamf.Solver(op, None, kind=igamf.SOLVE_PCG_F32_REFINE, rtol=1e-10, assembly=asm,
                             inner_f16_geometry=g16, fused=fused)
            u, st = s.solve(F)
            assert st.status == 0 and st.rel_residual <= 1.0001e-10, (g16, fused, st)
            d = ((u - u64).abs().max() / u64.abs().max()).item()
            assert d < 1e-8, (g16, fused, d)
            its[(g16, fused)] = st.iterations
    assert its[(1, 1)] <= 1.25 * its[(0, 1)], its
    with pytest.raises(igamf.IgaError):  # FP16 inner operator needs the FP32 inner solve
        igamf.Solver(op, None, kind=igamf.SOLVE_PCG_F64, inner_f16_geometry=1)
    with pytest.raises(igamf.IgaError):  # and an op that keeps the FP16 copy
        igamf.Solver(igamf.Operator(m), None, kind=igamf.SOLVE_PCG_F32_REFINE,
                     inner_f16_geometry=1)


def test_fem_preconditioned_solve():
    """The reference's FEM preconditioner (inner CG + Jacobi on the Q1 matrix):
    converged solutions agree with Jacobi-PCG, and its iteration count stays
    nearly flat under refinement while Jacobi's grows (rough right-hand side)."""
    its = {}
    for n in (32, 128):
        p = 3
        m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, p, n)
        op = igamf.Operator(m)
        fem = igamf.FemPC(n + p, n + p, igamf.fem_reference_nodes(p, p, n, n))
        mask = torch.from_numpy(m.arrays()["ID"] >= 0).cuda()
        F = torch.from_numpy(igamf.rng_normal(7, op.nfunc)).cuda() * mask
        sj = igamf.Solver(op, None, pc=igamf.PC_JACOBI, rtol=1e-10)
        uj, stj = sj.solve(F)
        its[("jacobi", n)] = stj.iterations
        for kind in (igamf.SOLVE_PCG_F64, igamf.SOLVE_PCG_F32_REFINE):
            sf = igamf.Solver(op, fem, kind=kind, pc=igamf.PC_FEM_CG, rtol=1e-10,
                              pc_inner_maxit=400, pc_inner_rtol=1e-8)
            uf, stf = sf.solve(F)
            assert stf.status == 0 and stf.rel_residual <= 1.0001e-10
            assert (uf - uj).abs().max().item() < 1e-7 * uj.abs().max().item()
            its[(kind, n)] = stf.iterations
    # kappa(D^-1 A) grows like h^-2 (about 64 at 32^2), kappa(K_fem^-1 A) stays near 98
    assert its[("jacobi", 128)] > 2 * its[("jacobi", 32)], its
    assert its[(igamf.SOLVE_PCG_F64, 128)] < 1.3 * its[(igamf.SOLVE_PCG_F64, 32)], its
    assert its[(igamf.SOLVE_PCG_F64, 128)] < its[("jacobi", 128)], its


def _pcg_count(A, F, prec, rtol):
    """Iterations of plain PCG (numpy, FP64, zero start) to ||r|| <= rtol ||F||."""
    x = np.zeros_like(F); r = F.copy(); z = prec(r); p = z.copy(); rz = r @ z
    nf = np.linalg.norm(F)
    for it in range(1, 10000):
        Ap = A @ p
        a = rz / (p @ Ap)
        x += a * p; r -= a * Ap
        if np.linalg.norm(r) <= rtol * nf:
            return it, x
        z = prec(r); rzn = r @ z; p = z + (rzn / rz) * p; rz = rzn
    return -1, x


def test_fem_ilu_preconditioned_solve():
    """ILU(0) of the Q1 FEM matrix as the preconditioner (the plain-method
    configuration: p = 2, 32^2, rtol 1e-10): both solve kinds converge to the
    direct solution, and the FP64 iteration count equals that of a numpy PCG
    with the oracle's ILU(0) of the oracle's FEM matrix (within 2)."""
    import scipy.sparse as sp
    p, n = 2, 32
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, p, n)
    om = _oracle_mesh(m)
    op = igamf.Operator(m)
    nodes = igamf.fem_reference_nodes(p, p, n, n)
    fem = igamf.FemPC(n + p, n + p, nodes)
    F = op.load(igamf.LOAD_SINSIN, 2 * np.pi ** 2)
    Fh = F.cpu().numpy()
    u_ref = _direct(m, om, Fh)
    its = {}
    for kind in (igamf.SOLVE_PCG_F64, igamf.SOLVE_PCG_F32_REFINE):
        s = igamf.Solver(op, fem, kind=kind, pc=igamf.PC_FEM_ILU, rtol=1e-10)
        u, st = s.solve(F)
        assert st.status == 0 and st.rel_residual <= 1.0001e-10, (kind, st)
        err = np.abs(u.cpu().numpy() - u_ref).max() / np.abs(u_ref).max()
        assert err < 1e-7, (kind, err)
        its[kind] = st.iterations
    nf = n + p
    rows, cols, _, _, val = O.fem_assemble(nf, nf, O.fem_nodes(p, p, n, n))
    K = sp.csr_matrix((val, (rows, cols)), shape=(nf * nf, nf * nf))
    K.sort_indices()
    lu, diag = O.ilu0(K.indptr, K.indices, K.data)
    A = O.assembled_matrix(om)
    it_ref, _ = _pcg_count(A, Fh, lambda r: O.ilu0_solve(K.indptr, K.indices, lu, diag, r),
                           1e-10)
    assert it_ref > 0 and abs(its[igamf.SOLVE_PCG_F64] - it_ref) <= 2, (its, it_ref)


def test_iteration_limit_and_bad_options():
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 2, 32)
    op = igamf.Operator(m)
    F = op.load(igamf.LOAD_REF_POLY)
    s = igamf.Solver(op, None, pc=igamf.PC_JACOBI, rtol=1e-12, maxit=3)
    with pytest.raises(igamf.IgaError) as e:
        s.solve(F)
    assert e.value.status == 11  # NOT_CONVERGED
    with pytest.raises(igamf.IgaError) as e:
        igamf.Solver(op, None, pc=igamf.PC_FEM_CG)  # FEM preconditioner without a matrix
    assert e.value.status == 1


def test_block_solve_matches_single_solves():
    """k systems in one block solve (FP32 PCG + FP64 reliable updates per
    column on the multi-RHS apply) against single FP64 solves; a zero column
    converges at once; a leading dimension above nfunc; unsupported kinds."""
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 41, 29)
    op = igamf.Operator(m)
    n, k = op.nfunc, 6
    mask = torch.from_numpy(m.arrays()["ID"] >= 0).cuda()
    ld = n + 5
    buf = torch.zeros(k, ld, dtype=torch.float64, device="cuda")
    for j in range(k - 1):
        buf[j, :n] = torch.from_numpy(igamf.rng_normal(40 + j, n)).cuda() * mask
    F = buf[:, :n]  # last column stays zero
    s = igamf.Solver(op, None, kind=igamf.SOLVE_PCG_F32_REFINE, pc=igamf.PC_JACOBI, rtol=1e-9)
    U, stats = s.solve_block(F, torch.full((k, ld), 5.0, dtype=torch.float64,
                                           device="cuda")[:, :n])
    s64 = igamf.Solver(op, None, kind=igamf.SOLVE_PCG_F64, pc=igamf.PC_JACOBI, rtol=1e-11)
    for j in range(k - 1):
        assert stats[j].status == 0 and stats[j].rel_residual <= 1e-9 * 1.0001, (j, stats[j])
        assert stats[j].outer_iterations >= 1
        u, _ = s64.solve(F[j].contiguous())
        assert (U[j] - u).abs().max().item() < 1e-7 * u.abs().max().item(), j
    assert stats[k - 1].status == 0 and stats[k - 1].iterations == 0
    assert U[k - 1].abs().max().item() == 0.0
    # a repeated block solve (graph replay) agrees to the tolerance; not bitwise,
    # since the multi-RHS apply adds seam sums with atomics
    U2, st2 = s.solve_block(F)
    assert all(x.status == 0 for x in st2)
    assert (U - U2).abs().max().item() < 1e-7 * U.abs().max().item()
    with pytest.raises(igamf.IgaError) as e:
        s64.solve_block(F)
    assert e.value.status == 6


def test_fem_chebyshev_preconditioned_solve():
    """The Chebyshev-Jacobi polynomial of the Q1 matrix as a fixed
    preconditioner: converged solutions equal the Jacobi-PCG ones and graph
    replays are bitwise equal; degree < 1 is rejected.  (How many iterations
    it saves depends on how well the parameter-grid Q1 matrix matches the
    geometry; on the quarter annulus Jacobi of A is the better of the two.)"""
    p, n = 3, 64
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, p, n)
    op = igamf.Operator(m)
    fem = igamf.FemPC(n + p, n + p, igamf.fem_reference_nodes(p, p, n, n))
    mask = torch.from_numpy(m.arrays()["ID"] >= 0).cuda()
    F = torch.from_numpy(igamf.rng_normal(17, op.nfunc)).cuda() * mask
    uj, stj = igamf.Solver(op, None, pc=igamf.PC_JACOBI, rtol=1e-10).solve(F)
    for kind in (igamf.SOLVE_PCG_F64, igamf.SOLVE_PCG_F32_REFINE):
        s = igamf.Solver(op, fem, kind=kind, pc=igamf.PC_FEM_CHEB, rtol=1e-10, pc_cheb_degree=8)
        u, st = s.solve(F)
        assert st.status == 0 and st.rel_residual <= 1.0001e-10, (kind, st.rel_residual)
        assert (u - uj).abs().max().item() < 1e-7 * uj.abs().max().item(), kind
        print("chebyshev-FEM iterations", kind, st.iterations, "jacobi", stj.iterations)
        u2, _ = s.solve(F)  # graph replay
        assert torch.equal(u, u2)
    with pytest.raises(igamf.IgaError) as e:
        igamf.Solver(op, fem, pc=igamf.PC_FEM_CHEB, pc_cheb_degree=0)
    assert e.value.status == 1


@pytest.mark.parametrize("fused", [0, 1])
def test_refine_solve_repeatable_on_full_vector_grid(fused):
    """A mesh whose vector kernels run the full grid of 512 workgroups: the
    FP32 + FP64 reliable-update solve (which stops its recursion with the
    p-update kernel's done flag) is bitwise repeatable and matches FP64."""
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 384)
    op = igamf.Operator(m)
    assert op.nfunc >= 512 * 256
    F = op.load(igamf.LOAD_CONST, 1.0)
    u64, _ = igamf.Solver(op, None, kind=igamf.SOLVE_PCG_F64, rtol=1e-9).solve(F)
    s = igamf.Solver(op, None, kind=igamf.SOLVE_PCG_F32_REFINE, rtol=1e-9, fused=fused)
    runs = [s.solve(F) for _ in range(3)]
    u0, st0 = runs[0]
    assert st0.status == 0 and st0.outer_iterations > 2, st0.outer_iterations
    for u, st in runs[1:]:
        assert torch.equal(u, u0) and st.iterations == st0.iterations
    d = ((u0 - u64).abs().max() / u64.abs().max()).item()
    assert d < 1e-7, d


# --------------------------------------------------------------------------- multigrid
def _fem_csr(nfx, nfy, nodes):
    import scipy.sparse as sp
    rows, cols, _, _, val = O.fem_assemble(nfx, nfy, nodes)
    return sp.csr_matrix((val, (rows, cols)), shape=(nfx * nfy, nfx * nfy))


def _annulus_with_fem(p, n, ney=None):
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, p, n, ney)
    d = m.desc
    return m, d.nelem_x + p, d.nelem_y + p, m.fem_nodes()


@pytest.mark.parametrize("geom", ["square", "annulus", "rect"])
def test_mg_vcycle_matches_cpu_model(geom):
    """One V-cycle of IGA_PC_FEM_MG (FP64 and FP32) against the scipy model
    of the same cycle, and symmetric: u.M^-1 v = v.M^-1 u.  The rectangle
    (37 x 150 nodes) coarsens its long direction alone at the end."""
    import mg_model
    p = 3
    if geom == "square":
        m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, p, 61)
        nfx = nfy = 64
        nodes = igamf.fem_reference_nodes(p, p, 61, 61)
    elif geom == "annulus":
        m, nfx, nfy, nodes = _annulus_with_fem(p, 70)
    else:
        m, nfx, nfy, nodes = _annulus_with_fem(p, 34, 147)
    op = igamf.Operator(m)
    fem = igamf.FemPC(nfx, nfy, nodes)
    model = mg_model.VCycle(_fem_csr(nfx, nfy, nodes), nfx, nfy)
    interior = m.arrays()["ID"] >= 0
    r = igamf.rng_normal(3, op.nfunc) * interior
    v = igamf.rng_normal(4, op.nfunc) * interior
    zref = model.apply(r)
    for kind, dt, tol in ((igamf.SOLVE_PCG_F64, torch.float64, 1e-11),
                          (igamf.SOLVE_PCG_F32_REFINE, torch.float32, 2e-5)):
        s = igamf.Solver(op, fem, kind=kind, pc=igamf.PC_FEM_MG)
        rt = torch.from_numpy(r).cuda().to(dt)
        vt = torch.from_numpy(v).cuda().to(dt)
        z = s.apply_pc(rt).double().cpu().numpy()
        err = np.abs(z - zref).max() / np.abs(zref).max()
        assert err < tol, (geom, dt, err)
        assert np.all(z[~interior] == 0.0)
        if dt == torch.float64:
            zv = s.apply_pc(vt).cpu().numpy()
            a, b = v @ z, r @ zv
            assert abs(a - b) <= 1e-12 * abs(a), (a, b)
            assert r @ z < 0  # K and so M^-1 are negative definite on the interior


@pytest.mark.parametrize("geom", ["square", "annulus"])
def test_mg_preconditioned_solve(geom):
    """PCG with one V-cycle on the Q1 matrix: FP64 and FP32 + FP64 reliable
    updates converge to the Jacobi-PCG solution, in an outer iteration count
    that stays flat under refinement (the exact K^-1's) while Jacobi's
    doubles."""
    p = 3
    its = {}
    for n in (48, 96):
        if geom == "square":
            m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, p, n)
            nodes = igamf.fem_reference_nodes(p, p, n, n)
            kind, val = igamf.LOAD_REF_POLY, 0.0
        else:
            m, _, _, nodes = _annulus_with_fem(p, n)
            kind, val = igamf.LOAD_CONST, 1.0
        op = igamf.Operator(m)
        fem = igamf.FemPC(n + p, n + p, nodes)
        F = op.load(kind, val)
        uj, stj = igamf.Solver(op, None, pc=igamf.PC_JACOBI, rtol=1e-10).solve(F)
        its[("jacobi", n)] = stj.iterations
        for sk in (igamf.SOLVE_PCG_F64, igamf.SOLVE_PCG_F32_REFINE):
            s = igamf.Solver(op, fem, kind=sk, pc=igamf.PC_FEM_MG, rtol=1e-10)
            u, st = s.solve(F)
            assert st.status == 0 and st.rel_residual <= 1.0001e-10, (sk, st.rel_residual)
            d = ((u - uj).abs().max() / uj.abs().max()).item()
            assert d < 1e-8, (geom, sk, d)
            its[(sk, n)] = st.iterations
            u2, st2 = s.solve(F)  # graph replay: bitwise
            assert torch.equal(u, u2) and st2.iterations == st.iterations
    # Jacobi grows like h^-1 (smooth right-hand sides converge a little
    # faster); the V-cycle's count stays flat.  On the unit square Jacobi is
    # still ahead at these sizes (kappa(K^-1 A) is about 60 whatever h is);
    # on the annulus the V-cycle already wins.
    assert its[("jacobi", 96)] > 1.4 * its[("jacobi", 48)], its
    for sk in (igamf.SOLVE_PCG_F64, igamf.SOLVE_PCG_F32_REFINE):
        assert its[(sk, 96)] <= its[(sk, 48)] + 4, its
        if geom == "annulus":
            assert its[(sk, 96)] < 0.6 * its[("jacobi", 96)], its


def test_non_finite_rhs_and_divergence():
    """Failure paths forced from the host: a right-hand side holding a NaN is
    IGA_ERR_NOT_FINITE (single and block solves, before any iteration); a
    divergence tolerance below 1 is IGA_ERR_DIVERGED at the first iteration
    (the test ||r|| > divtol ||f|| of PETSc's KSPConvergedDefault, divtol 1e4
    by default); an iteration cap below convergence is IGA_ERR_NOT_CONVERGED."""
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 24)
    op = igamf.Operator(m)
    F = op.load(igamf.LOAD_CONST, 1.0)
    bad = F.clone()
    bad[100] = float("nan")
    for kind in (igamf.SOLVE_PCG_F64, igamf.SOLVE_PCG_F32_REFINE):
        s = igamf.Solver(op, None, kind=kind)
        with pytest.raises(igamf.IgaError) as e:
            s.solve(bad)
        assert e.value.status == igamf.IGA_ERR_NOT_FINITE
        _, st = s.solve(bad, check_status=False)
        assert st.status == igamf.IGA_ERR_NOT_FINITE and st.iterations == 0
        # divtol below 1: the first iterate's residual already exceeds divtol ||f||
        sd = igamf.Solver(op, None, kind=kind, divtol=1e-3)
        _, st = sd.solve(F, check_status=False)
        assert st.status == igamf.IGA_ERR_DIVERGED, st.status
        sn = igamf.Solver(op, None, kind=kind, maxit=5)
        _, st = sn.solve(F, check_status=False)
        assert st.status == igamf.IGA_ERR_NOT_CONVERGED, st.status
    sb = igamf.Solver(op, None, kind=igamf.SOLVE_PCG_F32_REFINE)
    Fk = torch.stack([F, bad]).contiguous()
    _, sts = sb.solve_block(Fk, check_status=False)
    assert sts[0].status == 0 and sts[1].status == igamf.IGA_ERR_NOT_FINITE


def test_fempc_iterations_under_reference_ksp_semantics():
    """The reference's solver set-up driven through this library's entry
    points: PETSc's default outer KSPGMRES (igamf.ksp_gmres) on the MATSHELL
    apply (Operator.apply, FP64) with the PCSHELL apply (Solver.apply_pc:
    inner CG + Jacobi on the Q1 matrix to rtol = abstol = 1e-10), unit
    square p = 3, 32 x 32, f = x(1-x)y(1-y).  It takes the iteration count
    of the oracle's restatement on the assembled matrices (33; the
    reference logged 35, matrix_free/data/fempc_mf.txt:12) and reaches the
    direct solution."""
    p, n = 3, 32
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, p, n)
    op = igamf.Operator(m)
    fem = igamf.FemPC(n + p, n + p, igamf.fem_reference_nodes(p, p, n, n))
    F = op.load(igamf.LOAD_REF_POLY, 0.0)
    om = _oracle_mesh(m)
    dirs = torch.from_numpy(om.Dir.astype(np.int64)).cuda()
    assert F[dirs].abs().max().item() == 0.0
    s = igamf.Solver(op, fem, pc=igamf.PC_FEM_CG, pc_inner_rtol=1e-10, pc_inner_abstol=1e-10,
                     pc_inner_maxit=400)
    u, its = igamf.ksp_gmres(lambda v: op.apply(v.contiguous()), F, lambda v: s.apply_pc(v.contiguous()))
    its_oracle = O.fempc_reference_iterations(p, n)
    assert abs(its - its_oracle) <= 1, (its, its_oracle)
    assert abs(its - 35) <= 3, its
    u_ref = _direct(m, om, F.cpu().numpy())
    err = np.abs(u.cpu().numpy() - u_ref).max() / np.abs(u_ref).max()
    assert err < 1e-7, err
