"""A Dirichlet set other than the boundary ring (left and bottom edges only,
natural conditions on the others): the operator then carries an explicit
per-function mask, which every kernel path honours — the apply variants and
both assemblies against the oracle, the seam kernel, the CG update kernels
that sum the gather apply's seam functions, and the solves against a direct
solve of the oracle's assembled matrix.

Bar: as the other GPU parity tests (FP64 1e-12, FP32 2e-5 relative,
max-norm); solutions within 1e-8 of the direct solve.
"""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
spla = pytest.importorskip("scipy.sparse.linalg")
pytestmark = pytest.mark.gpu

import igamf  # noqa: E402


def mixed_mesh(kind, p, nex, ney):
    """The generator's mesh with Dirichlet rows on x = 0 and y = 0 only."""
    a = igamf.Mesh.build(kind, p, nex, ney).arrays()
    nfx = nex + p
    nf = a["ID"].size
    gx, gy = np.arange(nf) % nfx, np.arange(nf) // nfx
    dirs = (gx == 0) | (gy == 0)
    ID = np.where(dirs, -1, np.arange(nf))  # survivors keep their own number
    Dir = np.nonzero(dirs)[0]
    m = igamf.Mesh.from_arrays(a["p"], a["q"], a["nex"], a["ney"], a["CP"], a["W"], a["IEN"],
                               ID, Dir, a["C1"], a["C2"], a["h1"], a["h2"])
    om = O.Mesh(a["p"], a["q"], a["nex"], a["ney"], a["CP"], a["W"], a["IEN"], ID, Dir,
                a["C1"], a["C2"], a["h1"], a["h2"])
    return m, om


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [("annulus_p3_45x38", (igamf.GEOM_QUARTER_ANNULUS, 3, 45, 38)),
         ("square_p2_33x20", (igamf.GEOM_UNIT_SQUARE, 2, 33, 20))]


@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_mixed_dirichlet_apply(name, args):
    m, om = mixed_mesh(*args)
    x = igamf.rng_normal(3, om.nfunc)
    y_ref = O.apply_mf(om, x)
    xd = torch.from_numpy(x).cuda()
    variants = [0, 3] + ([4] if args[1] == 3 else []) + ([5] if args[1] <= 3 else [])
    for v in variants:
        op = igamf.Operator(m, variant_f64=3 if v == 3 else 0, variant_f32=v)
        for kern in (igamf.SF, igamf.EBE):
            for asm in (igamf.IEN_ATOMIC, igamf.INVLM_GATHER):
                y64 = op.apply(xd, kernel=kern, assembly=asm).cpu().numpy()
                assert rel(y64, y_ref) < 1e-12, (name, v, kern, asm, rel(y64, y_ref))
                y32 = op.apply(xd.float(), kernel=kern, assembly=asm).double().cpu().numpy()
                assert rel(y32, y_ref) < 2e-5, (name, v, kern, asm, rel(y32, y_ref))
                assert np.all(y64[om.Dir] == 0.0) and np.all(y32[om.Dir] == 0.0)
                # natural-boundary rows are not zeroed
                nat = (om.ID >= 0) & (np.abs(y_ref) > 1e-8 * np.abs(y_ref).max())
                assert nat.sum() > 0 and np.all(y64[nat] != 0.0)


@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_mixed_dirichlet_solves(name, args):
    m, om = mixed_mesh(*args)
    op = igamf.Operator(m)
    F = op.load(igamf.LOAD_CONST, 1.0)
    Fh = F.cpu().numpy()
    A = O.assembled_matrix(om)
    I = np.where(om.ID >= 0)[0]
    u_ref = np.zeros(om.nfunc)
    u_ref[I] = spla.spsolve(A[I][:, I].tocsc(), Fh[I])
    for skind in (igamf.SOLVE_PCG_F64, igamf.SOLVE_PCG_F32_REFINE):
        for asm in (igamf.INVLM_GATHER, igamf.IEN_ATOMIC):
            for fused in (0, 1, 2):
                s = igamf.Solver(op, None, kind=skind, pc=igamf.PC_JACOBI, rtol=1e-11,
                                 assembly=asm, fused=fused)
                u, st = s.solve(F)
                assert st.status == 0, (name, skind, asm, fused, st.status)
                e = rel(u.cpu().numpy(), u_ref)
                assert e < 1e-8, (name, skind, asm, fused, e)
