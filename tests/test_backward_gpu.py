"""GPU tests of the kappa-weighted operator, its backward kernel and the
torch.autograd bindings: against the oracle (which is pinned to the reference
at kappa = 1 and to finite differences of its own apply), against central
finite differences of the GPU forward, and against the assembled FP64 matrix."""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import igamf  # noqa: E402
from igamf import autograd as IA  # noqa: E402


def _om(m, kappa=None):
    a = m.arrays()
    return O.Mesh(a["p"], a["q"], a["nex"], a["ney"], a["CP"], a["W"], a["IEN"], a["ID"],
                  a["Dir"], a["C1"], a["C2"], a["h1"], a["h2"], kappa=kappa)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def kappa_field(m, seed):
    """kappa_q = exp(0.5 s(x_q)), s a sum of 16 Fourier modes
    sin(a pi x) sin(b pi y), a, b = 1..4, with N(0, 1/(a^2+b^2)) coefficients
    (splitmix64 normals of the seed, a-major); x_q from the oracle's element
    geometry.  The benchmark builds the same field on the device."""
    om = _om(m)
    z = igamf.rng_normal(seed, 16)
    coef = {(a, b): z[(a - 1) * 4 + b - 1] / np.sqrt(a * a + b * b) for a in range(1, 5)
            for b in range(1, 5)}
    out = []
    for ey in range(om.ney):
        for ex in range(om.nex):
            _, _, _, _, xq = O.element_qp(om, ex, ey)
            s = sum(c * np.sin(a * np.pi * xq[:, 0]) * np.sin(b * np.pi * xq[:, 1])
                    for (a, b), c in coef.items())
            out.append(np.exp(0.5 * s))
    return np.concatenate(out)


@pytest.fixture(scope="module")
def case():
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 3, 20)
    kappa = kappa_field(m, 1)
    return m, kappa


def test_kappa_apply_and_backward_match_oracle(case):
    m, kappa = case
    op = igamf.Operator(m, kappa=kappa)
    om = _om(m, kappa)
    x = igamf.rng_normal(3, op.nfunc)
    g = igamf.rng_normal(4, op.nfunc)
    xd, gd = torch.from_numpy(x).cuda(), torch.from_numpy(g).cuda()
    y = op.apply(xd).cpu().numpy()
    assert rel(y, O.apply_mf(om, x)) < 1e-12
    y32 = op.apply(xd.float()).double().cpu().numpy()
    assert rel(y32, O.apply_mf(om, x)) < 2e-5
    gx, gk = op.backward(xd, gd)
    assert rel(gk.cpu().numpy(), O.kappa_grad(om, x, g)) < 1e-12
    # grad_x = K~ (P_int g): assembled full-row matrix
    Ke = O.element_matrices(om)
    ien = om.IEN.reshape(om.nelem, om.nloc)
    gm = g * (om.ID >= 0)
    ref = np.zeros(om.nfunc)
    for e in range(om.nelem):
        ref[ien[e]] += Ke[e] @ gm[ien[e]]
    assert rel(gx.cpu().numpy(), ref) < 1e-12
    # set_kappa path (rebuild G on the device) equals create-time kappa
    op2 = igamf.Operator(m)
    op2.set_kappa(torch.from_numpy(kappa).cuda())
    assert torch.equal(op2.apply(xd), op.apply(xd))


def test_kappa_backward_general_rational():
    """kappa, grad_x and grad_kappa on a mesh with per-function weights."""
    m = igamf.Mesh.build(igamf.GEOM_PERTURBED_SQUARE, 3, 14, perturb=0.1, w_lo=0.8, w_hi=1.2,
                         seed=8)
    kappa = kappa_field(m, 2)
    op = igamf.Operator(m, kappa=kappa)
    assert op.info.rational == 2
    om = _om(m, kappa)
    x = igamf.rng_normal(13, op.nfunc)
    g = igamf.rng_normal(14, op.nfunc)
    xd, gd = torch.from_numpy(x).cuda(), torch.from_numpy(g).cuda()
    assert rel(op.apply(xd).cpu().numpy(), O.apply_mf(om, x)) < 1e-12
    gx, gk = op.backward(xd, gd)
    assert rel(gk.cpu().numpy(), O.kappa_grad(om, x, g)) < 1e-12
    Ke = O.element_matrices(om)
    ien = om.IEN.reshape(om.nelem, om.nloc)
    gm = g * (om.ID >= 0)
    ref = np.zeros(om.nfunc)
    for e in range(om.nelem):
        ref[ien[e]] += Ke[e] @ gm[ien[e]]
    assert rel(gx.cpu().numpy(), ref) < 1e-12


def test_autograd_apply_gradcheck(case):
    m, kappa = case
    op = igamf.Operator(m)
    x = torch.from_numpy(igamf.rng_normal(5, op.nfunc)).cuda().requires_grad_()
    k = torch.from_numpy(kappa).cuda().requires_grad_()
    w = torch.from_numpy(igamf.rng_normal(6, op.nfunc)).cuda()
    loss = (IA.apply_kappa(op, x, k) * w).sum()
    loss.backward()
    # central differences along random directions
    for seed in range(4):
        dx = torch.from_numpy(igamf.rng_normal(100 + seed, op.nfunc)).cuda()
        dk = torch.from_numpy(igamf.rng_normal(200 + seed, k.numel())).cuda()
        h = 1e-6
        with torch.no_grad():
            fp = (IA.apply_kappa(op, x + h * dx, k + h * dk) * w).sum().item()
            fm = (IA.apply_kappa(op, x - h * dx, k - h * dk) * w).sum().item()
        fd = (fp - fm) / (2 * h)
        an = (x.grad * dx).sum().item() + (k.grad * dk).sum().item()
        assert abs(fd - an) < 1e-6 * max(abs(an), 1.0), (fd, an)


def test_adjoint_gradient_of_solve(case):
    """J = 1/2 |u(kappa) - u_obs|^2 with A(kappa) u = f: adjoint gradient
    against FP64 central differences (step 1e-6) along 16 random directions."""
    m, kappa = case
    op = igamf.Operator(m)
    f = op.load(igamf.LOAD_CONST, 1.0)
    solver = igamf.Solver(op, None, pc=igamf.PC_JACOBI, rtol=1e-13, maxit=100000)
    k_obs = torch.from_numpy(kappa_field(m, 2)).cuda()
    with torch.no_grad():
        u_obs = IA.solve_kappa(solver, f, k_obs)
        u_obs = u_obs + 1e-3 * torch.from_numpy(igamf.rng_normal(9, op.nfunc)).cuda() * \
            IA.interior_mask(op)
    k = torch.from_numpy(kappa).cuda().requires_grad_()
    u = IA.solve_kappa(solver, f, k)
    J = 0.5 * ((u - u_obs) ** 2).sum()
    J.backward()
    grad = k.grad.clone()
    h = 1e-6
    for s in range(16):
        d = torch.from_numpy(igamf.rng_normal(300 + s, k.numel())).cuda()
        with torch.no_grad():
            Jp = 0.5 * ((IA.solve_kappa(solver, f, k + h * d) - u_obs) ** 2).sum().item()
            Jm = 0.5 * ((IA.solve_kappa(solver, f, k - h * d) - u_obs) ** 2).sum().item()
        fd = (Jp - Jm) / (2 * h)
        an = (grad * d).sum().item()
        assert abs(fd - an) < 1e-5 * abs(an) + 1e-12, (s, fd, an)


@pytest.mark.parametrize("geom", ["annulus", "rational"])
def test_qp_coords_match_oracle(geom):
    if geom == "annulus":
        m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 9, 7)
    else:
        m = igamf.Mesh.build(igamf.GEOM_PERTURBED_SQUARE, 2, 8, perturb=0.1, w_lo=0.8,
                             w_hi=1.2, seed=4)
    op = igamf.Operator(m)
    xq = op.qp_coords().cpu().numpy()
    om = _om(m)
    for ey in range(om.ney):
        for ex in range(om.nex):
            ref = O.element_qp(om, ex, ey)[4]
            assert np.abs(xq[ey * om.nex + ex] - ref).max() < 1e-14


def test_set_kappa_refreshes_f16_geometry():
    """set_kappa rebuilds the FP16 copy of G_q (and its power-of-two scale):
    the FP16-geometry apply follows the new kappa, also one 1000x larger
    than the first (beyond the first scale's FP16 range)."""
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 3, 20)
    op = igamf.Operator(m, f16_geometry=True)
    x = torch.from_numpy(igamf.rng_normal(4, op.nfunc)).cuda()
    k1 = torch.from_numpy(kappa_field(m, 1)).cuda()
    s0 = op.info.f16_scale
    for kap in (k1, 1000.0 * k1):
        op.set_kappa(kap)
        y64 = op.apply(x).cpu().numpy()
        y16 = op.apply(x.float(), g16=True).double().cpu().numpy()
        assert rel(y16, y64) < 3e-3, rel(y16, y64)
    ref = igamf.Operator(m, kappa=(1000.0 * k1).cpu().numpy(), f16_geometry=True)
    y_ref = ref.apply(x.float(), g16=True)
    assert torch.equal(op.apply(x.float(), g16=True), y_ref)
    assert s0 > 0
