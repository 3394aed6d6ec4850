"""GPU tests of the multi right-hand-side apply Y[j] = A X[j] (FP32).

Every compiled block configuration is checked against k single-vector FP32
applies of the same operator and against the oracle's FP64 apply, on ragged
meshes whose element counts are not multiples of any block tile, for k below,
at and above the rhs-chunk widths, and with a leading dimension above nfunc.
"""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import igamf  # noqa: E402

TOL32 = 2e-5  # FP32 storage of G_q, x and y: relative to max|y|


def _oracle(m):
    a = m.arrays()
    return O.Mesh(a["p"], a["q"], a["nex"], a["ney"], a["CP"], a["W"], a["IEN"], a["ID"],
                  a["Dir"], a["C1"], a["C2"], a["h1"], a["h2"])


def _mesh(name):
    if name == "annulus_p3":
        return igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 37, ney=29)
    if name == "square_p2":
        return igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 2, 21)
    if name == "perturbed_p4":
        return igamf.Mesh.build(igamf.GEOM_PERTURBED_SQUARE, 4, 26, ney=19, perturb=0.15)
    if name == "square_p1":
        return igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 1, 13)
    if name == "annulus_p5":
        return igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 5, 18, ney=11)
    raise KeyError(name)


def _set(op, k, cfg):
    """Select a block configuration; False when it is not compiled for the degree."""
    try:
        op.set_block_config(k, cfg)
        return True
    except igamf.IgaError as e:
        assert e.status == 6, (cfg, e.status)
        return False


def _rel(a, b):
    return (a - b).abs().max().item() / b.abs().max().item()


@pytest.mark.parametrize("name", ["annulus_p3", "square_p2", "perturbed_p4", "square_p1",
                                  "annulus_p5"])
def test_block_matches_single_applies_and_oracle(name):
    m = _mesh(name)
    op = igamf.Operator(m)
    n = op.nfunc
    k = 37
    X = torch.from_numpy(igamf.rng_normal(11, k * n).reshape(k, n)).cuda()
    X32 = X.float()
    single = torch.stack([op.apply(X32[j], assembly=igamf.IEN_ATOMIC) for j in range(k)])
    om = _oracle(m)
    ref = torch.stack([torch.from_numpy(O.apply_mf(om, X[j].cpu().numpy())) for j in (0, k - 1)])
    for cfg in range(len(igamf.block_configs())):
        if not _set(op, k, cfg):
            continue
        Y = op.apply_block(X32)
        torch.cuda.synchronize()
        assert torch.isfinite(Y).all()
        assert _rel(Y, single) < 1e-5, (name, cfg)
        for jj, j in enumerate((0, k - 1)):
            assert _rel(Y[j].double().cpu(), ref[jj]) < TOL32, (name, cfg, j)


@pytest.mark.parametrize("k", [1, 2, 5, 8, 16, 31, 32, 33, 64])
def test_block_rhs_counts(k):
    m = _mesh("annulus_p3")
    op = igamf.Operator(m)
    n = op.nfunc
    X = torch.from_numpy(igamf.rng_normal(3 + k, k * n).reshape(k, n)).cuda().float()
    single = torch.stack([op.apply(X[j], assembly=igamf.IEN_ATOMIC) for j in range(k)])
    cfgs = igamf.block_configs()
    for cfg in range(len(cfgs)):
        op.set_block_config(k, cfg)
        Y = op.apply_block(X)
        assert _rel(Y, single) < 1e-5, (k, cfg, cfgs[cfg])
    op.set_block_config(k, -1)  # built-in default
    assert _rel(op.apply_block(X), single) < 1e-5


def test_block_leading_dimension_and_untouched_padding():
    m = _mesh("perturbed_p4")
    op = igamf.Operator(m)
    n, k, ld = op.nfunc, 9, op.nfunc + 13
    buf = torch.from_numpy(igamf.rng_normal(5, k * ld).reshape(k, ld)).cuda().float()
    X = buf[:, :n]
    out = torch.full((k, ld), 7.0, device="cuda")
    Y = out[:, :n]
    op.apply_block(X, Y)
    single = torch.stack([op.apply(X[j].contiguous(), assembly=igamf.IEN_ATOMIC)
                          for j in range(k)])
    assert _rel(Y, single) < 1e-5
    assert bool((out[:, n:] == 7.0).all())  # padding between columns untouched


def test_block_tuner_and_errors():
    m = _mesh("annulus_p3")
    op = igamf.Operator(m)
    n, k = op.nfunc, 32
    X = torch.from_numpy(igamf.rng_normal(9, k * n).reshape(k, n)).cuda().float()
    cfg, ms = op.tune_block(X, reps=2)
    assert 0 <= cfg < len(igamf.block_configs()) and ms > 0
    assert op.block_config(k) == cfg
    single = torch.stack([op.apply(X[j], assembly=igamf.IEN_ATOMIC) for j in range(k)])
    assert _rel(op.apply_block(X), single) < 1e-5
    with pytest.raises(igamf.IgaError):
        op.set_block_config(k, len(igamf.block_configs()))
    with pytest.raises(ValueError):
        op.apply_block(X.double())
    with pytest.raises(ValueError):
        op.apply_block(X[:, : n - 1])


def test_block_large_annulus_linearity():
    """1024^2 elements p = 3, k = 8: Y is linear in X column by column and
    matches the single applies."""
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 1024)
    op = igamf.Operator(m)
    n, k = op.nfunc, 8
    X = torch.from_numpy(igamf.rng_normal(1, k * n).reshape(k, n)).cuda().float()
    Y = op.apply_block(X)
    single = torch.stack([op.apply(X[j], assembly=igamf.IEN_ATOMIC) for j in range(k)])
    assert _rel(Y, single) < 1e-5
    Y2 = op.apply_block(2.0 * X)
    assert _rel(Y2, 2.0 * Y) < 1e-6
