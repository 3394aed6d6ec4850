"""GPU parity of the operator apply against the CPU oracle.

Bar: FP64 within 1e-12 relative (max-norm over the output, relative to the
output's max-norm); FP32 within 2e-5 relative; the gather assembly bitwise
reproducible.  Meshes are the product's own generator, checked against the
oracle's generator in test_mesh_cpu.py; the oracle applies on the same arrays.
"""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import igamf  # noqa: E402

TOL64 = 1e-12
TOL32 = 2e-5


def oracle_mesh(m):
    a = m.arrays()
    return O.Mesh(a["p"], a["q"], a["nex"], a["ney"], a["CP"], a["W"], a["IEN"], a["ID"],
                  a["Dir"], a["C1"], a["C2"], a["h1"], a["h2"])


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    ("square_p2_n32", dict(kind=igamf.GEOM_UNIT_SQUARE, p=2, nex=32)),
    ("square_p3_n16", dict(kind=igamf.GEOM_UNIT_SQUARE, p=3, nex=16)),
    ("square_p3_ragged", dict(kind=igamf.GEOM_UNIT_SQUARE, p=3, nex=37, ney=21)),
    ("annulus_p3_n24", dict(kind=igamf.GEOM_QUARTER_ANNULUS, p=3, nex=24)),
    ("annulus_p2_n9x40", dict(kind=igamf.GEOM_QUARTER_ANNULUS, p=2, nex=9, ney=40)),
    ("square_p4_n20", dict(kind=igamf.GEOM_UNIT_SQUARE, p=4, nex=20)),
    ("annulus_p5_n12", dict(kind=igamf.GEOM_QUARTER_ANNULUS, p=5, nex=12)),
    ("square_p1_n8", dict(kind=igamf.GEOM_UNIT_SQUARE, p=1, nex=8)),
    ("square_p6_n9", dict(kind=igamf.GEOM_UNIT_SQUARE, p=6, nex=9)),
    # per-function NURBS weights (not a tensor product): the general rational kernels
    ("rational_p2_n20", dict(kind=igamf.GEOM_PERTURBED_SQUARE, p=2, nex=20, perturb=0.1,
                             w_lo=0.8, w_hi=1.2, seed=3)),
    ("rational_p3_ragged", dict(kind=igamf.GEOM_PERTURBED_SQUARE, p=3, nex=37, ney=21,
                                perturb=0.1, w_lo=0.8, w_hi=1.2, seed=4)),
    ("rational_p5_n12", dict(kind=igamf.GEOM_PERTURBED_SQUARE, p=5, nex=12, perturb=0.1,
                             w_lo=0.8, w_hi=1.2, seed=5)),
]
RATIONAL = [c for c in CASES if c[0].startswith("rational")]


@pytest.fixture(scope="module")
def built():
    cache = {}

    def get(name, kw):
        if name not in cache:
            m = igamf.Mesh.build(**kw)
            cache[name] = (m, oracle_mesh(m), igamf.Operator(m))
        return cache[name]
    return get


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_apply_matches_oracle(built, name, kw):
    m, om, op = built(name, kw)
    x = igamf.rng_normal(0, op.nfunc)
    y_ref = O.apply_mf(om, x)
    xd = torch.from_numpy(x).cuda()
    x32 = xd.float()
    for kern in (igamf.SF, igamf.EBE):
        for asm in (igamf.IEN_ATOMIC, igamf.INVLM_GATHER):
            y64 = op.apply(xd, kernel=kern, assembly=asm).cpu().numpy()
            assert rel(y64, y_ref) < TOL64, (name, kern, asm, rel(y64, y_ref))
            y32 = op.apply(x32, kernel=kern, assembly=asm).double().cpu().numpy()
            assert rel(y32, y_ref) < TOL32, (name, kern, asm, rel(y32, y_ref))
    # Dirichlet rows are exactly zero
    y64 = op.apply(xd).cpu().numpy()
    assert np.all(y64[om.Dir] == 0.0)


DOT_CASES = [c for c in CASES if c[0] in ("square_p3_ragged", "annulus_p3_n24", "square_p1_n8",
                                           "square_p6_n9", "rational_p3_ragged",
                                           "rational_p5_n12")]


@pytest.mark.parametrize("name,kw", DOT_CASES, ids=[c[0] for c in DOT_CASES])
def test_apply_dot_matches_oracle(built, name, kw):
    """iga_op_apply_dot: the energy product x.Ax accumulated inside the apply
    kernel equals x . y_oracle (Dirichlet rows of y are zero), for the default
    and split-element kernels, both assemblies, EBE and SF; y equals the plain
    apply's (bitwise for the gather assembly)."""
    m, om, op0 = built(name, kw)
    x = igamf.rng_normal(5, op0.nfunc)  # Dirichlet entries nonzero too
    y_ref = O.apply_mf(om, x)
    xy_ref = float(x @ y_ref)
    xd = torch.from_numpy(x).cuda()
    ops = [op0]
    if kw["p"] >= 3:
        ops.append(igamf.Operator(m, variant_f64=3, variant_f32=3))
    for op in ops:
        for kern in (igamf.SF, igamf.EBE):
            for asm in (igamf.IEN_ATOMIC, igamf.INVLM_GATHER):
                for xin, tol in ((xd, 1e-12), (xd.float(), 2e-6)):
                    y, xy = op.apply_dot(xin, kernel=kern, assembly=asm)
                    y0 = op.apply(xin, kernel=kern, assembly=asm)
                    if asm == igamf.INVLM_GATHER:
                        assert torch.equal(y, y0)
                    e = abs(xy.item() - xy_ref) / abs(xy_ref)
                    assert e < tol, (name, op.info.variant_f32, kern, asm, xin.dtype, e)
                    # against the product of the returned vectors themselves
                    e2 = abs(xy.item() - float(xin.double() @ y.double())) / abs(xy_ref)
                    assert e2 < tol, (name, kern, asm, xin.dtype, e2)


def test_apply_dot_unsupported_variant():
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 40)
    op = igamf.Operator(m, variant_f64=1, variant_f32=1)
    x = torch.ones(op.nfunc, dtype=torch.float64, device="cuda")
    with pytest.raises(igamf.IgaError) as e:
        op.apply_dot(x)
    assert e.value.status == 6  # UNSUPPORTED
    y, xy = op.apply_dot(x, kernel=igamf.EBE)  # the EBE kernel has no variant 1
    assert abs(xy.item() - float(x @ y)) <= 1e-12 * float(x.abs() @ y.abs())


# FP16-stored G_q (unit roundoff 2^-11 per value): the output error is at
# most 2^-11 times the sum of the contributions' magnitudes, a few times |y|
TOL16 = 3e-3
G16_CASES = [c for c in CASES if not c[0].startswith("rational")]


@pytest.mark.parametrize("name,kw", G16_CASES, ids=[c[0] for c in G16_CASES])
def test_f16_geometry_apply(built, name, kw):
    """IGA_F32_G16: FP32 arithmetic on the FP16 copy of G_q, against the
    FP64 oracle (TOL16), both assemblies; its x.Ax product; and the FP16
    operator is still symmetric to FP32 rounding (its perturbation of A is a
    fixed symmetric one), checked on interior-supported vectors."""
    m, om, _ = built(name, kw)
    op = igamf.Operator(m, f16_geometry=True)
    assert op.info.f16_geometry == 1 and op.info.f16_scale > 0
    x = igamf.rng_normal(0, op.nfunc)
    y_ref = O.apply_mf(om, x)
    x32 = torch.from_numpy(x).cuda().float()
    for asm in (igamf.IEN_ATOMIC, igamf.INVLM_GATHER):
        y = op.apply(x32, assembly=asm, g16=True).double().cpu().numpy()
        e = rel(y, y_ref)
        assert e < TOL16, (name, asm, e)
        if name.startswith("annulus"):
            # curved geometry: G_q are not FP16 numbers, the rounded copy is used
            # (on the unit square with 2-point Gauss rules they can be exact)
            y32 = op.apply(x32, assembly=asm).double().cpu().numpy()
            assert rel(y, y32) > 1e-5
        y2, xy = op.apply_dot(x32, assembly=asm, g16=True)
        e2 = abs(xy.item() - float(x32.double() @ y2.double())) / abs(float(x @ y_ref))
        assert e2 < 2e-6, (name, asm, e2)
    mask = torch.from_numpy((om.ID >= 0).astype(np.float32)).cuda()
    u = torch.from_numpy(igamf.rng_normal(1, op.nfunc)).cuda().float() * mask
    v = torch.from_numpy(igamf.rng_normal(2, op.nfunc)).cuda().float() * mask
    au = op.apply(u, g16=True).double()
    av = op.apply(v, g16=True).double()
    a, b = float(v.double() @ au), float(u.double() @ av)
    scale = float(v.double().abs() @ au.abs())
    assert abs(a - b) <= 1e-5 * scale, (name, a, b)
    with pytest.raises(igamf.IgaError):
        op.apply(x32, kernel=igamf.EBE, g16=True)


def test_f16_geometry_unsupported():
    m = igamf.Mesh.build(igamf.GEOM_PERTURBED_SQUARE, 2, 20, perturb=0.1, w_lo=0.8, w_hi=1.2,
                         seed=3)
    with pytest.raises(igamf.IgaError) as e:
        igamf.Operator(m, f16_geometry=True)  # per-function weights
    assert e.value.status == 6
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 3, 40)
    with pytest.raises(igamf.IgaError):
        igamf.Operator(m, f16_geometry=True, tile_f32=(32, 16))
    op = igamf.Operator(m)
    with pytest.raises(igamf.IgaError):
        op.apply(torch.zeros(op.nfunc, device="cuda"), g16=True)  # no FP16 copy


@pytest.mark.parametrize("tile", [(32, 16), (16, 16), (64, 8), (32, 32)])
def test_pipelined_variant_and_tiles(tile):
    """Persistent pipelined SF kernel, the four-wave FP32 kernel and every
    compiled tile shape, on a ragged mesh whose element counts are not
    multiples of the tile."""
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 77, 45)
    om = oracle_mesh(m)
    x = igamf.rng_normal(2, om.nfunc)
    y_ref = O.apply_mf(om, x)
    xd = torch.from_numpy(x).cuda()
    for variant in (0, 1, 2, 3):
        if variant == 2 and tile != (16, 16):
            continue  # the four-wave FP32 kernel exists for 16x16 only
        if variant == 3 and tile != (16, 16):
            continue  # the split-element kernel has its own 16x8 tile
        v64 = 1 if variant == 2 else variant
        t = None if variant == 3 else tile
        op = igamf.Operator(m, tile_f64=t, tile_f32=t, variant_f64=v64, variant_f32=variant)
        assert op.info.variant_f32 == variant
        for asm in (igamf.IEN_ATOMIC, igamf.INVLM_GATHER):
            y64 = op.apply(xd, assembly=asm).cpu().numpy()
            assert rel(y64, y_ref) < TOL64, (tile, variant, asm, rel(y64, y_ref))
            y32 = op.apply(xd.float(), assembly=asm).double().cpu().numpy()
            assert rel(y32, y_ref) < TOL32, (tile, variant, asm, rel(y32, y_ref))


@pytest.mark.parametrize("name,kw", [c for c in CASES if c[1]["p"] >= 3],
                         ids=[c[0] for c in CASES if c[1]["p"] >= 3])
def test_split_element_variant(built, name, kw):
    """Two threads per element (variant 3, 16x8 tiles) on every mesh of
    degree >= 3, general rational ones included; its EBE companion too."""
    m, om, _ = built(name, kw)
    op = igamf.Operator(m, variant_f64=3, variant_f32=3)
    assert (op.info.tile_x_f32, op.info.tile_y_f32) == (16, 8)
    x = igamf.rng_normal(0, op.nfunc)
    y_ref = O.apply_mf(om, x)
    xd = torch.from_numpy(x).cuda()
    for kern in (igamf.SF, igamf.EBE):
        for asm in (igamf.IEN_ATOMIC, igamf.INVLM_GATHER):
            y64 = op.apply(xd, kernel=kern, assembly=asm).cpu().numpy()
            assert rel(y64, y_ref) < TOL64, (name, kern, asm, rel(y64, y_ref))
            y32 = op.apply(xd.float(), kernel=kern, assembly=asm).double().cpu().numpy()
            assert rel(y32, y_ref) < TOL32, (name, kern, asm, rel(y32, y_ref))
    with pytest.raises(igamf.IgaError):
        igamf.Operator(m, variant_f32=3, tile_f32=(16, 16))


MFMA_CASES = [c for c in CASES if c[1]["p"] == 3 and not c[0].startswith("rational")]


@pytest.mark.parametrize("name,kw", MFMA_CASES, ids=[c[0] for c in MFMA_CASES])
def test_matrix_core_variant(built, name, kw):
    """Variant 4 (degree 3: the element operator on the matrix cores) against
    the oracle on FP32 and FP16-stored G_q, both assemblies, its x.Ax
    product, bitwise reproducibility of the gather form; EBE applies of such
    an operator and its FP64 applies use their own kernels."""
    m, om, _ = built(name, kw)
    op = igamf.Operator(m, variant_f32=4, f16_geometry=True)
    assert op.info.variant_f32 == 4 and (op.info.tile_x_f32, op.info.tile_y_f32) == (16, 16)
    x = igamf.rng_normal(7, op.nfunc)
    y_ref = O.apply_mf(om, x)
    xd = torch.from_numpy(x).cuda()
    x32 = xd.float()
    for asm in (igamf.IEN_ATOMIC, igamf.INVLM_GATHER):
        y32 = op.apply(x32, assembly=asm).double().cpu().numpy()
        assert rel(y32, y_ref) < TOL32, (name, asm, rel(y32, y_ref))
        y16 = op.apply(x32, assembly=asm, g16=True).double().cpu().numpy()
        assert rel(y16, y_ref) < TOL16, (name, asm, rel(y16, y_ref))
        y, xy = op.apply_dot(x32, assembly=asm)
        e = abs(xy.item() - float(x @ y_ref)) / abs(float(x @ y_ref))
        assert e < 2e-6, (name, asm, e)
        ye = op.apply(x32, kernel=igamf.EBE, assembly=asm).double().cpu().numpy()
        assert rel(ye, y_ref) < TOL32, (name, asm, rel(ye, y_ref))
        y64 = op.apply(xd, assembly=asm).cpu().numpy()
        assert rel(y64, y_ref) < TOL64, (name, asm, rel(y64, y_ref))
    a = op.apply(x32)
    b = op.apply(x32)
    assert torch.equal(a, b)
    # Dirichlet rows exactly zero
    assert np.all(op.apply(x32).cpu().numpy()[om.Dir] == 0.0)


def test_matrix_core_variant_unsupported():
    """Variant 4 exists for degree 3 on 16x16 tiles, FP32, not general rational."""
    for kw in (dict(kind=igamf.GEOM_UNIT_SQUARE, p=2, nex=20),
               dict(kind=igamf.GEOM_UNIT_SQUARE, p=4, nex=20),
               RATIONAL[1][1]):
        with pytest.raises(igamf.IgaError) as e:
            igamf.Operator(igamf.Mesh.build(**kw), variant_f32=4)
        assert e.value.status == 6, kw
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 3, 20)
    with pytest.raises(igamf.IgaError):
        igamf.Operator(m, variant_f32=4, tile_f32=(32, 16))
    with pytest.raises(igamf.IgaError):
        igamf.Operator(m, variant_f64=4)


LDS_CASES = [c for c in CASES if c[1]["p"] <= 3 and not c[0].startswith("rational")]


@pytest.mark.parametrize("name,kw", LDS_CASES, ids=[c[0] for c in LDS_CASES])
def test_lds_staged_variant(built, name, kw):
    """Variant 5 (persistent workgroups, the next tile's G_q staged in LDS by
    direct global-to-LDS loads) against the oracle, both assemblies, its x.Ax
    product, the FP16-G_q form at degree 3, bitwise reproducibility; the
    element arithmetic and reduction order are those of variant 0, so the
    gather results agree with it to FP32 rounding of the same sums."""
    m, om, op0 = built(name, kw)
    op = igamf.Operator(m, variant_f32=5, f16_geometry=kw["p"] == 3)
    assert op.info.variant_f32 == 5 and (op.info.tile_x_f32, op.info.tile_y_f32) == (16, 16)
    x = igamf.rng_normal(11, op.nfunc)
    y_ref = O.apply_mf(om, x)
    xd = torch.from_numpy(x).cuda()
    x32 = xd.float()
    for asm in (igamf.IEN_ATOMIC, igamf.INVLM_GATHER):
        y32 = op.apply(x32, assembly=asm).double().cpu().numpy()
        assert rel(y32, y_ref) < TOL32, (name, asm, rel(y32, y_ref))
        y, xy = op.apply_dot(x32, assembly=asm)
        e = abs(xy.item() - float(x @ y_ref)) / abs(float(x @ y_ref))
        assert e < 2e-6, (name, asm, e)
        if kw["p"] == 3:
            y16 = op.apply(x32, assembly=asm, g16=True).double().cpu().numpy()
            assert rel(y16, y_ref) < TOL16, (name, asm, rel(y16, y_ref))
        ye = op.apply(x32, kernel=igamf.EBE, assembly=asm).double().cpu().numpy()
        assert rel(ye, y_ref) < TOL32, (name, asm, rel(ye, y_ref))
    a = op.apply(x32)
    b = op.apply(x32)
    assert torch.equal(a, b)
    y0 = op0.apply(x32)
    assert rel(a.double().cpu().numpy(), y0.double().cpu().numpy()) < 1e-6
    assert np.all(a.cpu().numpy()[om.Dir] == 0.0)


def test_lds_staged_variant_unsupported():
    """Variant 5 exists for FP32 at degrees 1-3 on 16x16 tiles, not general
    rational meshes."""
    for kw in (dict(kind=igamf.GEOM_UNIT_SQUARE, p=4, nex=20), RATIONAL[1][1]):
        with pytest.raises(igamf.IgaError) as e:
            igamf.Operator(igamf.Mesh.build(**kw), variant_f32=5)
        assert e.value.status == 6, kw
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 3, 20)
    with pytest.raises(igamf.IgaError):
        igamf.Operator(m, variant_f32=5, tile_f32=(32, 16))
    with pytest.raises(igamf.IgaError):
        igamf.Operator(m, variant_f64=5)


STREAM_CASES = [c for c in CASES if c[0] in ("annulus_p3_n24", "annulus_p5_n12",
                                             "rational_p5_n12", "square_p6_n9")]


@pytest.mark.parametrize("name,kw", STREAM_CASES, ids=[c[0] for c in STREAM_CASES])
def test_streamed_geometry_loads(built, name, kw):
    """The FP32 split-element SF kernels' streaming instantiation (G_q loaded
    non-temporally; the default from 1 GiB of FP32 G_q) gives bitwise the
    results of normal loads, both assemblies, x.Ax included; small meshes
    default to normal loads; EBE applies keep their one kernel."""
    m, om, op0 = built(name, kw)
    assert op0.info.stream_g_f32 == 0
    ops = {k: igamf.Operator(m, stream_g_f32=k) for k in (0, 1)}
    assert (ops[0].info.stream_g_f32, ops[1].info.stream_g_f32) == (0, 1)
    assert (ops[1].info.tile_x_f32, ops[1].info.tile_y_f32) == (16, 8)
    x32 = torch.from_numpy(igamf.rng_normal(12, op0.nfunc)).cuda().float()
    y_ref = O.apply_mf(om, x32.double().cpu().numpy())
    for asm in (igamf.IEN_ATOMIC, igamf.INVLM_GATHER):
        for kern in (igamf.SF, igamf.EBE):
            y1 = ops[1].apply(x32, kernel=kern, assembly=asm)
            assert rel(y1.double().cpu().numpy(), y_ref) < TOL32, (name, asm, kern)
            if asm == igamf.INVLM_GATHER:
                assert torch.equal(y1, ops[0].apply(x32, kernel=kern, assembly=asm))
        y1, d1 = ops[1].apply_dot(x32, assembly=asm)
        y0, d0 = ops[0].apply_dot(x32, assembly=asm)
        if asm == igamf.INVLM_GATHER:
            assert torch.equal(y1, y0) and d1.item() == d0.item()
    # no split kernel on another tile: the option changes nothing there
    op16 = igamf.Operator(m, stream_g_f32=1, tile_f32=(16, 16))
    assert op16.info.stream_g_f32 == 0
    with pytest.raises(igamf.IgaError):
        igamf.Operator(m, stream_g_f32=2)


def test_general_rational_mode_and_tiles():
    """Non-tensor weights select the general rational kernels (5 geometry
    components per point); both of their tiles agree with the oracle; the
    persistent variant is not offered for them."""
    m = igamf.Mesh.build(**RATIONAL[1][1])
    om = oracle_mesh(m)
    x = igamf.rng_normal(6, om.nfunc)
    y_ref = O.apply_mf(om, x)
    xd = torch.from_numpy(x).cuda()
    for tile in ((16, 16), (32, 16)):
        op = igamf.Operator(m, tile_f64=tile, tile_f32=tile)
        assert op.info.rational == 2 and op.info.variant_f32 == 0
        for kern in (igamf.SF, igamf.EBE):
            y64 = op.apply(xd, kernel=kern).cpu().numpy()
            assert rel(y64, y_ref) < TOL64, (tile, kern, rel(y64, y_ref))
            y32 = op.apply(xd.float(), kernel=kern).double().cpu().numpy()
            assert rel(y32, y_ref) < TOL32, (tile, kern, rel(y32, y_ref))
    for bad in (dict(variant_f32=1), dict(tile_f64=(64, 8))):
        with pytest.raises(igamf.IgaError) as e:
            igamf.Operator(m, **bad)
        assert e.value.status == 6
    # the multi-RHS path has no general rational kernel
    op = igamf.Operator(m)
    with pytest.raises(igamf.IgaError) as e:
        op.apply_block(torch.zeros(2, op.nfunc, device="cuda"))
    assert e.value.status == 6


def test_gather_bitwise_reproducible(built):
    m, om, op = built("annulus_p3_n24", CASES[3][1])
    x = torch.from_numpy(igamf.rng_normal(5, op.nfunc)).cuda()
    for dt in (torch.float64, torch.float32):
        xs = x.to(dt)
        y0 = op.apply(xs, assembly=igamf.INVLM_GATHER)
        for _ in range(3):
            y1 = op.apply(xs, assembly=igamf.INVLM_GATHER)
            assert torch.equal(y0, y1)


def test_golden_reference_vectors(built):
    """The reference's own apply (compiled from its sources) on its C1 mesh."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "mf_p2_n32.npz"))
    p, n = int(g["p"]), int(g["nelem"])
    h = np.full(n, 1.0 / n)
    m = igamf.Mesh.from_arrays(p, p, n, n, g["CP"], None, g["IEN"], g["ID"], g["Dir"],
                               g["extraction"], g["extraction"], h, h)
    op = igamf.Operator(m)
    x = torch.from_numpy(g["x"]).cuda()
    for kern in (igamf.SF, igamf.EBE):
        for asm in (igamf.IEN_ATOMIC, igamf.INVLM_GATHER):
            y = op.apply(x, kernel=kern, assembly=asm).cpu().numpy()
            assert rel(y, g["y_mf"]) < TOL64
            assert rel(y, g["y_sf"]) < TOL64
    F = op.load(igamf.LOAD_REF_POLY).cpu().numpy()
    assert rel(F, g["F_ref"]) < TOL64


@pytest.mark.parametrize("name,kw", CASES[:5] + RATIONAL,
                         ids=[c[0] for c in CASES[:5] + RATIONAL])
def test_load_matches_oracle(built, name, kw):
    m, om, op = built(name, kw)
    for kind, val in ((igamf.LOAD_REF_POLY, 0.0), (igamf.LOAD_CONST, 1.0),
                      (igamf.LOAD_SINSIN, 2 * np.pi ** 2)):
        F = op.load(kind, val).cpu().numpy()
        Fr = O.load(om, kind, val)
        assert rel(F, Fr) < TOL64, (name, kind, rel(F, Fr))
    # f given at the quadrature points (any values, here random)
    fq = igamf.rng_normal(31, op.nelem * op.nqp)
    F = op.load_qp(torch.from_numpy(fq).cuda()).cpu().numpy()
    Fr = O.load_qp(om, fq)
    assert rel(F, Fr) < TOL64, (name, "qp", rel(F, Fr))


def test_large_annulus_invariants():
    """C2-sized (1024^2, p=3) quarter annulus: invariants that hold at any size.

    (i) A*1 = 0 on interior rows; (ii) linear reproduction: x_A = a + b X_A +
    c Y_A gives (A x)_i = 0 on every interior row; (iii) symmetry
    u^T A v = v^T A u for interior-supported u, v; (iv) FP32 vs FP64 agreement;
    (v) scatter-add vs gather agreement."""
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 1024)
    op = igamf.Operator(m)
    a = m.arrays()
    CP = torch.from_numpy(a["CP"].reshape(-1, 2)).cuda()
    interior = torch.from_numpy(a["ID"] >= 0).cuda()
    ones = torch.ones(op.nfunc, dtype=torch.float64, device="cuda")
    y1 = op.apply(ones)
    lin = 0.3 + 1.7 * CP[:, 0] - 0.9 * CP[:, 1]
    y2 = op.apply(lin.contiguous())
    u = torch.from_numpy(igamf.rng_normal(1, op.nfunc)).cuda() * interior
    v = torch.from_numpy(igamf.rng_normal(2, op.nfunc)).cuda() * interior
    Au, Av = op.apply(u), op.apply(v)
    scale = Au.abs().max().item()
    assert y1.abs().max().item() < 1e-10 * scale
    # rational basis: linear reproduction holds up to quadrature error
    assert y2.abs().max().item() < 1e-7 * scale
    s1, s2 = torch.dot(v, Au).item(), torch.dot(u, Av).item()
    assert abs(s1 - s2) < 1e-11 * abs(s1)
    Au32 = op.apply(u.float()).double()
    assert (Au32 - Au).abs().max().item() < 1e-4 * scale
    Au_at = op.apply(u, assembly=igamf.IEN_ATOMIC)
    assert (Au_at - Au).abs().max().item() < 1e-13 * scale
    torch.cuda.synchronize()


def test_operator_from_partition_file(tmp_path):
    """A mesh read back from the reference's partition format gives the same
    operator (control points within 1e-15 relative after 16 digits on disk)."""
    m = igamf.Mesh.build(igamf.GEOM_PERTURBED_SQUARE, 3, 40, 33, perturb=0.2, seed=6)
    fn = str(tmp_path / "part_0.txt")
    m.write_partition(fn)
    m2 = igamf.Mesh.read_partition(fn, 3, 3)
    op, op2 = igamf.Operator(m), igamf.Operator(m2)
    x = torch.from_numpy(igamf.rng_normal(4, op.nfunc)).cuda()
    y, y2 = op.apply(x).cpu().numpy(), op2.apply(x).cpu().numpy()
    assert rel(y2, y) < 1e-13
