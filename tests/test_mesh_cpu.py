"""CPU checks of the library's host side: its mesh generator against the
oracle's (and, through the oracle, the reference's) preprocessing, its random
numbers, and the status codes of invalid input (no GPU call is made)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
import igamf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "iga_mf.h")).read()
    names = set(re.findall(r"^\s*(?:int|void|const char \*)\s*(iga_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 24
    L = igamf.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_rng_matches_oracle():
    assert np.array_equal(igamf.rng_normal(7, 1001), O.rng_normal(7, 1001))
    assert np.array_equal(igamf.rng_uniform(3, 999, -0.1, 0.1),
                          O.rng_uniform(3, 999, -0.1, 0.1))


@pytest.mark.parametrize("p,nex,ney", [(1, 5, 7), (2, 32, 32), (3, 16, 9), (4, 6, 11),
                                       (5, 7, 7), (6, 8, 10), (8, 10, 10)])
def test_square_mesh_matches_oracle(p, nex, ney):
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, p, nex, ney).arrays()
    o = O.mesh_unit_square(p, nex, ney)
    assert np.array_equal(m["IEN"], o.IEN)
    assert np.array_equal(m["ID"], o.ID)
    assert np.array_equal(m["Dir"], o.Dir)
    assert np.abs(m["CP"] - o.CP).max() < 1e-15
    assert np.abs(m["C1"] - o.C1).max() < 1e-13
    assert np.abs(m["C2"] - o.C2).max() < 1e-13
    assert m["W"] is None


@pytest.mark.parametrize("p,nex,ney", [(2, 6, 9), (3, 16, 16), (3, 40, 7), (5, 9, 13)])
def test_annulus_mesh_matches_oracle(p, nex, ney):
    m = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, p, nex, ney, r0=1.0, r1=2.0).arrays()
    o = O.mesh_quarter_annulus(p, nex, ney, 1.0, 2.0)
    assert np.abs(m["CP"] - o.CP).max() < 1e-13
    assert np.abs(m["W"] - o.W).max() < 1e-13
    assert np.array_equal(m["IEN"], o.IEN)


def test_perturbed_square_draws():
    p, n = 3, 12
    m = igamf.Mesh.build(igamf.GEOM_PERTURBED_SQUARE, p, n, perturb=0.1, w_lo=0.8, w_hi=1.2,
                         seed=3).arrays()
    base = O.mesh_unit_square(p, n).CP.reshape(-1, 2)
    cp = m["CP"].reshape(-1, 2)
    nf = n + p
    ring = np.zeros((nf, nf), bool)
    ring[0, :] = ring[-1, :] = ring[:, 0] = ring[:, -1] = True
    ring = ring.reshape(-1)
    assert np.array_equal(cp[ring], base[ring])
    d = np.abs(cp[~ring] - base[~ring])
    assert d.max() <= 0.1 / n + 1e-15 and d.max() > 0.05 / n
    assert m["W"].min() >= 0.8 and m["W"].max() <= 1.2


def _op_status(**over):
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 3, 8)
    d = igamf.MeshDesc.from_buffer_copy(m.desc)
    keep = []
    for k, v in over.items():
        if isinstance(v, np.ndarray):
            keep.append(v)
            v = v.ctypes.data
        setattr(d, k, v)
    h = igamf.VP()
    rc = igamf.lib().iga_op_create(C.byref(d), None, C.byref(h))
    assert not h.value
    return rc


def test_create_rejects_invalid_input():
    assert _op_status(p=0) == 2          # INVALID_DEGREE
    assert _op_status(p=9) == 2
    assert _op_status(nqp_x=11) == 3     # INVALID_QUADRATURE
    assert _op_status(nqp_x=0) == 3
    assert _op_status(nelem_x=4) == 4    # TOO_FEW_ELEMENTS (p = 3 needs 5)
    assert _op_status(cp=None) == 1      # INVALID_ARGUMENT
    bad_ien = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 3, 8).arrays()["IEN"].copy()
    bad_ien[5] = 10 ** 6
    assert _op_status(ien=bad_ien) == 5  # BAD_MESH
    bad_id = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 3, 8).arrays()["ID"].copy()
    bad_id[12] = -1                      # interior function marked Dirichlet, not in Dir
    assert _op_status(id=bad_id) == 5
    assert igamf.lib().iga_status_string(5) == b"inconsistent mesh (IEN / ID / Dir / extraction)"


def test_mesh_build_rejects_invalid_input():
    h = igamf.VP()
    prm = igamf.MeshParams(0, 9, 20, 20, 1.0, 1.0, 2.0, 0.0, 1.0, 1.0, 0)
    assert igamf.lib().iga_mesh_build(C.byref(prm), C.byref(h)) == 2
    prm = igamf.MeshParams(0, 3, 4, 20, 1.0, 1.0, 2.0, 0.0, 1.0, 1.0, 0)
    assert igamf.lib().iga_mesh_build(C.byref(prm), C.byref(h)) == 4
    prm = igamf.MeshParams(1, 1, 20, 20, 1.0, 1.0, 2.0, 0.0, 1.0, 1.0, 0)
    assert igamf.lib().iga_mesh_build(C.byref(prm), C.byref(h)) == 2


def _same_mesh(a, b):
    for k in ("ID", "Dir", "IEN"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    for k in ("CP", "C1", "C2", "h1", "h2"):
        x, y = np.asarray(a[k], np.float64), np.asarray(b[k], np.float64)
        assert x.shape == y.shape, k
        # 16 significant digits on disk: within 1e-15 relative
        assert np.all(np.abs(x - y) <= 1e-15 * np.maximum(np.abs(x), 1e-300)), k


@pytest.mark.skipif(not O.ref_available(), reason="reference units not built (oracle/_ref)")
def test_partition_file_matches_reference_filemanager(tmp_path):
    """The reference's partition format (FileManager::WritePartition /
    ReadPartition, compiled from the reference): the library's writer emits
    the same bytes as the reference's writer for the same mesh, the
    reference's reader reads the library's file back, and the library's reader
    reads the reference's file back."""
    m = igamf.Mesh.build(igamf.GEOM_PERTURBED_SQUARE, 3, 10, 7, perturb=0.2, seed=5)
    a = m.arrays()
    assert a["W"] is None  # the format has no weights
    ours, refs = str(tmp_path / "ours_0.txt"), str(tmp_path / "ref_0.txt")
    m.write_partition(ours)
    O.ref_write_partition(refs, a)
    assert open(ours, "rb").read() == open(refs, "rb").read()
    r = O.ref_read_partition(ours, 4096)
    assert (r["nfunc"], r["nex"], r["ney"]) == (len(a["ID"]), 10, 7)
    _same_mesh(a, r)
    m2 = igamf.Mesh.read_partition(refs, 3, 3)
    _same_mesh(a, m2.arrays())


def test_partition_reader_rejects_bad_files(tmp_path):
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 2, 6)
    fn = str(tmp_path / "part_0.txt")
    m.write_partition(fn)
    m2 = igamf.Mesh.read_partition(fn, 2, 2)
    _same_mesh(m.arrays(), m2.arrays())
    with pytest.raises(igamf.IgaError) as e:  # IEN holds 9 functions per element, not 16
        igamf.Mesh.read_partition(fn, 3, 3)
    assert e.value.status == 5
    text = open(fn).read()
    short = str(tmp_path / "short_0.txt")  # the short form has no extraction operators
    open(short, "w").write(text.split("ElemSize1")[0] + "CP\n" + text.split("CP\n")[1].split("\n")[0]
                           + "\n")
    with pytest.raises(igamf.IgaError) as e:
        igamf.Mesh.read_partition(short, 2, 2)
    assert e.value.status == 5
    with pytest.raises(igamf.IgaError):
        igamf.Mesh.read_partition(str(tmp_path / "missing.txt"), 2, 2)
    with pytest.raises(igamf.IgaError) as e:  # weights cannot be written
        igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 2, 6).write_partition(fn)
    assert e.value.status == 6
    info = str(tmp_path / "info.txt")
    open(info, "w").write("p: 3\nq: 3\nLx: 1\nLy: 1.5\nnElemX: 64\nnElemY: 32\npart_num_1d: 1\n"
                          "dim: 2\nbase_name: part\n")
    d = igamf.read_preprocess_info(info)
    assert (d["p"], d["q"], d["nelem_x"], d["nelem_y"], d["Ly"], d["base_name"]) == \
        (3, 3, 64, 32, 1.5, "part")


def test_solver_opts_defaults():
    """iga_solver_opts_default (host code only) fills the values the header
    documents; reading them back through the ctypes mirror also checks that
    the mirror's field layout matches the C struct."""
    o = igamf.SolverOpts()
    o.pc_inner_abstol = 7.0  # the last field: overwritten only if the layouts agree
    igamf.lib().iga_solver_opts_default(C.byref(o))
    assert (o.kind, o.pc, o.kernel, o.assembly) == (
        igamf.SOLVE_PCG_F64, igamf.PC_JACOBI, igamf.SF, igamf.INVLM_GATHER)
    assert (o.rtol, o.abstol, o.divtol, o.maxit) == (1e-8, 0.0, 1e4, 100000)
    assert (o.inner_rtol, o.inner_maxit) == (0.0, 100000)  # 0: automatic
    assert (o.pc_inner_maxit, o.pc_inner_rtol, o.pc_cheb_degree) == (20, 1e-2, 8)
    assert (o.use_graph, o.inner_f16_geometry, o.fused) == (1, 0, 1)
    assert o.pc_inner_abstol == 0.0


def test_op_opts_defaults():
    """iga_op_opts_default (host code only): automatic variants and
    automatic streamed FP32 G_q loads; the last field doubles as the layout
    check of the ctypes mirror."""
    o = igamf.OpOpts()
    o.stream_g_f32 = 7
    igamf.lib().iga_op_opts_default(C.byref(o))
    assert (o.want_f64, o.want_f32, o.variant_f64, o.variant_f32) == (1, 1, -1, -1)
    assert (o.tile_x_f32, o.tile_y_f32, o.f16_geometry) == (0, 0, 0)
    assert o.stream_g_f32 == -1


def test_binding_rejects_bad_vectors_before_any_device_call():
    """The C ABI trusts bare device pointers: the binding checks dtype,
    device, contiguity and size first (no library call is reached here)."""
    import torch
    op = igamf.Operator.__new__(igamf.Operator)
    op._h, op.nfunc, op.nelem, op.nqp = None, 10, 2, 4
    op.device = torch.device("cuda", 0)
    cpu64 = torch.zeros(10, dtype=torch.float64)
    with pytest.raises(ValueError):
        op.apply(cpu64)
    with pytest.raises(ValueError):
        op.apply(torch.zeros(10, dtype=torch.int32))
    with pytest.raises(ValueError):
        op.apply(cpu64, g16=True)
    with pytest.raises(ValueError):
        op.set_kappa(torch.ones(8, dtype=torch.float64))
    with pytest.raises(ValueError):
        op.backward(cpu64, cpu64)
    with pytest.raises(ValueError):
        op.apply_block(torch.zeros((2, 10), dtype=torch.float32))
    s = igamf.Solver.__new__(igamf.Solver)
    s._h, s.op = None, op
    with pytest.raises(ValueError):
        s.solve(torch.zeros(10, dtype=torch.float32))
    with pytest.raises(ValueError):
        s.solve(cpu64)
    with pytest.raises(ValueError):
        s.solve_block(torch.zeros((2, 10), dtype=torch.float64))
    with pytest.raises(ValueError):
        igamf.check_vec(cpu64[::2], torch.float64, 5, op.device, "x")


# --------------------------------------------------------------------------- FEM-side files
def _ref_gpu_or_skip():
    if not O.ref_available():
        pytest.skip("reference CPU units not built (oracle/_ref)")
    return O.ref_gpu()


def _ref_fem_partition(R, fn, cap=1 << 20):
    counts = np.zeros(7, np.int32)
    ids = np.empty(cap, np.int32); dirs = np.empty(cap, np.int32)
    ien = np.empty(cap, np.int32); cp = np.empty(cap, np.float64)
    assert R.ref_fem_read_partition(fn.encode(), cap, counts, ids, dirs, ien, cp) == 0
    nf, nx, ny, nid, nd, nien, ncp = (int(c) for c in counts)
    return {"nlocalfunc": nf, "nlocalelemx": nx, "nlocalelemy": ny, "ID": ids[:nid],
            "Dir": dirs[:nd], "IEN": ien[:nien], "CP": cp[:ncp]}


@pytest.mark.parametrize("p,nex,ney,parts", [(3, 10, 12, 2), (4, 13, 9, 3)])
def test_fem_partition_files_match_reference(tmp_path, p, nex, ney, parts):
    """The reference's Partition::GeneratePartition writes part_fem_<r>.txt and
    new_to_old_mapping.txt; our readers return what its FileManager reads.
    The dry run's coordinate_nonzero_<r>.txt files (PETSc's MatGetRowIJ of
    each rank's rows, which cannot be built here) are written with the
    reference's FileManager from the oracle's pattern in partition numbering;
    the driver's merge (renumber, sort) then gives back the oracle's pattern
    bit for bit, and our writers write byte-identical files."""
    R = _ref_gpu_or_skip()
    d = str(tmp_path)
    nf = np.zeros(2, np.int32)
    assert R.ref_fem_write_partitions(d.encode(), p, p, nex, ney, parts, nf) == 0
    nfx, nfy = int(nf[0]), int(nf[1])
    nparts = parts * parts
    owned = []
    for r in range(nparts):
        fn = os.path.join(d, f"part_fem_{r}.txt")
        ours = igamf.read_fem_partition(fn)
        ref = _ref_fem_partition(R, fn)
        for k in ("nlocalfunc", "nlocalelemx", "nlocalelemy"):
            assert ours[k] == ref[k], k
        for k in ("ID", "Dir", "IEN"):
            assert np.array_equal(ours[k], ref[k]), k
        assert np.array_equal(ours["CP"], ref["CP"])
        owned.append(ours["nlocalfunc"])
    assert sum(owned) == nfx * nfy
    mfn = os.path.join(d, "new_to_old_mapping.txt")
    n2o = igamf.read_new_to_old(mfn)
    ref_map = np.empty(nfx * nfy, np.int32)
    assert R.ref_read_new_to_old(mfn.encode(), nfx * nfy, ref_map) == nfx * nfy
    assert np.array_equal(n2o, ref_map)
    assert np.array_equal(np.sort(n2o), np.arange(nfx * nfy))
    # the dry run's per-rank pattern in partition numbering: rank r owns a
    # contiguous block of new indices, each row's columns sorted
    rows, cols, _, _, _ = O.fem_assemble(nfx, nfy, O.fem_nodes(p, p, nex, ney))
    o2n = np.empty_like(n2o)
    o2n[n2o] = np.arange(len(n2o), dtype=np.int32)
    nr, nc = o2n[rows], o2n[cols]
    order = np.lexsort((nc, nr))
    nr, nc = nr[order], nc[order]
    start = np.concatenate([[0], np.cumsum(owned)])
    for r in range(nparts):
        sel = (nr >= start[r]) & (nr < start[r + 1])
        fn = os.path.join(d, f"coordinate_nonzero_{r}.txt")
        R.ref_write_nonzero(fn.encode(), int(sel.sum()), np.ascontiguousarray(nr[sel]),
                            np.ascontiguousarray(nc[sel]))
        a, b = igamf.read_nonzero_coordinate(fn)
        assert np.array_equal(a, nr[sel]) and np.array_equal(b, nc[sel])
        ours_fn = os.path.join(d, f"ours_{r}.txt")
        igamf.write_nonzero_coordinate(ours_fn, a, b)
        assert open(ours_fn, "rb").read() == open(fn, "rb").read()
    ours_map = os.path.join(d, "ours_map.txt")
    igamf.write_new_to_old(ours_map, n2o)
    assert open(ours_map, "rb").read() == open(mfn, "rb").read()
    pr, pc = igamf.fem_pattern_from_files(d, nparts)
    assert np.array_equal(pr, rows) and np.array_equal(pc, cols)


def test_fem_file_readers_reject_bad_input(tmp_path):
    bad = tmp_path / "bad.txt"
    bad.write_text("nnz: 3\nrows: 1 2 3 \ncols: 1 2 \n")
    with pytest.raises(igamf.IgaError) as e:
        igamf.read_nonzero_coordinate(str(bad))
    assert e.value.status == 5  # IGA_ERR_BAD_MESH
    with pytest.raises(igamf.IgaError):
        igamf.read_fem_partition(str(tmp_path / "missing.txt"))
    short = tmp_path / "part.txt"
    short.write_text("nlocalfunc: 4\nnlocalelemx: 1\nnlocalelemy: 1\nID\n0 1 2 3 \nDir\n\n"
                     "IEN\n0 1 2 9 \nCP\n0 0 1 0 0 1 1 1 \n")
    with pytest.raises(igamf.IgaError) as e:  # IEN entry outside the partition
        igamf.read_fem_partition(str(short))
    assert e.value.status == 5


def test_mesh_eval_reproduces_the_geometry():
    """iga_mesh_eval: the identity on the unit square (the Greville control
    points reproduce x(u) = u), exact circles on the quarter annulus
    (|x| = r0 + u (r1 - r0), the angle increasing from 0 to pi/2), the
    control points' own interpolation at the corners; outside: an error."""
    rng = np.random.default_rng(0)
    uv = rng.random((500, 2))
    uv[:4] = [[0, 0], [1, 0], [0, 1], [1, 1]]
    sq = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 3, 11, 7)
    assert np.abs(sq.eval(uv) - uv).max() < 1e-14
    an = igamf.Mesh.build(igamf.GEOM_QUARTER_ANNULUS, 3, 9, 13, r0=1.0, r1=2.0)
    xy = an.eval(uv)
    assert np.abs(np.hypot(xy[:, 0], xy[:, 1]) - (1.0 + uv[:, 0])).max() < 1e-13
    th = np.arctan2(xy[:, 1], xy[:, 0])
    order = np.argsort(uv[:, 1])
    assert np.all(np.diff(th[order]) >= -1e-15)
    assert np.abs(xy[:4] - [[1, 0], [2, 0], [0, 1], [0, 2]]).max() < 1e-14
    # the FEM node grid of the reference, mapped: nodes inside the annulus
    nodes = an.fem_nodes().reshape(-1, 2)
    r = np.hypot(nodes[:, 0], nodes[:, 1])
    assert r.min() >= 1 - 1e-14 and r.max() <= 2 + 1e-14
    with pytest.raises(igamf.IgaError):
        an.eval([[1.5, 0.5]])


def test_partition_readers_reject_non_finite_values(tmp_path):
    """NaN coordinates (what the reference's p = 2 FEM preprocessing writes)
    are IGA_ERR_NOT_FINITE, in both partition readers."""
    fem = tmp_path / "part_fem_0.txt"
    fem.write_text("nlocalfunc: 4\nnlocalelemx: 1\nnlocalelemy: 1\nID\n-1 -1 -1 -1 \nDir\n0 1 2 3 \n"
                   "IEN\n0 1 2 3 \nCP\nnan nan 1 0 0 1 1 1 \n")
    with pytest.raises(igamf.IgaError) as e:
        igamf.read_fem_partition(str(fem))
    assert e.value.status == igamf.IGA_ERR_NOT_FINITE
    m = igamf.Mesh.build(igamf.GEOM_UNIT_SQUARE, 2, 4)
    fn = tmp_path / "part_0.txt"
    m.write_partition(str(fn))
    txt = fn.read_text().split("\n")
    i = txt.index("CP")
    vals = txt[i + 1].split()
    vals[3] = "inf"
    txt[i + 1] = " ".join(vals) + " "
    fn.write_text("\n".join(txt))
    with pytest.raises(igamf.IgaError) as e:
        igamf.Mesh.read_partition(str(fn), 2, 2)
    assert e.value.status == igamf.IGA_ERR_NOT_FINITE
