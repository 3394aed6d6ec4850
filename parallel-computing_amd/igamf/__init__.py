"""Python binding of libigamf (the MI355X matrix-free IGA Poisson operator).

A thin ctypes layer over the C ABI declared in include/iga_mf.h; PyTorch
tensors provide device memory and streams.  There is no CPU fallback: every
call goes to the HIP library, and loading fails loudly when it is missing.

Mirrors the reference's operator interface: ``Operator.mult(x, y)`` is
``MyMatMult`` (matrix_free_gpu/matrixfreegpu.cu:52-66) and
``Operator.create_vecs()`` is ``MyMatCreateVecs`` (:68-78); ``Solver`` is the
KSP/PCSHELL set-up of :257-306.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libigamf.so")

F64, F32, F32_G16 = 0, 1, 2  # F32_G16: FP32 arithmetic on FP16-stored G_q
SF, EBE = 0, 1
IEN_ATOMIC, INVLM_GATHER = 0, 1
GEOM_UNIT_SQUARE, GEOM_QUARTER_ANNULUS, GEOM_PERTURBED_SQUARE = 0, 1, 2
PC_NONE, PC_JACOBI, PC_FEM_CG, PC_FEM_CHEB, PC_FEM_ILU, PC_FEM_MG = 0, 1, 2, 3, 4, 5
SOLVE_PCG_F64, SOLVE_PCG_F32_REFINE = 0, 1
LOAD_REF_POLY, LOAD_CONST, LOAD_SINSIN = 0, 1, 2
IGA_ERR_INVALID_ARGUMENT, IGA_ERR_UNSUPPORTED = 1, 6
IGA_ERR_NOT_FINITE, IGA_ERR_DIVERGED, IGA_ERR_NOT_CONVERGED, IGA_ERR_SIZE_LIMIT = 9, 10, 11, 12


class IgaError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = _lib().iga_status_string(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else msg)


class MeshDesc(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("p", "q", "nelem_x", "nelem_y", "nqp_x", "nqp_y")] + [
        ("cp", C.c_void_p), ("weights", C.c_void_p), ("ien", C.c_void_p), ("id", C.c_void_p),
        ("dir", C.c_void_p), ("ndir", C.c_int), ("extraction_x", C.c_void_p),
        ("extraction_y", C.c_void_p), ("elem_size_x", C.c_void_p),
        ("elem_size_y", C.c_void_p), ("kappa", C.c_void_p)]


class MeshParams(C.Structure):
    _fields_ = [("kind", C.c_int), ("p", C.c_int), ("nelem_x", C.c_int), ("nelem_y", C.c_int),
                ("L", C.c_double), ("r0", C.c_double), ("r1", C.c_double),
                ("perturb", C.c_double), ("w_lo", C.c_double), ("w_hi", C.c_double),
                ("seed", C.c_uint64)]


class OpOpts(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("device", "want_f64", "want_f32", "tile_x_f64",
                                       "tile_y_f64", "tile_x_f32", "tile_y_f32",
                                       "variant_f64", "variant_f32", "f16_geometry",
                                       "stream_g_f32")]


class OpInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nfunc", "nfx", "nfy", "nelem", "nqp", "nloc", "ndir",
                                       "rational", "tile_x_f64", "tile_y_f64", "tile_x_f32",
                                       "tile_y_f32")] + [("device_bytes", C.c_int64),
                                                         ("setup_ms", C.c_double),
                                                         ("variant_f64", C.c_int),
                                                         ("variant_f32", C.c_int),
                                                         ("f16_geometry", C.c_int),
                                                         ("f16_scale", C.c_double),
                                                         ("stream_g_f32", C.c_int)]


class SolverOpts(C.Structure):
    _fields_ = [("kind", C.c_int), ("pc", C.c_int), ("kernel", C.c_int), ("assembly", C.c_int),
                ("rtol", C.c_double), ("abstol", C.c_double), ("divtol", C.c_double),
                ("maxit", C.c_int), ("inner_rtol", C.c_double), ("inner_maxit", C.c_int),
                ("pc_inner_maxit", C.c_int), ("pc_inner_rtol", C.c_double),
                ("pc_cheb_degree", C.c_int), ("use_graph", C.c_int),
                ("inner_f16_geometry", C.c_int), ("fused", C.c_int),
                ("pc_inner_abstol", C.c_double)]


class PreprocessInfo(C.Structure):
    _fields_ = [("p", C.c_int), ("q", C.c_int), ("Lx", C.c_double), ("Ly", C.c_double),
                ("nelem_x", C.c_int), ("nelem_y", C.c_int), ("part_num_1d", C.c_int),
                ("dim", C.c_int), ("base_name", C.c_char * 256)]


class FemPartitionDesc(C.Structure):
    _fields_ = [("nlocalfunc", C.c_int), ("nlocalelemx", C.c_int), ("nlocalelemy", C.c_int),
                ("nid", C.c_int), ("ndir", C.c_int), ("id", C.c_void_p), ("dir", C.c_void_p),
                ("ien", C.c_void_p), ("cp", C.c_void_p)]


def _sized(call, what):
    """Two-call size protocol of the file readers: query the count, then
    read into arrays of that size."""
    n = C.c_int64()
    rc = call(0, None, None, C.byref(n))
    if rc not in (0, 12):  # IGA_ERR_SIZE_LIMIT: the count is known
        raise IgaError(rc, what)
    return n.value


def read_fem_partition(filename):
    """A FEM partition file (part_fem_<r>.txt): dict of the header counts
    and the ID, Dir, IEN, CP arrays."""
    L = _lib()
    h = VP()
    check(L.iga_fem_read_partition(filename.encode(), C.byref(h)), "fem_read_partition")
    try:
        d = FemPartitionDesc()
        check(L.iga_fem_partition_view(h, C.byref(d)), "fem_partition_view")

        def arr(ptr, n, ct, dt):
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), (n,)).astype(dt).copy() \
                if n else np.zeros(0, dt)
        nel = d.nlocalelemx * d.nlocalelemy
        return {"nlocalfunc": d.nlocalfunc, "nlocalelemx": d.nlocalelemx,
                "nlocalelemy": d.nlocalelemy,
                "ID": arr(d.id, d.nid, C.c_int, np.int32),
                "Dir": arr(d.dir, d.ndir, C.c_int, np.int32),
                "IEN": arr(d.ien, nel * 4, C.c_int, np.int32),
                "CP": arr(d.cp, 2 * d.nid, C.c_double, np.float64)}
    finally:
        L.iga_fem_partition_destroy(h)


def read_nonzero_coordinate(filename):
    """(rows, cols) of a coordinate_nonzero_<r>.txt file."""
    L = _lib()
    fn = filename.encode()
    n = _sized(lambda c, a, b, p: L.iga_read_nonzero_coordinate(fn, c, a, b, p),
               "read_nonzero_coordinate")
    rows = np.empty(n, np.int32)
    cols = np.empty(n, np.int32)
    m = C.c_int64()
    check(L.iga_read_nonzero_coordinate(fn, n, rows.ctypes.data, cols.ctypes.data,
                                        C.byref(m)), "read_nonzero_coordinate")
    return rows, cols


def write_nonzero_coordinate(filename, rows, cols):
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    assert rows.shape == cols.shape
    check(_lib().iga_write_nonzero_coordinate(filename.encode(), len(rows), rows.ctypes.data,
                                              cols.ctypes.data), "write_nonzero_coordinate")


def read_new_to_old(filename):
    L = _lib()
    fn = filename.encode()
    n = _sized(lambda c, a, b, p: L.iga_read_new_to_old(fn, c, a, p), "read_new_to_old")
    out = np.empty(n, np.int32)
    m = C.c_int64()
    check(L.iga_read_new_to_old(fn, n, out.ctypes.data, C.byref(m)), "read_new_to_old")
    return out


def write_new_to_old(filename, m):
    m = np.ascontiguousarray(m, np.int32)
    check(_lib().iga_write_new_to_old(filename.encode(), len(m), m.ctypes.data),
          "write_new_to_old")


def fem_pattern_from_files(directory, nparts):
    """The GPU driver's merged, renumbered and sorted FEM COO pattern."""
    L = _lib()
    d = directory.encode()
    n = _sized(lambda c, a, b, p: L.iga_fem_pattern_from_files(d, nparts, c, a, b, p),
               "fem_pattern_from_files")
    rows = np.empty(n, np.int32)
    cols = np.empty(n, np.int32)
    m = C.c_int64()
    check(L.iga_fem_pattern_from_files(d, nparts, n, rows.ctypes.data, cols.ctypes.data,
                                       C.byref(m)), "fem_pattern_from_files")
    return rows, cols


def read_preprocess_info(filename):
    """The reference's preprocess info file as a dict (iga_read_preprocess_info)."""
    info = PreprocessInfo()
    check(_lib().iga_read_preprocess_info(os.fsencode(filename), C.byref(info)),
          "read_preprocess_info")
    return {"p": info.p, "q": info.q, "Lx": info.Lx, "Ly": info.Ly, "nelem_x": info.nelem_x,
            "nelem_y": info.nelem_y, "part_num_1d": info.part_num_1d, "dim": info.dim,
            "base_name": info.base_name.decode()}


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer_iterations", C.c_int),
                ("rel_residual", C.c_double), ("solve_ms", C.c_double), ("status", C.c_int)]


class CsrView(C.Structure):
    _fields_ = [("nrows", C.c_int), ("nnz", C.c_int64), ("rowptr", C.c_void_p),
                ("col", C.c_void_p), ("val", C.c_void_p), ("coo_row", C.c_void_p),
                ("elem2coo", C.c_void_p), ("dir2coo", C.c_void_p), ("ndir", C.c_int)]


_LIB = None
VP = C.c_void_p


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libigamf.so not built at {LIB_PATH}; run "
                              "python parallel-computing_amd/build.py")
        L = C.CDLL(LIB_PATH)
        L.iga_status_string.restype = C.c_char_p
        L.iga_status_string.argtypes = [C.c_int]
        L.iga_mesh_build.argtypes = [C.POINTER(MeshParams), C.POINTER(VP)]
        L.iga_mesh_describe.argtypes = [VP, C.POINTER(MeshDesc)]
        L.iga_mesh_destroy.argtypes = [VP]
        L.iga_mesh_eval.argtypes = [VP, C.c_int64, VP, VP]
        L.iga_mesh_read_partition.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(VP)]
        L.iga_mesh_write_partition.argtypes = [VP, C.c_char_p]
        L.iga_read_preprocess_info.argtypes = [C.c_char_p, C.POINTER(PreprocessInfo)]
        L.iga_fem_read_partition.argtypes = [C.c_char_p, C.POINTER(VP)]
        L.iga_fem_partition_view.argtypes = [VP, C.POINTER(FemPartitionDesc)]
        L.iga_fem_partition_destroy.argtypes = [VP]
        L.iga_read_nonzero_coordinate.argtypes = [C.c_char_p, C.c_int64, VP, VP,
                                                  C.POINTER(C.c_int64)]
        L.iga_write_nonzero_coordinate.argtypes = [C.c_char_p, C.c_int64, VP, VP]
        L.iga_read_new_to_old.argtypes = [C.c_char_p, C.c_int64, VP, C.POINTER(C.c_int64)]
        L.iga_write_new_to_old.argtypes = [C.c_char_p, C.c_int64, VP]
        L.iga_fem_pattern_from_files.argtypes = [C.c_char_p, C.c_int, C.c_int64, VP, VP,
                                                 C.POINTER(C.c_int64)]
        L.iga_rng_normal.argtypes = [C.c_uint64, C.c_int64, VP]
        L.iga_rng_uniform.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double, VP]
        L.iga_op_opts_default.argtypes = [C.POINTER(OpOpts)]
        L.iga_op_opts_default.restype = None
        L.iga_op_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(OpOpts), C.POINTER(VP)]
        L.iga_op_destroy.argtypes = [VP]
        L.iga_op_get_info.argtypes = [VP, C.POINTER(OpInfo)]
        L.iga_op_apply.argtypes = [VP, C.c_int, C.c_int, C.c_int, VP, VP, VP]
        L.iga_op_apply_dot.argtypes = [VP, C.c_int, C.c_int, C.c_int, VP, VP, VP, VP]
        L.iga_op_load.argtypes = [VP, C.c_int, C.c_double, VP, VP]
        L.iga_op_load_qp.argtypes = [VP, VP, VP, VP]
        L.iga_op_profile.argtypes = [VP, C.c_int, C.c_int]
        L.iga_op_profile_read.argtypes = [VP, C.POINTER(C.c_int)] + [C.POINTER(C.c_double)] * 3
        L.iga_op_time_kernel.argtypes = [VP, C.c_int, C.c_int, C.c_int, VP, VP, C.c_int, VP,
                                         C.POINTER(C.c_double)]
        L.iga_op_qp_coords.argtypes = [VP, VP, VP]
        L.iga_op_set_kappa.argtypes = [VP, VP, VP]
        L.iga_op_apply_backward.argtypes = [VP, VP, VP, VP, VP, VP]
        L.iga_op_apply_block.argtypes = [VP, C.c_int, C.c_int, VP, VP, VP]
        L.iga_block_config_count.argtypes = []
        L.iga_block_config_matrix_core.argtypes = [C.c_int, C.POINTER(C.c_int)]
        L.iga_block_config_desc.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.POINTER(C.c_int)]
        L.iga_op_set_block_config.argtypes = [VP, C.c_int, C.c_int]
        L.iga_op_get_block_config.argtypes = [VP, C.c_int, C.POINTER(C.c_int)]
        L.iga_op_tune_block.argtypes = [VP, C.c_int, C.c_int, VP, VP, VP, C.c_int,
                                        C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.iga_fem_reference_nodes.argtypes = [C.c_int] * 4 + [C.c_double] * 2 + [VP]
        L.iga_fem_assemble.argtypes = [C.c_int, C.c_int, VP, C.c_int, C.POINTER(VP)]
        L.iga_fem_view.argtypes = [VP, C.POINTER(CsrView)]
        L.iga_fem_destroy.argtypes = [VP]
        L.iga_solver_opts_default.argtypes = [C.POINTER(SolverOpts)]
        L.iga_solver_opts_default.restype = None
        L.iga_solver_create.argtypes = [VP, VP, C.POINTER(SolverOpts), C.POINTER(VP)]
        L.iga_solver_solve.argtypes = [VP, VP, VP, VP, C.POINTER(SolveStats)]
        L.iga_solver_solve_block.argtypes = [VP, C.c_int, C.c_int, VP, VP, VP,
                                             C.POINTER(SolveStats)]
        L.iga_solver_destroy.argtypes = [VP]
        L.iga_solver_apply_pc.argtypes = [VP, C.c_int, VP, VP, VP]
        L.iga_dot_f64.argtypes = [VP, VP, C.c_int64, C.POINTER(C.c_double), VP]
        L.iga_dot_f64_noalloc.argtypes = [VP, VP, C.c_int64, VP, VP]
        _LIB = L
    return _LIB


def lib():
    return _lib()


def check(status, what=""):
    if status != 0:
        raise IgaError(status, what)


def rng_normal(seed, n):
    out = np.empty(n, np.float64)
    check(_lib().iga_rng_normal(seed, n, out.ctypes.data), "rng_normal")
    return out


def rng_uniform(seed, n, lo, hi):
    out = np.empty(n, np.float64)
    check(_lib().iga_rng_uniform(seed, n, lo, hi, out.ctypes.data), "rng_uniform")
    return out


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


def _ptr(t):
    return None if t is None else t.data_ptr()


def check_vec(t, dtype, n, device, name):
    """Raise ValueError unless t is a contiguous tensor of n values of dtype
    on the CUDA device `device` (the C ABI takes bare device pointers and
    trusts their size and type)."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_cuda or t.device != device:
        raise ValueError(f"{name} must live on {device}, got {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() != n:
        raise ValueError(f"{name} must have {n} values, got {t.numel()}")
    return t


# --------------------------------------------------------------------------- mesh
class Mesh:
    """Host mesh of one patch, built by the library (iga_mesh_build) or wrapped
    from caller arrays (``Mesh.from_arrays``)."""

    def __init__(self, handle=None, arrays=None):
        self._h = handle
        self._arrays = arrays
        self.desc = MeshDesc()
        if handle is not None:
            check(_lib().iga_mesh_describe(handle, C.byref(self.desc)), "mesh_describe")
        else:
            a = arrays
            d = self.desc
            d.p, d.q, d.nelem_x, d.nelem_y = a["p"], a["q"], a["nex"], a["ney"]
            d.nqp_x, d.nqp_y = a.get("nq1", a["p"] + 1), a.get("nq2", a["q"] + 1)
            for fld, key in (("cp", "CP"), ("weights", "W"), ("ien", "IEN"), ("id", "ID"),
                             ("dir", "Dir"), ("extraction_x", "C1"), ("extraction_y", "C2"),
                             ("elem_size_x", "h1"), ("elem_size_y", "h2"), ("kappa", "kappa")):
                v = a.get(key)
                setattr(d, fld, None if v is None else v.ctypes.data)
            d.ndir = len(a["Dir"])

    @classmethod
    def build(cls, kind, p, nex, ney=None, L=1.0, r0=1.0, r1=2.0, perturb=0.0, w_lo=1.0,
              w_hi=1.0, seed=0):
        prm = MeshParams(kind, p, nex, nex if ney is None else ney, L, r0, r1, perturb, w_lo,
                         w_hi, seed)
        h = VP()
        check(_lib().iga_mesh_build(C.byref(prm), C.byref(h)), "mesh_build")
        return cls(handle=h)

    @classmethod
    def read_partition(cls, filename, p, q):
        """A partition file in the reference's FileManager format
        (iga_mesh_read_partition)."""
        h = VP()
        check(_lib().iga_mesh_read_partition(os.fsencode(filename), p, q, C.byref(h)),
              "mesh_read_partition")
        return cls(handle=h)

    def write_partition(self, filename):
        """iga_mesh_write_partition (meshes built or read by the library)."""
        if self._h is None:
            raise ValueError("write_partition needs a library-owned mesh")
        check(_lib().iga_mesh_write_partition(self._h, os.fsencode(filename)),
              "mesh_write_partition")

    def eval(self, uv):
        """Geometry map at parametric points uv [n, 2] -> [n, 2]
        (iga_mesh_eval; library-owned meshes)."""
        if self._h is None:
            raise ValueError("eval needs a library-owned mesh")
        uv = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
        xy = np.empty_like(uv)
        check(_lib().iga_mesh_eval(self._h, len(uv), uv.ctypes.data, xy.ctypes.data),
              "mesh_eval")
        return xy

    def fem_nodes(self):
        """The reference's FEM-preconditioner node grid (iga_fem_reference_nodes
        on the unit parameter square) mapped through the geometry: on the
        unit square it is the reference's grid itself."""
        d = self.desc
        uv = fem_reference_nodes(d.p, d.q, d.nelem_x, d.nelem_y)
        return self.eval(uv).reshape(-1)

    @classmethod
    def from_arrays(cls, p, q, nex, ney, CP, W, IEN, ID, Dir, C1, C2, h1, h2, kappa=None):
        f64 = lambda v: None if v is None else np.ascontiguousarray(v, np.float64).reshape(-1)
        i32 = lambda v: None if v is None else np.ascontiguousarray(v, np.int32).reshape(-1)
        arrays = dict(p=p, q=q, nex=nex, ney=ney, CP=f64(CP), W=f64(W), IEN=i32(IEN),
                      ID=i32(ID), Dir=i32(Dir), C1=f64(C1), C2=f64(C2), h1=f64(h1), h2=f64(h2),
                      kappa=f64(kappa))
        return cls(arrays=arrays)

    def arrays(self):
        """numpy copies of the mesh arrays (reference layout)."""
        d = self.desc
        nfx, nfy = d.nelem_x + d.p, d.nelem_y + d.q
        nf, E, nloc = nfx * nfy, d.nelem_x * d.nelem_y, (d.p + 1) * (d.q + 1)

        def grab(ptr, n, dt):
            if not ptr:
                return None
            ct = C.c_double if dt == np.float64 else C.c_int
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()
        return dict(p=d.p, q=d.q, nex=d.nelem_x, ney=d.nelem_y,
                    CP=grab(d.cp, 2 * nf, np.float64), W=grab(d.weights, nf, np.float64),
                    IEN=grab(d.ien, E * nloc, np.int32), ID=grab(d.id, nf, np.int32),
                    Dir=grab(d.dir, d.ndir, np.int32),
                    C1=grab(d.extraction_x, d.nelem_x * (d.p + 1) ** 2, np.float64),
                    C2=grab(d.extraction_y, d.nelem_y * (d.q + 1) ** 2, np.float64),
                    h1=grab(d.elem_size_x, d.nelem_x, np.float64),
                    h2=grab(d.elem_size_y, d.nelem_y, np.float64))

    @property
    def nfunc(self):
        d = self.desc
        return (d.nelem_x + d.p) * (d.nelem_y + d.q)

    def __del__(self):
        if self._h is not None and _LIB is not None:
            _LIB.iga_mesh_destroy(self._h)
            self._h = None


# --------------------------------------------------------------------------- operator
class Operator:
    """The matrix-free operator y = A x on one GPU (iga_op)."""

    def __init__(self, mesh, device=0, want_f64=True, want_f32=True, tile_f64=None,
                 tile_f32=None, kappa=None, variant_f64=-1, variant_f32=-1, f16_geometry=False,
                 stream_g_f32=-1):
        import torch
        self.mesh = mesh
        o = OpOpts()
        _lib().iga_op_opts_default(C.byref(o))
        o.device, o.want_f64, o.want_f32 = device, int(want_f64), int(want_f32)
        o.f16_geometry = int(f16_geometry)
        if tile_f64:
            o.tile_x_f64, o.tile_y_f64 = tile_f64
        if tile_f32:
            o.tile_x_f32, o.tile_y_f32 = tile_f32
        o.variant_f64, o.variant_f32 = variant_f64, variant_f32
        o.stream_g_f32 = stream_g_f32
        desc = mesh.desc
        keep = None
        if kappa is not None:
            keep = np.ascontiguousarray(kappa, np.float64)
            desc = MeshDesc.from_buffer_copy(mesh.desc)
            desc.kappa = keep.ctypes.data
        self.device = torch.device("cuda", device)
        h = VP()
        check(_lib().iga_op_create(C.byref(desc), C.byref(o), C.byref(h)), "op_create")
        self._h = h
        info = OpInfo()
        check(_lib().iga_op_get_info(h, C.byref(info)), "op_get_info")
        self.info = info
        self.nfunc = info.nfunc
        self.nelem = info.nelem
        self.nqp = info.nqp
        self.p = mesh.desc.p

    def create_vecs(self, dtype=None):
        import torch
        dtype = dtype or torch.float64
        return (torch.zeros(self.nfunc, dtype=dtype, device=self.device),
                torch.zeros(self.nfunc, dtype=dtype, device=self.device))

    @staticmethod
    def _dtype(x, g16):
        import torch
        dt = {torch.float64: F64, torch.float32: F32}.get(getattr(x, "dtype", None))
        if dt is None:
            raise ValueError("vectors must be float64 or float32 tensors")
        if g16:
            if dt != F32:
                raise ValueError("FP16 geometry applies take float32 vectors")
            dt = F32_G16
        return dt

    def _xy(self, x, y, g16):
        """Checked (dtype code, x, y) of an apply; y allocated when None."""
        import torch
        dt = self._dtype(x, g16)
        check_vec(x, x.dtype, self.nfunc, self.device, "x")
        if y is None:
            y = torch.empty_like(x)
        check_vec(y, x.dtype, self.nfunc, self.device, "y")
        if y.data_ptr() == x.data_ptr():
            raise ValueError("x and y must not alias")
        return dt, x, y

    def apply(self, x, y=None, kernel=SF, assembly=INVLM_GATHER, stream=None, g16=False):
        """y = A x (iga_op_apply); g16: FP32 arithmetic on the FP16 G_q."""
        dt, x, y = self._xy(x, y, g16)
        check(_lib().iga_op_apply(self._h, dt, kernel, assembly, x.data_ptr(), y.data_ptr(),
                                  _stream_ptr(stream)), "op_apply")
        return y

    mult = apply  # MATOP_MULT

    def apply_dot(self, x, y=None, kernel=SF, assembly=INVLM_GATHER, stream=None, g16=False):
        """y = A x and x.y (float64 device scalar) from one kernel
        (iga_op_apply_dot)."""
        import torch
        dt, x, y = self._xy(x, y, g16)
        xy = torch.empty((), dtype=torch.float64, device=x.device)
        check(_lib().iga_op_apply_dot(self._h, dt, kernel, assembly, x.data_ptr(), y.data_ptr(),
                                      xy.data_ptr(), _stream_ptr(stream)), "op_apply_dot")
        return y, xy

    def profile(self, enable, max_calls=1024):
        check(_lib().iga_op_profile(self._h, int(enable), max_calls), "op_profile")

    def profile_read(self):
        n = C.c_int()
        tot, mn, mx = C.c_double(), C.c_double(), C.c_double()
        check(_lib().iga_op_profile_read(self._h, C.byref(n), C.byref(tot), C.byref(mn),
                                         C.byref(mx)), "op_profile_read")
        return n.value, tot.value, mn.value, mx.value

    def load(self, kind=LOAD_REF_POLY, value=0.0, stream=None):
        import torch
        F = torch.empty(self.nfunc, dtype=torch.float64, device=self.device)
        check(_lib().iga_op_load(self._h, kind, value, F.data_ptr(), _stream_ptr(stream)),
              "op_load")
        return F

    def load_qp(self, fq, stream=None):
        """Load vector of f given at the quadrature points (float64 CUDA,
        nelem * nqp values, the order of qp_coords)."""
        import torch
        fq = check_vec(fq.reshape(-1), torch.float64, self.nelem * self.nqp, self.device, "fq")
        F = torch.empty(self.nfunc, dtype=torch.float64, device=self.device)
        check(_lib().iga_op_load_qp(self._h, fq.data_ptr(), F.data_ptr(), _stream_ptr(stream)),
              "op_load_qp")
        return F

    def time_kernel(self, x, y, nlaunch=50, kernel=SF, assembly=INVLM_GATHER, stream=None,
                    g16=False):
        """Average ms of the element kernel alone over nlaunch back-to-back
        launches (y receives partial results)."""
        dt, x, y = self._xy(x, y, g16)
        ms = C.c_double()
        check(_lib().iga_op_time_kernel(self._h, dt, kernel, assembly, x.data_ptr(), y.data_ptr(),
                                        nlaunch, _stream_ptr(stream), C.byref(ms)), "time_kernel")
        return ms.value

    def qp_coords(self, stream=None):
        """Physical quadrature-point coordinates, float64 CUDA [E, nqp, 2]."""
        import torch
        xq = torch.empty((self.nelem, self.nqp, 2), dtype=torch.float64, device=self.device)
        check(_lib().iga_op_qp_coords(self._h, xq.data_ptr(), _stream_ptr(stream)), "qp_coords")
        return xq

    def set_kappa(self, kappa, stream=None):
        """kappa: float64 CUDA tensor of nelem * nqp values on the op's device
        (element-major, quadrature points in the reference's order)."""
        import torch
        k = check_vec(kappa.reshape(-1) if hasattr(kappa, "reshape") else kappa,
                      torch.float64, self.nelem * self.nqp, self.device, "kappa")
        check(_lib().iga_op_set_kappa(self._h, k.data_ptr(), _stream_ptr(stream)), "set_kappa")

    def backward(self, x, g, want_x=True, want_kappa=True, stream=None):
        import torch
        check_vec(g, torch.float64, self.nfunc, self.device, "g")
        if want_kappa or x is not None:
            check_vec(x, torch.float64, self.nfunc, self.device, "x")
        gx = torch.empty_like(g) if want_x else None
        gk = (torch.empty(self.nelem * self.nqp, dtype=torch.float64, device=self.device)
              if want_kappa else None)
        check(_lib().iga_op_apply_backward(self._h, _ptr(x), g.data_ptr(), _ptr(gx), _ptr(gk),
                                           _stream_ptr(stream)), "apply_backward")
        return gx, gk

    def _block_args(self, X, Y):
        import torch
        if (X.dtype != torch.float32 or not X.is_cuda or X.dim() != 2
                or X.device != self.device):
            raise ValueError("X must be a 2-D float32 tensor [k, nfunc] on the op's device")
        k, n = X.shape
        if n != self.nfunc or X.stride(1) != 1 or (k > 1 and X.stride(0) < n):
            raise ValueError("X rows must be contiguous vectors of length nfunc")
        if Y is None:
            Y = torch.empty_like(X)
        if Y.shape != X.shape or Y.stride() != X.stride() or Y.dtype != X.dtype:
            raise ValueError("Y must match X in shape, strides and dtype")
        return k, X.stride(0), Y

    def apply_block(self, X, Y=None, stream=None):
        """Y[j] = A X[j] for the k rows of X (FP32); a row stride above
        nfunc is the leading dimension."""
        k, ld, Y = self._block_args(X, Y)
        check(_lib().iga_op_apply_block(self._h, k, ld, X.data_ptr(), Y.data_ptr(),
                                        _stream_ptr(stream)), "apply_block")
        return Y

    def block_config(self, k):
        c = C.c_int()
        check(_lib().iga_op_get_block_config(self._h, k, C.byref(c)), "get_block_config")
        return c.value

    def set_block_config(self, k, cfg):
        check(_lib().iga_op_set_block_config(self._h, k, cfg), "set_block_config")

    def tune_block(self, X, Y=None, reps=5, stream=None):
        """Time every block configuration for k = X.shape[0]; the fastest is
        used by later apply_block calls with this k.  Returns (cfg, ms)."""
        k, ld, Y = self._block_args(X, Y)
        cfg, ms = C.c_int(), C.c_double()
        check(_lib().iga_op_tune_block(self._h, k, ld, X.data_ptr(), Y.data_ptr(),
                                       _stream_ptr(stream), reps, C.byref(cfg), C.byref(ms)),
              "tune_block")
        return cfg.value, ms.value

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.iga_op_destroy(self._h)
            self._h = None


def block_configs():
    """[(tile_x, tile_y, rhs_chunk)] of the compiled block-apply configurations."""
    L = _lib()
    out = []
    for i in range(L.iga_block_config_count()):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(L.iga_block_config_desc(i, C.byref(a), C.byref(b), C.byref(c)), "block_config_desc")
        out.append((a.value, b.value, c.value))
    return out


def block_config_matrix_core(cfg):
    """Whether block configuration cfg runs on the matrix cores."""
    v = C.c_int()
    check(_lib().iga_block_config_matrix_core(cfg, C.byref(v)), "block_config_matrix_core")
    return bool(v.value)


# --------------------------------------------------------------------------- FEM-PC
def fem_reference_nodes(p, q, nex, ney, Lx=1.0, Ly=1.0):
    nodes = np.empty(2 * (nex + p) * (ney + q))
    check(_lib().iga_fem_reference_nodes(p, q, nex, ney, Lx, Ly, nodes.ctypes.data),
          "fem_reference_nodes")
    return nodes


class FemPC:
    """Q1 FEM preconditioner matrix assembled on the device (iga_fem)."""

    def __init__(self, nfx, nfy, nodes, device=0):
        nodes = np.ascontiguousarray(nodes, np.float64)
        h = VP()
        check(_lib().iga_fem_assemble(nfx, nfy, nodes.ctypes.data, device, C.byref(h)),
              "fem_assemble")
        self._h = h
        v = CsrView()
        check(_lib().iga_fem_view(h, C.byref(v)), "fem_view")
        self.view = v
        self.nfx, self.nfy = nfx, nfy

    def to_torch(self):
        """Copies of the CSR (rowptr, col, val, coo_row, elem2coo, dir2coo)."""
        import torch
        v = self.view
        ne = (self.nfx - 1) * (self.nfy - 1)

        def wrap(ptr, n, dt):
            out = torch.empty(n, dtype=dt, device="cuda")
            torch.cuda.synchronize()
            _copy_dev(out, ptr, n * out.element_size())
            return out
        return dict(rowptr=wrap(v.rowptr, v.nrows + 1, torch.int32),
                    col=wrap(v.col, v.nnz, torch.int32),
                    val=wrap(v.val, v.nnz, torch.float64),
                    coo_row=wrap(v.coo_row, v.nnz, torch.int32),
                    elem2coo=wrap(v.elem2coo, ne * 16, torch.int32),
                    dir2coo=wrap(v.dir2coo, v.ndir, torch.int32))

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.iga_fem_destroy(self._h)
            self._h = None


_hip = None


def _copy_dev(dst_tensor, src_ptr, nbytes):
    """Device-to-device copy of nbytes from a raw device pointer."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [VP, VP, C.c_size_t, C.c_int]
    rc = _hip.hipMemcpy(dst_tensor.data_ptr(), src_ptr, nbytes, 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")


# --------------------------------------------------------------------------- solver
# statuses that describe the solve's outcome rather than a failed call
_SOLVE_OUTCOMES = (IGA_ERR_NOT_CONVERGED, IGA_ERR_DIVERGED, IGA_ERR_NOT_FINITE)


class Solver:
    def __init__(self, op, fem=None, **kw):
        o = SolverOpts()
        _lib().iga_solver_opts_default(C.byref(o))
        names = {f[0] for f in SolverOpts._fields_}
        for k, v in kw.items():
            if k not in names:
                raise TypeError(f"unknown solver option {k!r}")
            setattr(o, k, v)
        h = VP()
        check(_lib().iga_solver_create(op._h, None if fem is None else fem._h, C.byref(o),
                                       C.byref(h)), "solver_create")
        self._h = h
        self.op = op
        self.opts = o

    def solve(self, f, u=None, stream=None, check_status=True):
        """Solve A u = f; returns (u, stats).  A solve that does not converge
        (or an f holding a NaN or an infinity) raises IgaError unless
        check_status is False (stats.status then carries the code)."""
        import torch
        check_vec(f, torch.float64, self.op.nfunc, self.op.device, "f")
        if u is None:
            u = torch.zeros_like(f)
        check_vec(u, torch.float64, self.op.nfunc, self.op.device, "u")
        if u.data_ptr() == f.data_ptr():
            raise ValueError("f and u must not alias")
        st = SolveStats()
        rc = _lib().iga_solver_solve(self._h, f.data_ptr(), u.data_ptr(), _stream_ptr(stream),
                                     C.byref(st))
        if rc != 0 and (check_status or rc not in _SOLVE_OUTCOMES):
            raise IgaError(rc, "solver_solve")
        return u, st

    def apply_pc(self, r, z=None, stream=None):
        """z = M^-1 r, one application of the FEM preconditioner
        (iga_solver_apply_pc): float64 for an FP64 solver, float32 for an
        FP32-refinement one."""
        import torch
        dt = {torch.float64: F64, torch.float32: F32}.get(r.dtype)
        if dt is None:
            raise ValueError("r must be float64 or float32")
        check_vec(r, r.dtype, self.op.nfunc, self.op.device, "r")
        if z is None:
            z = torch.empty_like(r)
        check_vec(z, r.dtype, self.op.nfunc, self.op.device, "z")
        check(_lib().iga_solver_apply_pc(self._h, dt, r.data_ptr(), z.data_ptr(),
                                         _stream_ptr(stream)), "solver_apply_pc")
        return z

    def solve_block(self, F, U=None, stream=None, check_status=True):
        """Solve A U[j] = F[j] for the k rows of F (FP64 [k, nfunc], rows
        contiguous; a row stride above nfunc is the leading dimension).
        Returns (U, [stats per row])."""
        import torch
        if (F.dtype != torch.float64 or F.dim() != 2 or F.stride(1) != 1 or not F.is_cuda
                or F.device != self.op.device):
            raise ValueError("F must be a 2-D float64 tensor on the op's device with "
                             "contiguous rows")
        k, n = F.shape
        if n != self.op.nfunc or k < 1 or (k > 1 and F.stride(0) < n):
            raise ValueError("F rows must be vectors of length nfunc")
        if U is None:  # same leading dimension as F
            U = torch.zeros((k, F.stride(0)), dtype=F.dtype, device=F.device)[:, :n]
        if U.shape != F.shape or U.stride() != F.stride():
            raise ValueError("U must match F in shape and strides")
        st = (SolveStats * k)()
        rc = _lib().iga_solver_solve_block(self._h, k, F.stride(0), F.data_ptr(), U.data_ptr(),
                                           _stream_ptr(stream), st)
        if rc != 0 and (check_status or rc not in _SOLVE_OUTCOMES):
            raise IgaError(rc, "solver_solve_block")
        return U, list(st)

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.iga_solver_destroy(self._h)
            self._h = None


def ksp_gmres(matvec, b, pc, rtol=1e-10, atol=1e-10, restart=30, maxit=10000):
    """Host-driven GMRES on device vectors with PETSc 3.18's KSPGMRES
    defaults -- the outer KSP the reference's drivers get from
    KSPSetFromOptions (matrix_free/iteration_number_fempc_mf.cpp:190-221,
    matrix_free_gpu/matrixfreegpu.cu:264-301): zero initial guess, left
    preconditioning, the preconditioned residual norm, classical Gram-Schmidt
    without refinement, restart 30, and the stop ||r_k|| <= max(rtol ||r_0||,
    atol).  matvec(v) and pc(v) return new float64 tensors (e.g.
    Operator.apply, the MATSHELL MatMult, and Solver.apply_pc, the PCSHELL
    apply).  For users without PETSc: it drives the same entry points a
    PETSc KSP would.  Returns (x, iterations)."""
    import numpy as np
    import torch
    x = torch.zeros_like(b)
    its, norm0 = 0, None
    while True:
        r = pc(b - matvec(x))
        beta = torch.linalg.vector_norm(r).item()
        if norm0 is None:
            norm0 = beta
        tol = max(rtol * norm0, atol)
        if beta <= tol or its >= maxit:
            return x, its
        V = torch.zeros((restart + 1, b.numel()), dtype=b.dtype, device=b.device)
        V[0] = r / beta
        H = np.zeros((restart + 1, restart))
        g = np.zeros(restart + 1)
        g[0] = beta
        cs, sn = np.zeros(restart), np.zeros(restart)
        k = 0
        for j in range(restart):
            w = pc(matvec(V[j]))
            h = V[:j + 1] @ w
            w = w - V[:j + 1].T @ h
            H[:j + 1, j] = h.cpu().numpy()
            H[j + 1, j] = torch.linalg.vector_norm(w).item()
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            d = float(np.hypot(H[j, j], H[j + 1, j]))
            cs[j], sn[j] = H[j, j] / d, H[j + 1, j] / d
            H[j, j], H[j + 1, j] = d, 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] *= cs[j]
            its += 1
            k = j + 1
            if H[j, j] == 0.0 or abs(g[j + 1]) <= tol or its >= maxit:
                break
            V[j + 1] = w / torch.linalg.vector_norm(w)
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
        x = x + V[:k].T @ torch.from_numpy(y).to(b.device, b.dtype)
        if abs(g[k]) <= tol or its >= maxit:
            return x, its
