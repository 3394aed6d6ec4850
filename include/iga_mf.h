/*
 * iga_mf.h -- C ABI of the MI355X (gfx950) matrix-free isogeometric Poisson
 * operator: Bezier-extracted (NURBS) basis, Gauss quadrature, element action
 * and global scatter/gather on one GPU, plus the solver around it.
 *
 * Plain C types only: device buffers are passed as pointers the caller owns,
 * streams as `void*` (a hipStream_t; NULL = the null stream), sizes as ints.
 * Every call returns an iga_status; nothing prints or exits.
 *
 * Conventions (those of the reference, SeaVegetable/Parallel-Computing):
 *   - functions (control points) numbered x-fastest, A = j*nfx + i,
 *     nfx = nelem_x + p, nfy = nelem_y + q; vectors have nfx*nfy entries,
 *     Dirichlet functions included (matrix_free_gpu/matrixfreegpu.cu:68-78);
 *   - elements numbered x-fastest, local basis a = jloc*(p+1) + iloc,
 *     quadrature points y-outer, x-inner (matrix_free/src/ElementMF.cpp:
 *     270-276);
 *   - the operator is the reference's, sign included:
 *       y = P_int * sum_e P_e^T K_e x[IEN_e],
 *       K_e = -sum_q kappa_q w_q det(J_q) grad R_a . grad R_b,
 *     x gathered for all functions, rows of Dirichlet functions zeroed
 *     (matrix_free/src/GlobalAssemblyMF.cpp:141-210).
 */
#ifndef IGA_MF_H
#define IGA_MF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status -------------------------------------------------------------
 * Replaces the reference's print-and-exit error handling
 * (matrix_free_gpu/include/memory.cuh:9-12, matrix_free_gpu/src/mult.cu:
 * 510-515, 528-532, 550-557; matrix_free/src/BernsteinBasis.cpp:67-70;
 * QuadraturePoint.cpp:48-51; ControlPointGenerator.cpp:11-15). */
typedef enum {
  IGA_OK = 0,
  IGA_ERR_INVALID_ARGUMENT = 1,   /* NULL where a buffer is required, bad enum */
  IGA_ERR_INVALID_DEGREE = 2,     /* p or q outside [1,8] */
  IGA_ERR_INVALID_QUADRATURE = 3, /* quadrature points outside [1,10] */
  IGA_ERR_TOO_FEW_ELEMENTS = 4,   /* nelem < p + 2 (nfunc < 2p + 2) */
  IGA_ERR_BAD_MESH = 5,           /* IEN / ID / Dir / extraction inconsistent */
  IGA_ERR_UNSUPPORTED = 6,        /* valid input with no kernel built for it */
  IGA_ERR_HIP = 7,                /* a HIP runtime call failed */
  IGA_ERR_OUT_OF_MEMORY = 8,
  IGA_ERR_NOT_FINITE = 9,         /* non-finite value in an input vector */
  IGA_ERR_DIVERGED = 10,          /* residual grew past divtol */
  IGA_ERR_NOT_CONVERGED = 11,     /* maxit reached */
  IGA_ERR_SIZE_LIMIT = 12         /* request larger than a fixed capacity */
} iga_status;

const char *iga_status_string(int status);

/* ---- mesh description ---------------------------------------------------
 * Host arrays, read during iga_op_create only.  The fields are what the
 * reference's MatMulMF / MatrixFreeMatMultCUDA receive
 * (matrix_free_gpu/include/GlobalAssemblyMF.cuh:36-48,
 * matrix_free_gpu/include/mult.cuh:16-24); the Bernstein tables and Gauss
 * rule the reference also passes are rebuilt from p and nqp. */
typedef struct {
  int p, q;                    /* degrees */
  int nelem_x, nelem_y;        /* elements per direction */
  int nqp_x, nqp_y;            /* Gauss points per direction (reference: p+1) */
  const double *cp;            /* [nfunc][2] control points (x,y) */
  const double *weights;       /* [nfunc] NURBS weights, NULL = all one */
  const int *ien;              /* [E][(p+1)(q+1)], NULL = tensor-product IEN */
  const int *id;               /* [nfunc] equation numbers, -1 = Dirichlet */
  const int *dir;              /* [ndir] Dirichlet functions */
  int ndir;
  const double *extraction_x;  /* [nelem_x][(p+1)^2], row = B-spline */
  const double *extraction_y;  /* [nelem_y][(q+1)^2] */
  const double *elem_size_x;   /* [nelem_x] parametric element sizes */
  const double *elem_size_y;   /* [nelem_y] */
  const double *kappa;         /* [E][nqp_x*nqp_y] conductivity, NULL = 1 */
} iga_mesh_desc;

/* ---- mesh generation (the reference's preprocessing, in memory) ---------
 * Replaces matrix_free/preprocess.cpp:16-98 (knots, IEN, ID, control points,
 * extraction) and matrix_free_gpu/src/Partition.cpp:238-281 (serial
 * renumbering); files are not written. */
typedef struct iga_mesh iga_mesh;

typedef enum {
  IGA_GEOM_UNIT_SQUARE = 0,     /* [0,L]^2, identity map, Greville points */
  IGA_GEOM_QUARTER_ANNULUS = 1, /* r in [r0,r1], theta in [0,pi/2], exact NURBS */
  IGA_GEOM_PERTURBED_SQUARE = 2 /* interior points moved U[-a h, a h], weights U[w0,w1] */
} iga_geometry;

typedef struct {
  iga_geometry kind;
  int p;                  /* degree in both directions */
  int nelem_x, nelem_y;
  double L;               /* unit square side */
  double r0, r1;          /* annulus radii */
  double perturb;         /* perturbation amplitude as a fraction of h */
  double w_lo, w_hi;      /* weight range of the perturbed square */
  uint64_t seed;          /* splitmix64 seed of the perturbed square */
} iga_mesh_params;

int iga_mesh_build(const iga_mesh_params *params, iga_mesh **out);
/* Fill a descriptor that points into the mesh's arrays (valid until
 * iga_mesh_destroy). */
int iga_mesh_describe(const iga_mesh *m, iga_mesh_desc *desc);
int iga_mesh_destroy(iga_mesh *m);
/* The patch's geometry map at n parametric points uv[n][2] (host), xy[n][2]
 * out: x(u,v) = sum_A w_A N_A(u,v) P_A / sum_A w_A N_A(u,v), the B-splines
 * from the mesh's Bezier extraction operators, the parameter domain
 * [0, sum h1] x [0, sum h2].  No counterpart in the reference, whose domains
 * are rectangles; it maps the reference's FEM-preconditioner node grid
 * (iga_fem_reference_nodes on [0,1]^2) onto curved patches such as the
 * quarter annulus.  IGA_ERR_INVALID_ARGUMENT for a point outside the domain. */
int iga_mesh_eval(const iga_mesh *m, int64_t n, const double *uv, double *xy);

/* ---- the reference's on-disk partition format --------------------------
 * A partition file of FileManager::WritePartition
 * (matrix_free/src/FileManager.cpp:3-69, read there by ReadPartition
 * :114-280): lines "nlocalfunc: N", "nlocalelemx: N", "nlocalelemy: N", then
 * named sections (ID, ghostID, Dir, IEN, ElemSize1, ElemSize2, CP,
 * NURBSExtraction1, NURBSExtraction2), each followed by one line of
 * space-separated values; reals with 16 significant digits.  The degrees
 * are not in it (they are in the preprocess info file).  A file without the
 * element sizes or extraction operators (the short WritePartition form,
 * :71-112) gives IGA_ERR_BAD_MESH. */
int iga_mesh_read_partition(const char *filename, int p, int q, iga_mesh **out);  /* reals
   that are not finite: IGA_ERR_NOT_FINITE */
/* Write m in the full form (ghostID empty: one partition).  The format has
 * no NURBS weights: IGA_ERR_UNSUPPORTED for a mesh that has them. */
int iga_mesh_write_partition(const iga_mesh *m, const char *filename);
typedef struct {
  int p, q;
  double Lx, Ly;
  int nelem_x, nelem_y, part_num_1d, dim;
  char base_name[256];
} iga_preprocess_info;
/* The preprocess info file (FileManager::WritePreprocessInfo / Read...,
 * FileManager.cpp:375-448); the partition of rank r is
 * base_name + "_" + r + ".txt" (:450-452). */
int iga_read_preprocess_info(const char *filename, iga_preprocess_info *info);

/* ---- the FEM preconditioner's files --------------------------------------
 * Replaces FileManager's readers of the GPU driver's FEM set-up
 * (matrix_free_gpu/matrixfreegpu.cu:139-188).
 *
 * A FEM partition file, the short WritePartition form
 * (matrix_free/src/FileManager.cpp:71-112, read by ReadPartition :282-373):
 * "nlocalfunc: N", "nlocalelemx: N", "nlocalelemy: N", then the sections ID,
 * Dir, IEN (4 local indices per Q1 element) and CP (x, y; 16 significant
 * digits).  ID and CP list the owned functions first, then the ghosts
 * (matrix_free_gpu/src/Partition.cpp:395-466). */
typedef struct iga_fem_partition iga_fem_partition;
typedef struct {
  int nlocalfunc, nlocalelemx, nlocalelemy;
  int nid;            /* owned + ghost functions: entries of id, pairs of cp */
  int ndir;
  const int *id;      /* [nid] partition-numbered IDs, -1 = Dirichlet */
  const int *dir;     /* [ndir] */
  const int *ien;     /* [nlocalelemx * nlocalelemy][4] indices into id / cp */
  const double *cp;   /* [nid][2] */
} iga_fem_partition_desc;
/* IGA_ERR_BAD_MESH for missing sections or inconsistent sizes, IGA_ERR_NOT_FINITE
 * for a non-finite coordinate (the reference's p = 2 FEM preprocessing
 * writes "nan": AbscissaeGenerator averages over zero knots). */
int iga_fem_read_partition(const char *filename, iga_fem_partition **out);
int iga_fem_partition_view(const iga_fem_partition *fp, iga_fem_partition_desc *desc);
int iga_fem_partition_destroy(iga_fem_partition *fp);
/* The dry run's sparsity file coordinate_nonzero_<rank>.txt ("nnz: N",
 * "rows: ...", "cols: ...": matrix_free_gpu/src/FileManager.cpp:684-749,
 * written by GlobalAssemblyDR.cpp:3-26).  *nnz receives the count; with
 * cap < *nnz nothing is copied and IGA_ERR_SIZE_LIMIT returned (call with
 * cap = 0 to size the buffers). */
int iga_read_nonzero_coordinate(const char *filename, int64_t cap, int *rows, int *cols,
                                int64_t *nnz);
int iga_write_nonzero_coordinate(const char *filename, int64_t nnz, const int *rows,
                                 const int *cols);
/* new_to_old_mapping.txt, one line of the partition numbering's original
 * indices (matrix_free_gpu/src/FileManager.cpp:751-786, written by
 * Partition.cpp:342-365); same size protocol. */
int iga_read_new_to_old(const char *filename, int64_t cap, int *map, int64_t *n);
int iga_write_new_to_old(const char *filename, int64_t n, const int *map);
/* The GPU driver's merge (matrix_free_gpu/matrixfreegpu.cu:149-188): the
 * coordinate_nonzero_<r>.txt files of ranks 0..nparts-1 in directory dir,
 * mapped through dir/new_to_old_mapping.txt to the natural numbering and
 * sorted lexicographically: the FEM matrix's COO pattern, the one
 * iga_fem_assemble computes.  The serial preprocessing writes no mapping
 * file; without one the numbering is taken as natural (the reference would
 * read an empty map).  Same size protocol. */
int iga_fem_pattern_from_files(const char *dir, int nparts, int64_t cap, int *rows, int *cols,
                               int64_t *nnz);

/* splitmix64 -> Box-Muller, the generator of every synthetic input */
int iga_rng_normal(uint64_t seed, int64_t n, double *out);
int iga_rng_uniform(uint64_t seed, int64_t n, double lo, double hi, double *out);

/* ---- the operator -------------------------------------------------------
 * Replaces GlobalAssemblyMF::MatMulMF (.cu) and MatrixFreeMatMultCUDA
 * (matrix_free_gpu/src/GlobalAssemblyMF.cu:141-242,
 * matrix_free_gpu/src/mult.cu:518-558), which re-upload every mesh array on
 * every apply: here the mesh is uploaded once and the per-quadrature-point
 * factors G_q = kappa_q w_q det(J_q) J_q^-1 J_q^-T are precomputed in FP64. */
typedef struct iga_op iga_op;

typedef enum {
  IGA_F64 = 0,
  IGA_F32 = 1,
  IGA_F32_G16 = 2  /* FP32 vectors and arithmetic on the FP16 copy of the
                      geometry factors (iga_op_opts.f16_geometry; SF kernel,
                      16x16 tiles or variant 3, not general rational meshes):
                      half the operator bytes, G_q rounded to 11 bits (a
                      symmetric perturbation of A of relative size <= 2^-11) */
} iga_dtype;

typedef enum {
  IGA_KERNEL_SF = 0,  /* sum-factorised element action */
  IGA_KERNEL_EBE = 1  /* element-by-element dense action */
} iga_kernel;

typedef enum {
  IGA_ASSEMBLY_IEN_ATOMIC = 0,  /* scatter-add through IEN
                                   (matrix_free_gpu/src/mult.cu:454-461) */
  IGA_ASSEMBLY_INVLM_GATHER = 1 /* atomics-free gather over the inverse LM
                                   (matrix_free_gpu_dbd/src/mult.cu:341-450);
                                   bitwise reproducible */
} iga_assembly;

typedef struct {
  int device;       /* HIP device ordinal */
  int want_f64;     /* keep FP64 G_q (default 1) */
  int want_f32;     /* keep FP32 G_q (default 1) */
  int tile_x_f64, tile_y_f64; /* element tile of the FP64 kernels, 0 = default
                                 (16x16; 16x8 with variant 3); 16x16 and 32x16
                                 for every degree, 32x32 and 64x8 for p <= 4;
                                 general rational meshes 16x16, and 32x16 for
                                 p <= 4 */
  int tile_x_f32, tile_y_f32; /* element tile of the FP32 kernels, same
                                 choices and default */
  int variant_f64, variant_f32; /* 0: one workgroup per tile; 1: persistent
                                   workgroups prefetching the next tile (SF);
                                   2: one workgroup per tile with a register
                                   budget of four waves per SIMD (FP32 SF,
                                   16x16, p <= 4; EBE applies use 0);
                                   3: two threads per element (SF) on 16x8
                                   tiles, for high degrees and FP64; general
                                   rational meshes accept 0 and 3;
                                   4 (FP32 only): the element operator on the
                                   matrix cores (SF, p = 3, 16x16, not general
                                   rational; EBE applies use 0);
                                   5 (FP32 only): persistent workgroups that
                                   stage the next tile's G_q in LDS by direct
                                   global-to-LDS loads (SF, p <= 3, 16x16, not
                                   general rational; its FP16-G_q form at
                                   p = 3; EBE applies use 0);
                                   -1 (default): 3 for FP32 at p = 3 (not
                                   general rational) and p >= 5, and for FP64
                                   at p = 3 (not general rational); 0 when a
                                   tile other than 16x8 is given, and
                                   otherwise */
  int f16_geometry;   /* also keep an FP16 copy of the FP32 G_q, divided by a
                         power of two that maps their largest magnitude to
                         2^14 (IGA_F32_G16 applies; default 0) */
  int stream_g_f32;   /* FP32 SF applies on 16x8 tiles (variant 3): 1 load
                         G_q non-temporally, 0 with normal loads, -1
                         (default) non-temporally when the FP32 G_q take
                         1 GiB or more (four times the device's last-level
                         cache, so that none of them stays cached from one
                         apply to the next); the result is the same */
} iga_op_opts;

typedef struct {
  int nfunc, nfx, nfy, nelem, nqp, nloc, ndir;
  int rational;               /* 0 unit weights, 1 tensor-product weights, 2 general */
  int tile_x_f64, tile_y_f64, tile_x_f32, tile_y_f32;
  int64_t device_bytes;       /* device memory held by the operator */
  double setup_ms;            /* wall time of iga_op_create */
  int variant_f64, variant_f32; /* apply variant in use (see iga_op_opts) */
  int f16_geometry;           /* the FP16 copy of G_q exists */
  double f16_scale;           /* FP16 G_q = G_q / f16_scale (a power of two) */
  int stream_g_f32;           /* FP32 SF applies load G_q non-temporally */
} iga_op_info;

void iga_op_opts_default(iga_op_opts *o);
int iga_op_create(const iga_mesh_desc *m, const iga_op_opts *o, iga_op **out);
int iga_op_destroy(iga_op *op);
int iga_op_get_info(const iga_op *op, iga_op_info *info);

/* y = A x on the device, y overwritten (the caller's VecSet(y,0) of
 * MyMatMult, matrix_free_gpu/matrixfreegpu.cu:52-66, is folded in).
 * x, y: nfunc entries of the dtype, device memory, not aliased.
 * Asynchronous on `stream`; allocates nothing and never synchronises, so it
 * can be captured in a HIP graph.  Calls on one op must not overlap when
 * they use the gather assembly (the op owns its halo buffer). */
int iga_op_apply(iga_op *op, iga_dtype dt, iga_kernel kern, iga_assembly asmb,
                 const void *x, void *y, void *stream);
/* y = A x as iga_op_apply, and the energy product x.y = x^T P_int A x (the
 * Dirichlet rows of y are zero) accumulated inside the same kernel, written
 * to *xy (one double, device memory) on `stream`.  Replaces the MatMult +
 * VecDot pair (p, Ap) of the reference's KSPCG iteration around MyMatMult
 * (matrix_free_gpu/matrixfreegpu.cu:52-66, 264-301).  The per-tile partials
 * live in the op: calls on one op must not overlap.  IGA_ERR_UNSUPPORTED for
 * apply variant 1. */
int iga_op_apply_dot(iga_op *op, iga_dtype dt, iga_kernel kern, iga_assembly asmb,
                     const void *x, void *y, double *xy, void *stream);

/* Average duration of the element kernel alone (k_apply or its variant,
 * without the seam kernel), from HIP events around nlaunch back-to-back
 * launches on the stream after one warm-up launch.  y (and, in gather
 * form, the halo slots) receives partial results: a timing call, not an
 * apply.  No counterpart in the reference. */
int iga_op_time_kernel(iga_op *op, int dt, int kern, int asmb, const void *x, void *y,
                       int nlaunch, void *stream, double *avg_ms);
/* Kernel timing of the element-action kernel (k_apply) with HIP events on
 * the stream it runs on.  No counterpart in the reference, which times whole
 * solves only (PetscTime, matrix_free_gpu/matrixfreegpu.cu:299-306).
 * enable = 1 starts recording (up to max_calls applies, earlier records
 * dropped), 0 stops; read synchronises on the recorded events. */
int iga_op_profile(iga_op *op, int enable, int max_calls);
int iga_op_profile_read(iga_op *op, int *count, double *total_ms, double *min_ms,
                        double *max_ms);

/* Load vector F (FP64, device) with the reference's F[Dir] = 0
 * (matrix_free_gpu/src/mult.cu:233-336, matrix_free/src/LocalAssemblyMF.cpp:
 * 3-29).  kind: 0 f = x(1-x)y(1-y) (the reference's), 1 f = value,
 * 2 f = value*sin(pi x)sin(pi y). */
int iga_op_load(iga_op *op, int kind, double value, double *F, void *stream);
/* The same load vector for a right-hand side given by its values at the
 * quadrature points: fq [E][nqp] FP64 device, elements and points in the
 * reference's order (those of iga_op_qp_coords).  The reference hard-codes
 * f (matrix_free/include/LocalAssemblyMF.hpp:52-55, matrix_free_gpu/src/
 * mult.cu:228-231); this form takes any f, e.g. one defined on the
 * parametric coordinates. */
int iga_op_load_qp(iga_op *op, const double *fq, double *F, void *stream);

/* Physical coordinates of every quadrature point, xq[e][q][2] (FP64,
 * device, E * nqp * 2 values), elements and points in the reference's order
 * (the x_q of matrix_free/src/ElementMF.cpp:106-207): for building a
 * per-point kappa field on the device. */
int iga_op_qp_coords(iga_op *op, double *xq, void *stream);
/* ---- per-qp conductivity and its gradient (extension; kappa = 1 is the
 * reference).  kappa: [E][nqp] FP64 device; G_q is rebuilt on the device in
 * every precision the op keeps.  The FP16 copy (f16_geometry) gets a new
 * power-of-two scale, which is read back to the host: the call then
 * synchronises the stream.  IGA_ERR_UNSUPPORTED for an op without FP64
 * geometry (want_f64 = 0); IGA_ERR_NOT_FINITE when the new factors are not
 * finite. */
int iga_op_set_kappa(iga_op *op, const double *kappa, void *stream);
/* For y = A(kappa) x and an output cotangent g: grad_x = A^T g
 * (= K~ P_int g; equals A g on interior rows when g vanishes on Dirichlet
 * rows) and grad_kappa[e][q] = -w_q det(J_q) grad x_q . grad (P_int g)_q.
 * Either output may be NULL.  FP64 device buffers. */
int iga_op_apply_backward(iga_op *op, const double *x, const double *g,
                          double *grad_x, double *grad_kappa, void *stream);

/* ---- multi right-hand side ----------------------------------------------
 * Y[:,j] = A X[:,j] for k right-hand sides stored column-major with leading
 * dimension ld >= nfunc (FP32, device), Y overwritten.  No counterpart in the
 * reference (one MatMult per vector there); the operator's G_q and tables
 * are read once per element for each chunk of columns.  Seam sums use
 * atomics, so the result is not bitwise reproducible (FP32 rounding order).
 * Not available on general rational meshes (IGA_ERR_UNSUPPORTED).  Uses the
 * configuration chosen by iga_op_tune_block for this k when there is one. */
int iga_op_apply_block(iga_op *op, int k, int ld, const float *X, float *Y,
                       void *stream);
/* Autotuner: times every compiled (element tile, rhs chunk) configuration
 * valid for the degree and k on the given buffers (reps launches each, HIP
 * events), keeps the fastest for this k and reports it. */
int iga_op_tune_block(iga_op *op, int k, int ld, const float *X, float *Y,
                      void *stream, int reps, int *best_cfg, double *best_ms);
/* Block configurations: count, shape (element tile, rhs chunk per
 * workgroup), and the per-k choice (set cfg = -1 to return k to the
 * built-in default), so a tuning result can be stored and restored. */
int iga_block_config_count(void);
int iga_block_config_desc(int cfg, int *tile_x, int *tile_y, int *rhs_chunk);
/* *matrix_core = 1 when configuration cfg runs on the matrix cores
 * (v_mfma_f32_16x16x4f32, the element operator as two dense products with
 * the 16 right-hand sides of a chunk as columns; degrees up to 4), 0 for the
 * VALU kernel.  No counterpart in the reference, which applies one vector
 * per MatMult (matrix_free_gpu/matrixfreegpu.cu:52-66). */
int iga_block_config_matrix_core(int cfg, int *matrix_core);
int iga_op_set_block_config(iga_op *op, int k, int cfg);
int iga_op_get_block_config(const iga_op *op, int k, int *cfg);

/* ---- FEM preconditioner (elem2coo -> coo2aij) ---------------------------
 * Q1 stiffness on the nfx x nfy node grid `nodes` ([nfx*nfy][2] host),
 * assembled on the device by a deterministic sort-reduce into CSR:
 * replaces Elem2COOGenerator + AssembleStiffnessKernel + DirichletBCKKernel +
 * MatSetPreallocationCOO/MatSetValuesCOO (matrix_free_gpu/src/
 * Elem2COOGenerator.cpp:3-69, assembly.cu:41-99, GlobalAssembly.cu:3-108).
 * Row/column indices are those of the reference's sorted COO pattern. */
typedef struct iga_fem iga_fem;
typedef struct {
  int nrows;
  int64_t nnz;
  const int *rowptr;      /* device [nrows+1] */
  const int *col;         /* device [nnz] */
  const double *val;      /* device [nnz] */
  const int *coo_row;     /* device [nnz] COO row of each entry */
  const int *elem2coo;    /* device [E_fem][4][4], -1 = dropped */
  const int *dir2coo;     /* device [ndir] */
  int ndir;
} iga_csr_view;
/* Reference FEM-PC node grid (matrix_free_gpu/matrixfreegpu.cu:104-122) for a
 * p x q patch with nelem_x x nelem_y IGA elements on [0,Lx]x[0,Ly]; nodes is
 * [(nelem_x+p)(nelem_y+q)][2] host memory. */
int iga_fem_reference_nodes(int p, int q, int nelem_x, int nelem_y, double Lx,
                            double Ly, double *nodes);
int iga_fem_assemble(int nfx, int nfy, const double *nodes, int device,
                     iga_fem **out);
int iga_fem_view(const iga_fem *f, iga_csr_view *v);
int iga_fem_destroy(iga_fem *f);

/* ---- solver -------------------------------------------------------------
 * Replaces the KSP + MATSHELL + PCSHELL set-up of
 * matrix_free_gpu/matrixfreegpu.cu:257-306 (zero initial guess, rtol/abstol
 * convergence test on the unpreconditioned residual). */
typedef enum {
  IGA_PC_NONE = 0,
  IGA_PC_JACOBI = 1,     /* diagonal of A */
  IGA_PC_FEM_CG = 2,     /* inner CG + Jacobi on the Q1 FEM matrix
                            (matrix_free_gpu/matrixfreegpu.cu:18-24,273-289) */
  IGA_PC_FEM_CHEB = 3,   /* fixed-degree Chebyshev-Jacobi polynomial of the Q1
                            matrix (a linear preconditioner; interval
                            [lmax/30, lmax], lmax the Gershgorin bound of
                            D^-1 K); no counterpart in the reference */
  IGA_PC_FEM_ILU = 4,    /* ILU(0) of the Q1 matrix in natural ordering (the
                            -pc_type ilu of the reference's plain-method runs,
                            PETSc PCILU with 0 levels), factored on the host
                            at set-up; triangular solves level-scheduled in
                            one workgroup, so meant for small meshes (tens of
                            thousands of functions) */
  IGA_PC_FEM_MG = 5      /* one V-cycle of geometric multigrid on the Q1
                            matrix: the approximation of K^-1 that the
                            reference's inner CG computes, at a cost linear in
                            the grid (structured 9-point stencils, Galerkin
                            coarse grids, Chebyshev-Jacobi smoothing of degree
                            2, coarsest grid solved exactly); a symmetric
                            linear preconditioner, so the outer CG keeps the
                            standard beta.  The Q1 matrix's Dirichlet set must
                            be the boundary ring (as iga_fem_assemble builds
                            it); FP64 V-cycle in the FP64 solve, FP32 in the
                            FP32 inner solve */
} iga_pc_kind;

typedef enum {
  IGA_SOLVE_PCG_F64 = 0,        /* plain FP64 preconditioned CG */
  IGA_SOLVE_PCG_F32_REFINE = 1  /* FP32 PCG under FP64 defect correction
                                   (reliable updates of the true residual) */
} iga_solve_kind;

typedef struct {
  iga_solve_kind kind;
  iga_pc_kind pc;
  iga_kernel kernel;
  iga_assembly assembly;
  double rtol, abstol, divtol;  /* on ||f - A u||_2 (FP64) */
  int maxit;                    /* total inner iterations cap */
  double inner_rtol;            /* F32_REFINE: residual reduction between
                                   reliable (FP64) updates; 0 (default) =
                                   automatic: 0.1 with Jacobi or none (fewer
                                   FP64 residuals than 0.5 at the same inner
                                   iteration count, fewer FP32 rounding
                                   losses than longer recursions), 1e-4 with
                                   a FEM preconditioner (few, costly
                                   iterations) */
  int inner_maxit;
  int pc_inner_maxit;           /* FEM_CG: inner CG iteration cap */
  double pc_inner_rtol;         /* FEM_CG: inner CG tolerance */
  int pc_cheb_degree;           /* FEM_CHEB: polynomial degree */
  int use_graph;                /* capture the iteration body in a HIP graph */
  int inner_f16_geometry;       /* F32_REFINE: the inner FP32 PCG applies
                                   IGA_F32_G16 (the op must keep the FP16
                                   copy); the FP64 residual of the reliable
                                   updates removes its perturbation (default
                                   0) */
  int fused;                    /* 0: the p.Ap product by its own kernel
                                   after the apply;
                                   1 (default): p.Ap accumulated inside the
                                   apply kernel (same recurrences);
                                   2: single-reduction CG for Jacobi / no
                                   preconditioner (Chronopoulos-Gear
                                   recurrences, p.Ap formed from u.Au and
                                   u.s, u.Au inside the apply): the apply plus
                                   one vector kernel per iteration.
                                   Apply variant 1 cannot fuse: 0 is used. */
  double pc_inner_abstol;       /* FEM_CG: absolute floor of the inner CG's
                                   stopping test, which (like PETSc's KSPCG
                                   default, KSP_NORM_PRECONDITIONED) compares
                                   ||D^-1 r|| with max(pc_inner_rtol
                                   ||D^-1 b||, pc_inner_abstol); the reference
                                   drivers set 1e-10 for both
                                   (matrix_free/iteration_number_fempc_mf.cpp:
                                   214-220); default 0 */
} iga_solver_opts;

typedef struct {
  int iterations;        /* total (inner) CG iterations */
  int outer_iterations;  /* defect-correction steps (1 for plain FP64) */
  double rel_residual;   /* final ||f - A u|| / ||f|| in FP64 */
  double solve_ms;       /* wall time of the solve on the stream */
  int status;
} iga_solve_stats;

typedef struct iga_solver iga_solver;
/* Defaults: PCG_F64, Jacobi, SF kernel, INVLM gather; rtol 1e-8, abstol 0,
   divtol 1e4, maxit 100000; inner_rtol 0 (automatic), inner_maxit 100000;
   pc_inner_maxit 20, pc_inner_rtol 1e-2, pc_cheb_degree 8; use_graph 1,
   inner_f16_geometry 0, fused 1, pc_inner_abstol 0.  Host code only (no
   device call). */
void iga_solver_opts_default(iga_solver_opts *o);
int iga_solver_create(iga_op *op, const iga_fem *fem, const iga_solver_opts *o,
                      iga_solver **out);
/* Solve A u = f; f, u FP64 device (u overwritten, zero initial guess). */
int iga_solver_solve(iga_solver *s, const double *f, double *u, void *stream,
                     iga_solve_stats *stats);
/* Block solve: A U[:,j] = F[:,j] for k FP64 columns (leading dimension ld >=
 * nfunc, U overwritten, zero initial guesses).  No counterpart in the
 * reference (one KSPSolve per right-hand side there).  Each column is an
 * FP32 PCG with FP64 reliable updates (the solver's kind must be
 * IGA_SOLVE_PCG_F32_REFINE, preconditioner Jacobi or none); all columns
 * advance together on the multi-RHS apply, a converged column is frozen.
 * stats: k entries (solve_ms is the whole block's) or NULL.  Returns
 * IGA_OK or the status of the first column that failed. */
int iga_solver_solve_block(iga_solver *s, int k, int ld, const double *F, double *U,
                           void *stream, iga_solve_stats *stats);
/* z = M^-1 r: one application of the solver's FEM preconditioner, the
 * PCSHELL apply callback MyPCApply(PC, Vec, Vec) of the reference
 * (matrix_free_gpu/matrixfreegpu.cu:18-24, 273-289), for callers running
 * their own Krylov method.  dt: IGA_F64 for a PCG_F64 solver, IGA_F32 for a
 * PCG_F32_REFINE one (the precision its inner solve uses); r, z nfunc device
 * values, not aliased.  Asynchronous on `stream`; must not overlap a solve
 * of the same solver.  IGA_ERR_UNSUPPORTED for Jacobi and none. */
int iga_solver_apply_pc(iga_solver *s, int dt, const void *r, void *z, void *stream);
int iga_solver_destroy(iga_solver *s);

/* ---- device vector helpers (FP64 / FP32, used by drivers and tests) ----- */
int iga_dot_f64(const double *a, const double *b, int64_t n, double *out_host,
                void *stream);
/* Allocation-free form of iga_dot_f64.  No counterpart in the reference; it
 * is for callers that cannot allocate or synchronise (code under HIP graph
 * capture, tight serving loops).  The partial sums (one per 1,024 values)
 * live in a static device array of the library, 65,536 doubles (512 KiB) per
 * GPU held until the process ends, so n <= 67,108,864 (IGA_ERR_SIZE_LIMIT
 * above).  a . b goes to *out (one double, device memory) on `stream`,
 * asynchronously: nothing is allocated and nothing synchronises.  Calls that
 * overlap on one GPU (other streams or threads) share the array and race. */
int iga_dot_f64_noalloc(const double *a, const double *b, int64_t n, double *out,
                        void *stream);

#ifdef __cplusplus
}
#endif
#endif /* IGA_MF_H */
