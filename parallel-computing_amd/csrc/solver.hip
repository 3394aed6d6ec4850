// solver.hip -- preconditioned CG around the matrix-free operator.
//
// Replaces the reference's KSP + MATSHELL + PCSHELL set-up
// (matrix_free_gpu/matrixfreegpu.cu:257-306; CPU twin
// matrix_free/iteration_number_fempc_mf.cpp:193-233): zero initial guess,
// convergence on the unpreconditioned residual ||f - A u||_2 <=
// max(rtol ||f||_2, abstol), divergence when it exceeds divtol ||f||_2.
//
// Everything stays on the device: the CG scalars live in device memory, each
// reduction is a two-level (wavefront, then workgroup) sum whose last
// workgroup folds the partials in index order (deterministic) and resets its
// own counter (graph-replay safe), and every kernel returns at once when the
// device "done" flag is set, so the host can enqueue -- or replay as a HIP
// graph -- a fixed chunk of iterations and check convergence once per chunk.
//
// Two solve kinds:
//   PCG_F64         plain FP64 PCG;
//   PCG_F32_REFINE  FP64 defect correction: r = f - A u in FP64, an FP32 PCG
//                   solves A d = r / ||r|| to inner_rtol (FP32 operator,
//                   FP32 vectors, FP64 dot-product accumulation), u += ||r|| d.
// Preconditioners: Jacobi (diag A), or the reference's FEM preconditioner
// applied by an inner CG + Jacobi on the Q1 matrix (a nonlinear
// preconditioner: the outer iteration then uses the flexible, Polak-Ribiere
// form of beta).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "op_internal.hpp"

struct iga_fem;  // defined in fem.hip; only its CSR view is used here

namespace {

constexpr int RB = 256;    // threads per reduction workgroup
constexpr int RG = 1024;   // reduction workgroups (grid-stride)
constexpr int CHUNK = 16;  // iterations per enqueued chunk / graph

// scalar slots (doubles) of one CG instance
enum { S_RZ = 0, S_PAP, S_ALPHA, S_BETA, S_RR, S_TOL2, S_DIV2, S_RR0, S_NS };
// integer slots
enum { I_DONE = 0, I_ITER, I_MAXIT, I_CNT0, I_CNT1, I_NI };
// I_DONE values: 0 running, 1 converged, 2 iteration limit, 3 diverged / breakdown

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the workgroup; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sm[NV][RB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    v[k] = wave_sum(v[k]);
    if (lane == 0) sm[k][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int i = 0; i < RB / 64; ++i) s += sm[k][i];
      v[k] = s;
    }
  }
}

// Publish per-workgroup partials; returns true in thread 0 of the last
// workgroup to arrive, which then owns the totals (summed in index order).
template <int NV>
__device__ __forceinline__ bool last_block(double (&v)[NV], double *part, int *cnt,
                                           double (&tot)[NV]) {
  __shared__ bool is_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      __hip_atomic_store(&part[k * RG + blockIdx.x], v[k], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const int prev = atomicAdd(cnt, 1);
    is_last = (prev == (int)gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  if (threadIdx.x == 0) {
    __threadfence();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int i = 0; i < (int)gridDim.x; ++i)
        s += __hip_atomic_load(&part[k * RG + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      tot[k] = s;
    }
    *cnt = 0;
  }
  return threadIdx.x == 0;
}

// ---------------------------------------------------------------- CG kernels
// init: r = f (u = 0), z = M r (Jacobi or identity), p = z when z is local;
// rr0 = |f|^2, rz = r.z, tolerances
template <typename V, typename F>
__global__ void k_cg_init(const F *f, double fscale, V *u, V *r, V *z, V *p, const double *dinv,
                          int use_z, int64_t n, double *sc, int *ic, double *part, double rtol2,
                          double abstol2, double divtol2, int maxit) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    const double ri = (double)f[i] * fscale;
    u[i] = V(0);
    r[i] = (V)ri;
    const double zi = dinv ? ri * dinv[i] : ri;
    if (use_z) {
      z[i] = (V)zi;
      p[i] = (V)zi;
    }
    v[0] += (double)(V)ri * (double)(V)ri;
    v[1] += (double)(V)ri * (double)(V)zi;
  }
  block_sum(v);
  double tot[2];
  if (last_block(v, part, ic + I_CNT0, tot)) {
    sc[S_RR0] = tot[0];
    sc[S_RR] = tot[0];
    sc[S_RZ] = tot[1];
    sc[S_TOL2] = fmax(rtol2 * tot[0], abstol2);
    sc[S_DIV2] = divtol2 * tot[0];
    ic[I_ITER] = 0;
    ic[I_MAXIT] = maxit;
    ic[I_DONE] = (tot[0] <= sc[S_TOL2]) ? 1 : 0;
    if (!(tot[0] == tot[0]) || isinf(tot[0])) ic[I_DONE] = 3;
  }
}

// alpha = rz / (p . Ap)
template <typename V>
__global__ void k_cg_pap(const V *p, const V *Ap, int64_t n, double *sc, int *ic, double *part) {
  if (ic[I_DONE]) return;
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    v[0] += (double)p[i] * (double)Ap[i];
  block_sum(v);
  double tot[1];
  if (last_block(v, part, ic + I_CNT0, tot)) {
    sc[S_PAP] = tot[0];
    const double a = sc[S_RZ] / tot[0];
    sc[S_ALPHA] = a;
    if (!(a == a) || isinf(a) || tot[0] == 0.0) ic[I_DONE] = 3;
  }
}

// u += alpha p; r -= alpha Ap; (rold = r before the update when flexible);
// z = dinv r when the preconditioner is local; sums r.r and r.z
template <typename V>
__global__ void k_cg_update(V *u, V *r, const V *p, const V *Ap, V *z, V *rold,
                            const double *dinv, int local_pc, int64_t n, double *sc, int *ic,
                            double *part) {
  if (ic[I_DONE]) return;
  const double a = sc[S_ALPHA];
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    u[i] = (V)((double)u[i] + a * (double)p[i]);
    if (rold) rold[i] = r[i];
    const V ri = (V)((double)r[i] - a * (double)Ap[i]);
    r[i] = ri;
    v[0] += (double)ri * (double)ri;
    if (local_pc) {
      const V zi = dinv ? (V)((double)ri * dinv[i]) : ri;
      z[i] = zi;
      v[1] += (double)ri * (double)zi;
    }
  }
  block_sum(v);
  double tot[2];
  if (last_block(v, part, ic + I_CNT1, tot)) {
    sc[S_RR] = tot[0];
    const int it = ic[I_ITER] + 1;
    ic[I_ITER] = it;
    if (tot[0] <= sc[S_TOL2]) ic[I_DONE] = 1;
    else if (!(tot[0] == tot[0]) || tot[0] > sc[S_DIV2]) ic[I_DONE] = 3;
    else if (it >= ic[I_MAXIT]) ic[I_DONE] = 2;
    if (local_pc) {
      sc[S_BETA] = tot[1] / sc[S_RZ];
      sc[S_RZ] = tot[1];
    }
  }
}

// for a preconditioner applied by its own kernels: beta from z (flexible form
// z.(r - rold) when rold is given)
template <typename V>
__global__ void k_cg_rz(const V *r, const V *z, const V *rold, int64_t n, double *sc, int *ic,
                        double *part) {
  if (ic[I_DONE]) return;
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    v[0] += (double)r[i] * (double)z[i];
    if (rold) v[1] += (double)rold[i] * (double)z[i];
  }
  block_sum(v);
  double tot[2];
  if (last_block(v, part, ic + I_CNT0, tot)) {
    sc[S_BETA] = (tot[0] - tot[1]) / sc[S_RZ];
    sc[S_RZ] = tot[0];
  }
}

// rz after the first preconditioner application (no beta)
template <typename V>
__global__ void k_cg_rz0(const V *r, const V *z, V *p, int64_t n, double *sc, int *ic,
                         double *part) {
  if (ic[I_DONE]) return;
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    v[0] += (double)r[i] * (double)z[i];
    p[i] = z[i];
  }
  block_sum(v);
  double tot[1];
  if (last_block(v, part, ic + I_CNT0, tot)) sc[S_RZ] = tot[0];
}

template <typename V>
__global__ void k_cg_pupdate(V *p, const V *z, int64_t n, const double *sc, const int *ic) {
  if (ic[I_DONE]) return;
  const double b = sc[S_BETA];
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    p[i] = (V)((double)z[i] + b * (double)p[i]);
}

// ---------------------------------------------------------------- outer (FP64)
// r = f - Au, returns |r|^2 and |f|^2 in sc_out[0], sc_out[1]
__global__ void k_residual(const double *f, const double *Au, double *r, int64_t n,
                           double *out, int *cnt, double *part) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    const double ri = f[i] - Au[i];
    r[i] = ri;
    v[0] += ri * ri;
    v[1] += f[i] * f[i];
  }
  block_sum(v);
  double tot[2];
  if (last_block(v, part, cnt, tot)) {
    out[0] = tot[0];
    out[1] = tot[1];
  }
}

// u += scale * d (FP32 correction into the FP64 iterate)
__global__ void k_axpy_f32(double *u, const float *d, double scale, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    u[i] += scale * (double)d[i];
}

__global__ void k_inv(double *d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    d[i] = 1.0 / d[i];
}

__global__ void k_dot(const double *a, const double *b, int64_t n, double *out, int *cnt,
                      double *part) {
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    v[0] += a[i] * b[i];
  block_sum(v);
  double tot[1];
  if (last_block(v, part, cnt, tot)) out[0] = tot[0];
}

// ---------------------------------------------------------------- FEM preconditioner
// CSR SpMV y = K x (one thread per row; the Q1 rows hold at most 9 entries)
template <typename V>
__global__ void k_spmv(const int *rowptr, const int *col, const V *val, const V *x, V *y,
                       int nrows, const int *skip) {
  if (skip && *skip) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  double s = 0.0;
  for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += (double)val[k] * (double)x[col[k]];
  y[i] = (V)s;
}

__global__ void k_csr_diag_inv(const int *rowptr, const int *col, const double *val, double *d,
                               int nrows) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  double v = 0.0;
  for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
    if (col[k] == i) v = val[k];
  d[i] = (v != 0.0) ? 1.0 / v : 1.0;
}

template <typename V>
__global__ void k_cast(const double *a, V *b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    b[i] = (V)a[i];
}

// inner CG on the FEM matrix solving K z = r (z0 = 0, Jacobi, PETSc's default
// preconditioned-norm test |B r_k| <= max(rtol |B r_0|, abstol)); its
// scalars/flags are a second CG instance whose I_DONE is reset by init.
template <typename V>
__global__ void k_pc_init(const V *r, V *zz, V *rin, V *w, V *pp, const double *dinv,
                          int64_t n, double *sc, int *ic, double *part, double rtol2,
                          double abstol2, int maxit, const int *outer_done) {
  if (*outer_done) return;
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    const double ri = (double)r[i];
    zz[i] = V(0);
    rin[i] = r[i];
    const V wi = (V)(ri * dinv[i]);
    w[i] = wi;
    pp[i] = wi;
    v[0] += ri * (double)wi;
    v[1] += (double)wi * (double)wi;
  }
  block_sum(v);
  double tot[2];
  if (last_block(v, part, ic + I_CNT0, tot)) {
    sc[S_RZ] = tot[0];
    sc[S_RR0] = tot[1];
    sc[S_TOL2] = fmax(rtol2 * tot[1], abstol2);
    ic[I_ITER] = 0;
    ic[I_MAXIT] = maxit;
    ic[I_DONE] = (tot[1] <= sc[S_TOL2]) ? 1 : 0;
  }
}

template <typename V>
__global__ void k_pc_update(V *zz, V *rin, V *w, const V *pp, const V *q, const double *dinv,
                            int64_t n, double *sc, int *ic, double *part) {
  if (ic[I_DONE]) return;
  const double a = sc[S_ALPHA];
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    zz[i] = (V)((double)zz[i] + a * (double)pp[i]);
    const double ri = (double)rin[i] - a * (double)q[i];
    rin[i] = (V)ri;
    const V wi = (V)(ri * dinv[i]);
    w[i] = wi;
    v[0] += ri * (double)wi;
    v[1] += (double)wi * (double)wi;
  }
  block_sum(v);
  double tot[2];
  if (last_block(v, part, ic + I_CNT1, tot)) {
    const int it = ic[I_ITER] + 1;
    ic[I_ITER] = it;
    sc[S_BETA] = tot[0] / sc[S_RZ];
    sc[S_RZ] = tot[0];
    if (tot[1] <= sc[S_TOL2] || it >= ic[I_MAXIT]) ic[I_DONE] = 1;
  }
}

inline dim3 rgrid(int64_t n) {
  int64_t b = (n + RB - 1) / RB;
  if (b > RG) b = RG;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

}  // namespace

struct iga_solver {
  iga_op *op = nullptr;
  iga_solver_opts o{};
  int64_t n = 0;
  int device = 0;
  // FP64 state
  double *r = nullptr, *p = nullptr, *Ap = nullptr, *z = nullptr, *rold = nullptr;
  double *Au = nullptr, *dinv = nullptr;  // Jacobi of A (inverse diagonal)
  double *x64 = nullptr;                  // FP64 PCG iterate (owned: graphs capture it)
  // FP32 inner state
  float *d32 = nullptr, *r32 = nullptr, *p32 = nullptr, *Ap32 = nullptr, *z32 = nullptr,
        *rold32 = nullptr;
  // scalars: [0..S_NS) main CG, [S_NS..2S_NS) inner PC CG, then outer slots
  double *sc = nullptr;
  int *ic = nullptr;
  double *part = nullptr;
  // FEM preconditioner (CSR of the reference's Q1 matrix)
  int fem_rows = 0;
  const int *rowptr = nullptr, *col = nullptr;
  const double *val64 = nullptr;
  float *val32 = nullptr;
  double *fdinv = nullptr;
  double *zz = nullptr, *rin = nullptr, *w = nullptr, *pp = nullptr, *q = nullptr;
  float *zz32 = nullptr, *rin32 = nullptr, *w32 = nullptr, *pp32 = nullptr, *q32 = nullptr;
  std::vector<void *> allocs;
  int gen = -1;  // operator generation the Jacobi diagonal was built for
  // graphs of one chunk of iterations (FP64 PCG, FP32 inner PCG)
  hipGraphExec_t g64 = nullptr, g32 = nullptr;
  hipStream_t gstream64 = nullptr, gstream32 = nullptr;
};

namespace {

int salloc(iga_solver *s, void **p, size_t bytes) {
  if (hipMalloc(p, bytes ? bytes : 256) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return IGA_ERR_OUT_OF_MEMORY;
  }
  s->allocs.push_back(*p);
  return IGA_OK;
}

bool local_pc(const iga_solver *s) { return s->o.pc == IGA_PC_NONE || s->o.pc == IGA_PC_JACOBI; }

// one application of the FEM preconditioner z = K^-1 r (inner CG + Jacobi)
template <typename V>
int fem_pc(iga_solver *s, const V *r, V *z, V *rin, V *w, V *pp, V *q, const V *val,
           hipStream_t st) {
  double *sc = s->sc + S_NS;
  int *ic = s->ic + I_NI;
  const int64_t n = s->n;
  const double rt = s->o.pc_inner_rtol;
  hipLaunchKernelGGL((k_pc_init<V>), rgrid(n), dim3(RB), 0, st, r, z, rin, w, pp, s->fdinv, n, sc,
                     ic, s->part, rt * rt, 0.0, s->o.pc_inner_maxit,
                     s->ic + I_DONE);
  const int bs = 256;
  for (int it = 0; it < s->o.pc_inner_maxit; ++it) {
    hipLaunchKernelGGL((k_spmv<V>), dim3((s->fem_rows + bs - 1) / bs), dim3(bs), 0, st, s->rowptr,
                       s->col, val, pp, q, s->fem_rows, ic + I_DONE);
    hipLaunchKernelGGL((k_cg_pap<V>), rgrid(n), dim3(RB), 0, st, pp, q, n, sc, ic, s->part);
    hipLaunchKernelGGL((k_pc_update<V>), rgrid(n), dim3(RB), 0, st, z, rin, w, pp, q, s->fdinv, n,
                       sc, ic, s->part);
    hipLaunchKernelGGL((k_cg_pupdate<V>), rgrid(n), dim3(RB), 0, st, pp, w, n, sc, ic);
  }
  return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
}

// One CG iteration on vectors of type V (FP64 main solve or FP32 inner solve).
template <typename V>
int cg_iteration(iga_solver *s, bool f64, V *u, V *r, V *p, V *Ap, V *z, V *rold,
                 hipStream_t st) {
  const int64_t n = s->n;
  const bool gather = s->o.assembly == IGA_ASSEMBLY_INVLM_GATHER;
  const bool ebe = s->o.kernel == IGA_KERNEL_EBE;
  int rc = igai::apply_raw(s->op, f64, ebe, gather, p, Ap, 0, 1, st, s->ic + I_DONE);
  if (rc) return rc;
  hipLaunchKernelGGL((k_cg_pap<V>), rgrid(n), dim3(RB), 0, st, p, Ap, n, s->sc, s->ic, s->part);
  const bool loc = local_pc(s);
  const double *dinv = s->o.pc == IGA_PC_JACOBI ? s->dinv : nullptr;
  hipLaunchKernelGGL((k_cg_update<V>), rgrid(n), dim3(RB), 0, st, u, r, p, Ap, z,
                     loc ? (V *)nullptr : rold, dinv, loc ? 1 : 0, n, s->sc, s->ic, s->part);
  if (!loc) {
    if (f64)
      rc = fem_pc<double>(s, (const double *)r, (double *)z, s->rin, s->w, s->pp, s->q, s->val64,
                          st);
    else
      rc = fem_pc<float>(s, (const float *)r, (float *)z, s->rin32, s->w32, s->pp32, s->q32,
                         s->val32, st);
    if (rc) return rc;
    hipLaunchKernelGGL((k_cg_rz<V>), rgrid(n), dim3(RB), 0, st, r, z, rold, n, s->sc, s->ic,
                       s->part);
  }
  hipLaunchKernelGGL((k_cg_pupdate<V>), rgrid(n), dim3(RB), 0, st, p, z, n, s->sc, s->ic);
  return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
}

template <typename V>
int cg_chunk(iga_solver *s, bool f64, V *u, V *r, V *p, V *Ap, V *z, V *rold, hipStream_t st) {
  for (int k = 0; k < CHUNK; ++k) {
    int rc = cg_iteration<V>(s, f64, u, r, p, Ap, z, rold, st);
    if (rc) return rc;
  }
  return IGA_OK;
}

// Run a CG to completion: init must already have been enqueued on st.
template <typename V>
int cg_run(iga_solver *s, bool f64, V *u, V *r, V *p, V *Ap, V *z, V *rold, hipStream_t st,
           hipGraphExec_t *gexec, int *iters, int *done) {
  const int64_t n = s->n;
  if (!local_pc(s)) {
    int rc = f64 ? fem_pc<double>(s, (const double *)r, (double *)z, s->rin, s->w, s->pp, s->q,
                                  s->val64, st)
                 : fem_pc<float>(s, (const float *)r, (float *)z, s->rin32, s->w32, s->pp32,
                                 s->q32, s->val32, st);
    if (rc) return rc;
    hipLaunchKernelGGL((k_cg_rz0<V>), rgrid(n), dim3(RB), 0, st, r, z, p, n, s->sc, s->ic, s->part);
  }
  if (s->o.use_graph && !*gexec) {
    // capture one chunk; the graph reads every scalar from device memory
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return IGA_ERR_HIP;
    hipGraph_t g;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipStreamDestroy(cs);
      return IGA_ERR_HIP;
    }
    int rc = cg_chunk<V>(s, f64, u, r, p, Ap, z, rold, cs);
    hipError_t e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc || e != hipSuccess) return rc ? rc : IGA_ERR_HIP;
    e = hipGraphInstantiate(gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      *gexec = nullptr;
      return IGA_ERR_HIP;
    }
  }
  int hic[2];
  for (;;) {
    if (s->o.use_graph) {
      if (hipGraphLaunch(*gexec, st) != hipSuccess) return IGA_ERR_HIP;
    } else {
      int rc = cg_chunk<V>(s, f64, u, r, p, Ap, z, rold, st);
      if (rc) return rc;
    }
    if (hipMemcpyAsync(hic, s->ic, 2 * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return IGA_ERR_HIP;
    if (hic[I_DONE]) break;
  }
  *iters = hic[I_ITER];
  *done = hic[I_DONE];
  return IGA_OK;
}

}  // namespace

extern "C" {

void iga_solver_opts_default(iga_solver_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->kind = IGA_SOLVE_PCG_F64;
  o->pc = IGA_PC_JACOBI;
  o->kernel = IGA_KERNEL_SF;
  o->assembly = IGA_ASSEMBLY_INVLM_GATHER;
  o->rtol = 1e-8;
  o->abstol = 0.0;
  o->divtol = 1e4;
  o->maxit = 100000;
  o->inner_rtol = 1e-4;
  o->inner_maxit = 100000;
  o->pc_inner_maxit = 20;
  o->pc_inner_rtol = 1e-2;
  o->pc_cheb_degree = 8;
  o->use_graph = 1;
}

int iga_solver_destroy(iga_solver *s) {
  if (!s) return IGA_OK;
  (void)hipSetDevice(s->device);
  if (s->g64) (void)hipGraphExecDestroy(s->g64);
  if (s->g32) (void)hipGraphExecDestroy(s->g32);
  for (void *p : s->allocs) (void)hipFree(p);
  delete s;
  return IGA_OK;
}

int iga_solver_create(iga_op *op, const iga_fem *fem, const iga_solver_opts *o_in,
                      iga_solver **out) {
  if (!op || !out) return IGA_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  iga_solver_opts o;
  if (o_in) o = *o_in; else iga_solver_opts_default(&o);
  if (o.kind != IGA_SOLVE_PCG_F64 && o.kind != IGA_SOLVE_PCG_F32_REFINE) return IGA_ERR_INVALID_ARGUMENT;
  if (o.pc != IGA_PC_NONE && o.pc != IGA_PC_JACOBI && o.pc != IGA_PC_FEM_CG)
    return IGA_ERR_UNSUPPORTED;
  if (o.pc == IGA_PC_FEM_CG && !fem) return IGA_ERR_INVALID_ARGUMENT;
  if (o.kind == IGA_SOLVE_PCG_F32_REFINE && !op->G32) return IGA_ERR_UNSUPPORTED;
  if (!(o.rtol >= 0) || !(o.abstol >= 0) || o.maxit < 1 || !(o.inner_rtol > 0) ||
      o.inner_maxit < 1 || (o.pc == IGA_PC_FEM_CG && o.pc_inner_maxit < 1))
    return IGA_ERR_INVALID_ARGUMENT;
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  iga_solver *s = new (std::nothrow) iga_solver;
  if (!s) return IGA_ERR_OUT_OF_MEMORY;
  s->op = op;
  s->o = o;
  s->n = op->nfunc;
  s->device = op->device;
  auto fail = [&](int r) {
    iga_solver_destroy(s);
    return r;
  };
  const size_t n = (size_t)s->n;
#define SA(ptr, bytes)                                  \
  do {                                                  \
    int r_ = salloc(s, (void **)&(ptr), (bytes));       \
    if (r_) return fail(r_);                            \
  } while (0)
  SA(s->r, n * 8);
  SA(s->p, n * 8);
  SA(s->Ap, n * 8);
  SA(s->z, n * 8);
  SA(s->rold, n * 8);
  SA(s->Au, n * 8);
  SA(s->dinv, n * 8);
  SA(s->x64, n * 8);
  SA(s->sc, 64 * 8);
  SA(s->ic, 32 * 4);
  SA(s->part, (size_t)4 * RG * 8);
  if (o.kind == IGA_SOLVE_PCG_F32_REFINE) {
    SA(s->d32, n * 4);
    SA(s->r32, n * 4);
    SA(s->p32, n * 4);
    SA(s->Ap32, n * 4);
    SA(s->z32, n * 4);
    SA(s->rold32, n * 4);
  }
  if (hipMemset(s->ic, 0, 32 * 4) != hipSuccess || hipMemset(s->sc, 0, 64 * 8) != hipSuccess)
    return fail(IGA_ERR_HIP);
  if (o.pc == IGA_PC_FEM_CG) {
    iga_csr_view v;
    if (iga_fem_view(fem, &v) != IGA_OK) return fail(IGA_ERR_INVALID_ARGUMENT);
    if (v.nrows != s->n) return fail(IGA_ERR_BAD_MESH);
    s->fem_rows = v.nrows;
    s->rowptr = v.rowptr;
    s->col = v.col;
    s->val64 = v.val;
    SA(s->fdinv, n * 8);
    SA(s->zz, n * 8);
    SA(s->rin, n * 8);
    SA(s->w, n * 8);
    SA(s->pp, n * 8);
    SA(s->q, n * 8);
    const int bs = 256;
    hipLaunchKernelGGL(k_csr_diag_inv, dim3((v.nrows + bs - 1) / bs), dim3(bs), 0, nullptr,
                       v.rowptr, v.col, v.val, s->fdinv, v.nrows);
    if (o.kind == IGA_SOLVE_PCG_F32_REFINE) {
      SA(s->val32, (size_t)v.nnz * 4);
      SA(s->zz32, n * 4);
      SA(s->rin32, n * 4);
      SA(s->w32, n * 4);
      SA(s->pp32, n * 4);
      SA(s->q32, n * 4);
      hipLaunchKernelGGL((k_cast<float>), rgrid(v.nnz), dim3(RB), 0, nullptr, v.val, s->val32,
                         (int64_t)v.nnz);
    }
  }
#undef SA
  if (hipDeviceSynchronize() != hipSuccess) return fail(IGA_ERR_HIP);
  *out = s;
  return IGA_OK;
}

int iga_solver_solve(iga_solver *s, const double *f, double *u, void *stream,
                     iga_solve_stats *stats) {
  if (!s || !f || !u || f == u) return IGA_ERR_INVALID_ARGUMENT;
  int rc = igai::set_device(s->device);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = s->n;
  const iga_solver_opts &o = s->o;
  iga_solve_stats stt;
  std::memset(&stt, 0, sizeof(stt));
  const bool gather = o.assembly == IGA_ASSEMBLY_INVLM_GATHER;
  const bool ebe = o.kernel == IGA_KERNEL_EBE;
  const bool loc = local_pc(s);
  const double *dinv = o.pc == IGA_PC_JACOBI ? s->dinv : nullptr;
  double *osc = s->sc + 2 * S_NS;  // outer slots: |r|^2, |f|^2
  int *ocnt = s->ic + 2 * I_NI;
  if (s->gen != s->op->gen) {
    // Jacobi of A for the operator's current G_q
    iga_op *op = s->op;
    if (op->ops->diag(s->dinv, op->G64, op->tabx64, op->taby64, op->dirmask, op->nex, op->ney,
                      op->nfx, op->nfy, op->L64, st) != hipSuccess)
      return IGA_ERR_HIP;
    hipLaunchKernelGGL(k_inv, rgrid(n), dim3(RB), 0, st, s->dinv, n);
    s->gen = op->gen;
  }
  if (hipStreamSynchronize(st) != hipSuccess) return IGA_ERR_HIP;
  const auto t0 = std::chrono::steady_clock::now();
  int status = IGA_OK;
  if (o.kind == IGA_SOLVE_PCG_F64) {
    hipLaunchKernelGGL((k_cg_init<double, double>), rgrid(n), dim3(RB), 0, st, f, 1.0, s->x64,
                       s->r, s->z, s->p, dinv, loc ? 1 : 0, n, s->sc, s->ic, s->part,
                       o.rtol * o.rtol, o.abstol * o.abstol, o.divtol * o.divtol, o.maxit);
    int iters = 0, done = 0;
    rc = cg_run<double>(s, true, s->x64, s->r, s->p, s->Ap, s->z, s->rold, st, &s->g64, &iters,
                        &done);
    if (rc) return rc;
    if (hipMemcpyAsync(u, s->x64, n * 8, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return IGA_ERR_HIP;
    stt.iterations = iters;
    stt.outer_iterations = 1;
    status = done == 1 ? IGA_OK : (done == 2 ? IGA_ERR_NOT_CONVERGED : IGA_ERR_DIVERGED);
  } else {
    // FP64 defect correction around an FP32 PCG
    if (hipMemsetAsync(u, 0, n * 8, st) != hipSuccess) return IGA_ERR_HIP;
    hipLaunchKernelGGL(k_residual, rgrid(n), dim3(RB), 0, st, f, u /* = 0 */, s->r, n, osc, ocnt,
                       s->part);
    double h[2];
    if (hipMemcpyAsync(h, osc, 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return IGA_ERR_HIP;
    const double ff = h[1];
    double rr = h[0];
    const double tol = std::fmax(o.rtol * std::sqrt(ff), o.abstol);
    int total = 0, outer = 0;
    status = IGA_ERR_NOT_CONVERGED;
    for (;;) {
      const double rn = std::sqrt(rr);
      if (!(rn == rn) || rn > o.divtol * std::sqrt(ff)) { status = IGA_ERR_DIVERGED; break; }
      if (rn <= tol) { status = IGA_OK; break; }
      if (total >= o.maxit) break;
      // inner: A32 d = r / |r|
      const int imax = std::min(o.inner_maxit, o.maxit - total);
      hipLaunchKernelGGL((k_cg_init<float, double>), rgrid(n), dim3(RB), 0, st, s->r, 1.0 / rn,
                         s->d32, s->r32, s->z32, s->p32, dinv, loc ? 1 : 0, n, s->sc, s->ic,
                         s->part, o.inner_rtol * o.inner_rtol, 0.0, 1e300, imax);
      int iters = 0, done = 0;
      rc = cg_run<float>(s, false, s->d32, s->r32, s->p32, s->Ap32, s->z32, s->rold32, st, &s->g32,
                         &iters, &done);
      if (rc) return rc;
      total += iters;
      ++outer;
      hipLaunchKernelGGL(k_axpy_f32, rgrid(n), dim3(RB), 0, st, u, s->d32, rn, n);
      rc = igai::apply_raw(s->op, true, ebe, gather, u, s->Au, 0, 1, st, nullptr);
      if (rc) return rc;
      hipLaunchKernelGGL(k_residual, rgrid(n), dim3(RB), 0, st, f, s->Au, s->r, n, osc, ocnt,
                         s->part);
      if (hipMemcpyAsync(h, osc, 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return IGA_ERR_HIP;
      rr = h[0];
      if (done == 3) { status = IGA_ERR_DIVERGED; break; }
    }
    stt.iterations = total;
    stt.outer_iterations = outer;
  }
  // final FP64 residual
  rc = igai::apply_raw(s->op, true, ebe, gather, u, s->Au, 0, 1, st, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_residual, rgrid(n), dim3(RB), 0, st, f, s->Au, s->Ap, n, osc, ocnt,
                     s->part);
  double h[2];
  if (hipMemcpyAsync(h, osc, 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return IGA_ERR_HIP;
  const auto t1 = std::chrono::steady_clock::now();
  stt.rel_residual = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
  stt.solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  stt.status = status;
  if (stats) *stats = stt;
  return status;
}

int iga_dot_f64(const double *a, const double *b, int64_t n, double *out_host, void *stream) {
  if (!a || !b || !out_host || n < 0) return IGA_ERR_INVALID_ARGUMENT;
  hipStream_t st = (hipStream_t)stream;
  double *buf = nullptr;
  if (hipMallocAsync((void **)&buf, (RG + 8) * 8, st) != hipSuccess) return IGA_ERR_OUT_OF_MEMORY;
  int *cnt = (int *)(buf + RG + 4);
  if (hipMemsetAsync(cnt, 0, 4, st) != hipSuccess) return IGA_ERR_HIP;
  hipLaunchKernelGGL(k_dot, rgrid(n), dim3(RB), 0, st, a, b, n, buf + RG, cnt, buf);
  double r = 0.0;
  if (hipMemcpyAsync(&r, buf + RG, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return IGA_ERR_HIP;
  (void)hipFreeAsync(buf, st);
  *out_host = r;
  return IGA_OK;
}

}  // extern "C"
