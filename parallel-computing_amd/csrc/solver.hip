// solver.hip -- preconditioned CG around the matrix-free operator.
//
// Replaces the reference's KSP + MATSHELL + PCSHELL set-up
// (matrix_free_gpu/matrixfreegpu.cu:257-306; CPU twin
// matrix_free/iteration_number_fempc_mf.cpp:193-233): zero initial guess,
// convergence on the unpreconditioned residual ||f - A u||_2 <=
// max(rtol ||f||_2, abstol), divergence when it exceeds divtol ||f||_2.
//
// Everything stays on the device.  A CG iteration is
//   apply (Ap = A p) -> K1 partials of p.Ap
//   -> K2 alpha = rz / sum(partials), u, r, z updates, partials of r.r, r.z
//   -> [preconditioner kernels and their r.z partials when it is not local]
//   -> K3 beta, convergence test, p = z + beta p.
// Each kernel writes per-workgroup partial sums; the NEXT kernel folds them
// in every workgroup, in the same fixed order, so all workgroups see the same
// bitwise scalar without atomics, fences or counters (the per-XCD L2s make
// device-scope fences expensive on MI355X).  Scalars that the next iteration
// needs are written by workgroup 0 into slots no workgroup of the same kernel
// reads.  Every kernel returns at once when the device "done" flag is set, so
// the host enqueues -- or replays as a HIP graph -- a fixed chunk of
// iterations and checks convergence once per chunk.
//
// Two solve kinds:
//   PCG_F64         plain FP64 PCG;
//   PCG_F32_REFINE  FP32 PCG (FP32 operator and vectors, FP64 dot-product
//                   accumulation) under FP64 defect correction by reliable
//                   updates: each time its recursive residual has dropped by
//                   inner_rtol, the FP32 iterate is folded into the FP64
//                   solution u, r = f - A u is recomputed with the FP64
//                   operator and the recursion continues from it with the
//                   same search direction.
// Preconditioners: Jacobi (diag A), or the reference's FEM preconditioner
// applied by an inner CG + Jacobi on the Q1 matrix (a nonlinear
// preconditioner: the outer iteration then uses the flexible, Polak-Ribiere
// form of beta).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "op_internal.hpp"
#include "outcome.hpp"

struct iga_fem;  // defined in fem.hip; only its CSR view is used here

namespace {

constexpr int RB = 256;    // threads per workgroup of the vector kernels
constexpr int RG = 512;    // workgroups of the vector kernels (grid-stride)
constexpr int CHUNK = 16;  // iterations per enqueued chunk / graph
// With a preconditioner applied by its own kernels (a V-cycle or an inner
// solve) an iteration is tens of kernels and an order of magnitude longer:
// shorter chunks keep the iterations launched past a stop (the rest of the
// stopping chunk and the one enqueued ahead, each kernel returning at once
// but still dispatched) few, and still hide the host's flag check.
constexpr int CHUNK_PC = 4;
// iterations per chunk near a predicted stop (run_chunks), enqueued as plain
// launches
constexpr int CHUNK_SHORT = 4;

// scalar slots (doubles) of one CG instance
enum { S_RZ = 0, S_RZ_OLD, S_RR, S_RR0, S_TOL2, S_DIV2, S_ALPHA, S_BETA, S_UPD2, S_NS };
// integer slots; I_TOK: one per workgroup of k_cg_pupdate, the iteration it
// last processed (see there)
enum { I_DONE = 0, I_ITER, I_NEXT, I_MAXIT, I_TOK, I_NI = I_TOK + RG };
// I_DONE values: 0 running, 1 converged, 2 iteration limit, 3 diverged /
// breakdown, 4 residual reduced by the reliable-update factor (FP32 solve)
// partial-sum slots (each RG doubles)
enum { P_PAP = 0, P_RR, P_RZ, P_RZO, P_NP };

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the workgroup, returned to every thread (fixed tree).
__device__ __forceinline__ double block_sum_all(double v) {
  __shared__ double sm[RB / 64];
  __shared__ double res;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sm[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < RB / 64; ++i) s += sm[i];
    res = s;
  }
  __syncthreads();
  return res;
}

// Sum over the workgroup of N values at once (fixed order), to every thread.
template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N]) {
  __shared__ double sm[N][RB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) sm[k][w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < RB / 64; ++i) s += sm[k][i];
    v[k] = s;
  }
}

// Write this workgroup's partial sum of v.
__device__ __forceinline__ void put_partial(double *slot, double v) {
  v = block_sum_all(v);
  if (threadIdx.x == 0) slot[blockIdx.x] = v;
}

// Fold nb partials written by an earlier kernel (same value in every
// workgroup and thread).
__device__ __forceinline__ double fold(const double *slot, int nb) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += RB) s += slot[i];
  return block_sum_all(s);
}

// This thread's share of the sum of n partials, eight loads in flight at a
// time (a serial chain of dependent loads would expose the L2 latency once
// per partial); fixed order.
template <int NL = 8>
__device__ __forceinline__ double fold_share(const double *__restrict__ a, int n) {
  static_assert(NL == 8 || NL == 16, "8 or 16 loads in flight");
  double acc[NL];
#pragma unroll
  for (int k = 0; k < NL; ++k) acc[k] = 0.0;
  int i = threadIdx.x;
  for (; i + (NL - 1) * RB < n; i += NL * RB) {
#pragma unroll
    for (int k = 0; k < NL; ++k) acc[k] += a[i + k * RB];
  }
  for (; i < n; i += RB) acc[0] += a[i];
  if constexpr (NL == 16) {
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] += acc[k + 8];
  }
  return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

// The same with 16-byte loads of consecutive pairs, NL in flight: the
// 8,192 per-tile partials of the FP32 split-element apply at 1024^2 take
// one round of loads per thread instead of two (each round waits for the
// last-level cache, where the apply left them); fixed order.  a: 16-byte
// aligned.
template <int NL>
__device__ __forceinline__ double fold_pairs(const double *__restrict__ a, int n) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 *a2 = reinterpret_cast<const d2 *>(a);
  const int m = n >> 1;
  double acc[NL];
#pragma unroll
  for (int k = 0; k < NL; ++k) acc[k] = 0.0;
  int i = threadIdx.x;
  for (; i + (NL - 1) * RB < m; i += NL * RB) {
    d2 v[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) v[k] = a2[i + k * RB];
#pragma unroll
    for (int k = 0; k < NL; ++k) acc[k] += v[k].x + v[k].y;
  }
  for (; i < m; i += RB) {
    const d2 v = a2[i];
    acc[0] += v.x + v.y;
  }
  if ((n & 1) && threadIdx.x == 0) acc[0] += a[n - 1];
#pragma unroll
  for (int w = NL / 2; w > 0; w >>= 1)
#pragma unroll
    for (int k = 0; k < w; ++k) acc[k] += acc[k + w];
  return acc[0];
}

__device__ __forceinline__ bool finite(double x) { return x == x && !isinf(x); }

#define GRID_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < (n); i += (int64_t)gridDim.x * RB)

// ---------------------------------------------------------------- CG kernels
// init: u = 0, r = f*fscale, z = dinv r (or r), p = z when use_z;
// partials r.r and r.z
// TU: type of the iterate u (FP64 also for the FP32 inner recursion: an FP32
// iterate's rounding returns through the true residual of every reliable
// update and delays convergence)
template <typename V, typename F, typename D, typename TU = V>
__global__ void k_cg_init(const F *__restrict__ f, double fscale, TU *__restrict__ u,
                          V *__restrict__ r, V *__restrict__ z, V *__restrict__ p,
                          const D *__restrict__ dinv, int use_z, int64_t n, double *part) {
  double rr = 0.0, rz = 0.0;
  GRID_LOOP(i, n) {
    const V ri = (V)((double)f[i] * fscale);
    u[i] = TU(0);
    r[i] = ri;
    const V zi = dinv ? (V)((double)ri * (double)dinv[i]) : ri;
    if (use_z) {
      z[i] = zi;
      p[i] = zi;
    }
    rr += (double)ri * (double)ri;
    rz += (double)ri * (double)zi;
  }
  put_partial(part + P_RR * RG, rr);
  put_partial(part + P_RZ * RG, rz);
}

// one workgroup: tolerances and the first r.z
__global__ void k_cg_init_fin(const double *part, int nb, double *sc, int *ic, double rtol2,
                              double abstol2, double divtol2, int maxit, double upd2) {
  const double rr = fold(part + P_RR * RG, nb);
  const double rz = fold(part + P_RZ * RG, nb);
  if (threadIdx.x == 0) {
    sc[S_RR0] = rr;
    sc[S_RR] = rr;
    sc[S_RZ] = rz;
    sc[S_TOL2] = fmax(rtol2 * rr, abstol2);
    sc[S_DIV2] = divtol2 * rr;
    sc[S_UPD2] = upd2 * rr;
    ic[I_ITER] = 0;
    ic[I_NEXT] = 0;
    ic[I_MAXIT] = maxit;
    ic[I_DONE] = !finite(rr) ? 3 : (rr <= sc[S_TOL2] ? 1 : 0);
  }
  for (int i = threadIdx.x; i < RG; i += blockDim.x) ic[I_TOK + i] = 0;
}

// K1: partials of p.Ap
// Column j = blockIdx.y of a block solve: vectors at j * ld, scalars, flags
// and partial slots of instance j (ld = 0 and one column otherwise).
#define CG_COLUMN(ld)                                 \
  const int col_ = blockIdx.y;                        \
  const int64_t voff_ = (int64_t)col_ * (ld);         \
  part += (size_t)col_ * P_NP * RG;                   \
  sc += (size_t)col_ * S_NS;                          \
  ic += (size_t)col_ * I_NI;

template <typename V>
__global__ void k_cg_pap(const V *__restrict__ p, const V *__restrict__ Ap, int64_t n,
                         double *part, double *sc, int *ic, const int *outer_done, int64_t ld) {
  CG_COLUMN(ld)
  p += voff_;
  Ap += voff_;
  if (ic[I_DONE] || (outer_done && *outer_done)) return;
  constexpr int U = 4;
  double s = 0.0;
  for (int64_t b0 = (int64_t)blockIdx.x * RB * U + threadIdx.x; b0 < n;
       b0 += (int64_t)gridDim.x * RB * U) {
    V pv[U], av[U];
    if (b0 + (int64_t)(U - 1) * RB < n) {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        pv[k] = p[b0 + (int64_t)k * RB];
        av[k] = Ap[b0 + (int64_t)k * RB];
      }
    } else {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int64_t i = b0 + (int64_t)k * RB;
        pv[k] = i < n ? p[i] : V(0);
        av[k] = i < n ? Ap[i] : V(0);
      }
    }
#pragma unroll
    for (int k = 0; k < U; ++k) s += (double)pv[k] * (double)av[k];
  }
  put_partial(part + P_PAP * RG, s);
}

// K2: alpha; u += alpha p; r -= alpha Ap (rold keeps r for the flexible
// beta); z = dinv r when the preconditioner is local; partials of r.r (or
// z.z with prec_norm) and r.z
// SEAM: Ap comes from a gather apply that left its seam partials in the
// function slots sg describes (igai::CgHook::defer_seam); this kernel sums them.
template <typename V, typename D, bool LOCAL, bool ROLD, typename TU = V, bool SEAM = false>
__global__ __launch_bounds__(RB) void k_cg_update(TU *__restrict__ u, V *__restrict__ r, const V *__restrict__ p,
                            const V *__restrict__ Ap, V *__restrict__ z, V *__restrict__ rold,
                            const D *__restrict__ dinv, int prec_norm, int64_t n, double *part,
                            double *sc, int *ic, int nb, const int *outer_done, int64_t ld,
                            const double *__restrict__ tdot, int ntiles, igai::SeamGeom sg) {
  CG_COLUMN(ld)
  u += voff_;
  r += voff_;
  p += voff_;
  Ap += voff_;
  z += voff_;
  if (ROLD) rold += voff_;
  // Four elements per thread and step, every load of a step issued before
  // any store; the first step's loads are issued before the done flags are
  // read and the p.Ap fold, and each next step's before the current one is
  // processed, so their latency overlaps the flags, the fold and the stores.
  constexpr int U = 4;
  const int64_t step = (int64_t)gridDim.x * RB * U;
  using SP = igai::SeamPart<V>;
  auto ldap = [&](int64_t i, SP &q) {
    if constexpr (SEAM)
      return igai::seam_load(Ap, sg, i, q);
    else
      return Ap[i];
  };
  auto load = [&](int64_t b, TU (&uu)[U], V (&pv)[U], V (&rv)[U], V (&av)[U], D (&dv)[U],
                  SP (&sp)[U]) {
    if (b + (int64_t)(U - 1) * RB < n) {  // whole step in range: unconditional loads
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int64_t i = b + (int64_t)k * RB;
        uu[k] = u[i];
        pv[k] = p[i];
        rv[k] = r[i];
        av[k] = ldap(i, sp[k]);
        if (LOCAL) dv[k] = dinv ? dinv[i] : D(1);
      }
    } else {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int64_t i = b + (int64_t)k * RB;
        const bool in = i < n;
        sp[k].a1 = sp[k].a2 = sp[k].a3 = V(0);
        uu[k] = in ? u[i] : TU(0);
        pv[k] = in ? p[i] : V(0);
        rv[k] = in ? r[i] : V(0);
        av[k] = in ? ldap(i, sp[k]) : V(0);
        if (LOCAL) dv[k] = (dinv && in) ? dinv[i] : D(1);
      }
    }
  };
  int64_t b0 = (int64_t)blockIdx.x * RB * U + threadIdx.x;
  TU uu[U];
  V pv[U], rv[U], av[U];
  D dv[U];
  SP sp[U];
  if (b0 < n) load(b0, uu, pv, rv, av, dv, sp);
  const bool stop = ic[I_DONE] || (outer_done && *outer_done);
  const double rz_old = sc[S_RZ];
  // p.Ap: partials of k_cg_pap, or the apply kernel's per-tile partials
  // (FP32: the split-element apply's 8,192 tile partials at 1024^2, sixteen
  // 16-byte loads in flight), summed over the workgroup together with the
  // stop flag, so that every thread returns or none
  double red[2];
  if (tdot) {
    red[0] = sizeof(V) == 4 ? fold_pairs<16>(tdot, ntiles) : fold_share<8>(tdot, ntiles);
  } else {
    double sp0 = 0.0;
    for (int i = threadIdx.x; i < nb; i += RB) sp0 += part[P_PAP * RG + i];
    red[0] = sp0;
  }
  red[1] = stop ? 1.0 : 0.0;
  block_sum_n<2>(red);
  if (red[1] != 0.0) return;
  const double pap = red[0];
  const double a = rz_old / pap;
  if (!finite(a) || pap == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ic[I_DONE] = 3;
    return;
  }
  double rr = 0.0, rz = 0.0;
  while (b0 < n) {
    const int64_t bn = b0 + step;
    TU uu2[U];
    V pv2[U], rv2[U], av2[U];
    D dv2[U];
    SP sp2[U];
    if (bn < n) load(bn, uu2, pv2, rv2, av2, dv2, sp2);
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t i = b0 + (int64_t)k * RB;
      if (i >= n) continue;
      const V api = SEAM ? igai::seam_value(av[k], sp[k]) : av[k];
      u[i] = (TU)((double)uu[k] + a * (double)pv[k]);
      if (ROLD) rold[i] = rv[k];
      const V ri = (V)((double)rv[k] - a * (double)api);
      r[i] = ri;
      if (LOCAL) {
        const V zi = (V)((double)ri * (double)dv[k]);
        z[i] = zi;
        rz += (double)ri * (double)zi;
        rr += prec_norm ? (double)zi * (double)zi : (double)ri * (double)ri;
      } else {
        rr += (double)ri * (double)ri;
      }
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      uu[k] = uu2[k];
      pv[k] = pv2[k];
      rv[k] = rv2[k];
      av[k] = av2[k];
      dv[k] = dv2[k];
      sp[k] = sp2[k];
    }
    b0 = bn;
  }
  if constexpr (LOCAL) {
    // both partials through one set of barriers (the sums of put_partial)
    double v[2] = {rr, rz};
    block_sum_n<2>(v);
    if (threadIdx.x == 0) {
      part[P_RR * RG + blockIdx.x] = v[0];
      part[P_RZ * RG + blockIdx.x] = v[1];
    }
  } else {
    put_partial(part + P_RR * RG, rr);
  }
  // bookkeeping for K3 (slots no workgroup of this kernel reads)
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[S_ALPHA] = a;
    sc[S_RZ_OLD] = rz_old;
    ic[I_NEXT] = ic[I_ITER] + 1;
  }
}

// after a preconditioner applied by its own kernels: partials of r.z and
// (flexible) rold.z
template <typename V>
__global__ void k_cg_rz(const V *r, const V *z, const V *rold, int64_t n, double *part,
                        const int *ic) {
  if (ic[I_DONE]) return;
  double rz = 0.0, ro = 0.0;
  GRID_LOOP(i, n) {
    rz += (double)r[i] * (double)z[i];
    if (rold) ro += (double)rold[i] * (double)z[i];
  }
  put_partial(part + P_RZ * RG, rz);
  if (rold) put_partial(part + P_RZO * RG, ro);
}

// K3: convergence test, beta, p = z + beta p
// With Apz, also zero Ap for the next scatter-add apply (IEN atomic form),
// which then needs no seam-zeroing kernel; Ap is dead once this kernel runs.
// When the iteration stops, Ap is left as it is: cg_run zeroes it on entry.
// Workgroup 0 writes the done flag, and on a reliable-update stop (st = 4)
// the other workgroups must still form p: so this kernel never reads the
// done flag.  It runs for iteration I_NEXT when K2 ran (K2 advances I_NEXT
// and returns at once on a set flag), which each workgroup tells from its
// own token, the iteration it last processed (no other workgroup reads or
// writes it).
template <typename V>
__global__ void k_cg_pupdate(V *__restrict__ p, const V *__restrict__ z, int64_t n,
                             const double *part, double *sc, int *ic, int nb, int flexible,
                             const int *outer_done, int64_t ld, V *__restrict__ Apz) {
  CG_COLUMN(ld)
  p += voff_;
  z += voff_;
  if (Apz) Apz += voff_;
  constexpr int U = 4;
  const int64_t step = (int64_t)gridDim.x * RB * U;
  auto load = [&](int64_t bq, V (&zv)[U], V (&pv)[U]) {
    if (bq + (int64_t)(U - 1) * RB < n) {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        zv[k] = z[bq + (int64_t)k * RB];
        pv[k] = p[bq + (int64_t)k * RB];
      }
    } else {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int64_t i = bq + (int64_t)k * RB;
        zv[k] = i < n ? z[i] : V(0);
        pv[k] = i < n ? p[i] : V(0);
      }
    }
  };
  // first step's loads before the flags, scalars and folds, each next
  // step's before the current one is processed (latency overlap, as in
  // k_cg_update); the scalars the test below reads are loaded with the flags
  int64_t b0 = (int64_t)blockIdx.x * RB * U + threadIdx.x;
  V zv[U], pv[U];
  if (b0 < n) load(b0, zv, pv);
  if (outer_done && *outer_done) return;
  const int it = ic[I_NEXT];
  if (ic[I_TOK + blockIdx.x] == it) return;
  const double tol2 = sc[S_TOL2], div2 = sc[S_DIV2], upd2 = sc[S_UPD2], rzo = sc[S_RZ_OLD];
  const int maxit = ic[I_MAXIT];
  double f[3];
  {
    // partials in 16-byte pairs: the update kernel's 512 in one round of
    // loads per thread (the slots are RG-aligned)
    typedef double d2 __attribute__((ext_vector_type(2)));
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const int np = nb >> 1;
    for (int i = threadIdx.x; i < np; i += RB) {
      const d2 v0 = reinterpret_cast<const d2 *>(part + P_RR * RG)[i];
      const d2 v1 = reinterpret_cast<const d2 *>(part + P_RZ * RG)[i];
      a0 += v0.x + v0.y;
      a1 += v1.x + v1.y;
      if (flexible) {
        const d2 v2 = reinterpret_cast<const d2 *>(part + P_RZO * RG)[i];
        a2 += v2.x + v2.y;
      }
    }
    if ((nb & 1) && threadIdx.x == 0) {
      a0 += part[P_RR * RG + nb - 1];
      a1 += part[P_RZ * RG + nb - 1];
      if (flexible) a2 += part[P_RZO * RG + nb - 1];
    }
    f[0] = a0;
    f[1] = a1;
    f[2] = a2;
  }
  block_sum_n<3>(f);  // its barriers order every thread's token read before the write
  if (threadIdx.x == 0) ic[I_TOK + blockIdx.x] = it;
  const double rr = f[0], rz = f[1], ro = f[2];
  int st = 0;
  if (rr <= tol2) st = 1;
  else if (!finite(rr) || rr > div2) st = 3;
  else if (it >= maxit) st = 2;
  else if (rr <= upd2) st = 4;
  const double b = (rz - ro) / rzo;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[S_RR] = rr;
    sc[S_RZ] = rz;
    sc[S_BETA] = b;
    ic[I_ITER] = it;
    if (st) ic[I_DONE] = st;
  }
  if (st && st != 4) return;  // a reliable update keeps the search direction
  while (b0 < n) {
    const int64_t bn = b0 + step;
    V zv2[U], pv2[U];
    if (bn < n) load(bn, zv2, pv2);
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t i = b0 + (int64_t)k * RB;
      if (i < n) {
        p[i] = (V)((double)zv[k] + b * (double)pv[k]);
        if (Apz) Apz[i] = V(0);
      }
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      zv[k] = zv2[k];
      pv[k] = pv2[k];
    }
    b0 = bn;
  }
}

// first preconditioner application of a non-local preconditioner:
// partials of r.z and p = z
template <typename V>
__global__ void k_cg_rz0(const V *r, const V *z, V *p, int64_t n, double *part, const int *ic) {
  if (ic[I_DONE]) return;
  double s = 0.0;
  GRID_LOOP(i, n) {
    s += (double)r[i] * (double)z[i];
    p[i] = z[i];
  }
  put_partial(part + P_RZ * RG, s);
}

__global__ void k_cg_rz0_fin(const double *part, int nb, double *sc, const int *ic) {
  if (ic[I_DONE]) return;
  const double rz = fold(part + P_RZ * RG, nb);
  if (threadIdx.x == 0) sc[S_RZ] = rz;
}

// ---------------------------------------------------------------- reliable update
// x64 += x, x = 0
template <typename TU>
__global__ void k_ru_accum(double *x64, TU *x, int64_t n) {
  GRID_LOOP(i, n) {
    x64[i] += (double)x[i];
    x[i] = TU(0);
  }
}

// r32 = r64 (true residual); z32 = dinv r32 when local; partials r.z
__global__ void k_ru_reset(const double *__restrict__ r64, float *__restrict__ r32,
                           float *__restrict__ z32, const float *__restrict__ dinv, int local_pc,
                           int64_t n, double *part) {
  double rz = 0.0;
  GRID_LOOP(i, n) {
    const float ri = (float)r64[i];
    r32[i] = ri;
    if (local_pc) {
      const float zi = dinv ? (float)((double)ri * (double)dinv[i]) : ri;
      z32[i] = zi;
      rz += (double)ri * (double)zi;
    }
  }
  if (local_pc) put_partial(part + P_RZ * RG, rz);
}

__global__ void k_ru_clear(int *ic) {
  if (threadIdx.x == 0) ic[I_DONE] = 0;
}

// partials of r.z after a non-local preconditioner
__global__ void k_ru_rz(const float *r, const float *z, int64_t n, double *part) {
  double rz = 0.0;
  GRID_LOOP(i, n) rz += (double)r[i] * (double)z[i];
  put_partial(part + P_RZ * RG, rz);
}

// restart the recursion from the true residual: rz, update threshold, done = 0
__global__ void k_ru_fin(const double *part, int nb, double *sc, int *ic, double upd2) {
  const double rz = fold(part + P_RZ * RG, nb);
  if (threadIdx.x == 0) {
    sc[S_RZ] = rz;
    sc[S_UPD2] = upd2;
    ic[I_DONE] = 0;
  }
}

// ---------------------------------------------------------------- fused iteration
// Single-reduction PCG for local preconditioners (Chronopoulos & Gear, 1989),
// with p.Ap formed explicitly from products that need no second reduction:
//   w = A u                      (apply; u.w = sum of its per-tile partials)
//   beta = g / g_old,  g = r.u
//   p.Ap = u.w + 2 beta u.s_old + beta^2 (p.Ap)_old     (A symmetric, s = A p)
//   alpha = g / p.Ap
//   p = u + beta p;  s = w + beta s;  x += alpha p;  r -= alpha s;  u = M r
// and the partials of the next r.u, r.r and u.s in the same pass.  An
// iteration is the apply plus one vector kernel.  The update kernel reads the
// state it also rewrites, so the state is double-buffered: the kernel reads
// CUR and writes NEW; the apply kernel between two updates copies NEW to CUR
// (its workgroup 0, a kernel no update kernel overlaps).
// State layout (doubles), CUR at 0 and NEW at CG_L:
enum { CG_RZO = 0, CG_PAP, CG_IT, CG_RR = 4, CG_RZ = CG_RR + RG, CG_US = CG_RZ + RG,
       CG_L = CG_US + RG };

// init: x = 0, r = f*fscale, u = M r, p = s = 0; NEW partials of r.r, r.u
// and u.s (= 0)
template <typename V, typename F, typename D, typename TU = V>
__global__ void k_cgf_init(const F *__restrict__ f, double fscale, TU *__restrict__ x,
                           V *__restrict__ r, V *__restrict__ u, V *__restrict__ p,
                           V *__restrict__ sv, const D *__restrict__ dinv, int64_t n,
                           double *cgs) {
  double v[3] = {0.0, 0.0, 0.0};
  GRID_LOOP(i, n) {
    const V ri = (V)((double)f[i] * fscale);
    const V ui = dinv ? (V)((double)ri * (double)dinv[i]) : ri;
    x[i] = TU(0);
    r[i] = ri;
    u[i] = ui;
    p[i] = V(0);
    sv[i] = V(0);
    v[0] += (double)ri * (double)ri;
    v[1] += (double)ri * (double)ui;
  }
  block_sum_n<3>(v);
  if (threadIdx.x == 0) {
    double *nw = cgs + CG_L;
    nw[CG_RR + blockIdx.x] = v[0];
    nw[CG_RZ + blockIdx.x] = v[1];
    nw[CG_US + blockIdx.x] = v[2];
  }
}

// one workgroup: tolerances (as k_cg_init_fin) and the NEW scalars of a
// fresh recursion (g_old = inf makes the first beta 0)
__global__ void k_cgf_init_fin(double *cgs, int nb, double *sc, int *ic, double rtol2,
                               double abstol2, double divtol2, int maxit, double upd2) {
  double *nw = cgs + CG_L;
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nb; i += RB) {
    v[0] += nw[CG_RR + i];
    v[1] += nw[CG_RZ + i];
  }
  block_sum_n<2>(v);
  if (threadIdx.x == 0) {
    const double rr = v[0];
    sc[S_RR0] = rr;
    sc[S_RR] = rr;
    sc[S_RZ] = v[1];
    sc[S_TOL2] = fmax(rtol2 * rr, abstol2);
    sc[S_DIV2] = divtol2 * rr;
    sc[S_UPD2] = upd2 * rr;
    nw[CG_RZO] = HUGE_VAL;
    nw[CG_PAP] = 0.0;
    nw[CG_IT] = 0.0;
    ic[I_ITER] = 0;
    ic[I_NEXT] = 0;
    ic[I_MAXIT] = maxit;
    ic[I_DONE] = !finite(rr) ? 3 : (rr <= sc[S_TOL2] ? 1 : 0);
  }
}

// The vector kernel of a fused iteration (see above).  w = A u comes from the
// apply, the partials of u.w from its tiles (tdot).  Stops (done flag) on
// convergence, the iteration cap, divergence or breakdown, and -- FP32 inner
// solve -- when r.r fell below the reliable-update threshold; the state is
// then that of the last completed iteration.  zero_w: clear w for the next
// scatter-add apply.
// SEAM: as in k_cg_update, w's seam functions are summed here.
template <typename V, typename D, bool JAC, typename TU = V, bool SEAM = false>
__global__ __launch_bounds__(RB) void k_cgf_update(TU *__restrict__ x, V *__restrict__ r, V *__restrict__ p,
                             V *__restrict__ sv, V *__restrict__ u, V *__restrict__ w,
                             const D *__restrict__ dinv, int zero_w, int64_t n, double *cgs,
                             const double *__restrict__ tdot, int ntiles, double *sc, int *ic,
                             int nb, igai::SeamGeom sg) {
  if (ic[I_DONE]) return;
  const double *cur = cgs;
  double *nw = cgs + CG_L;
  constexpr int U = 4;
  const int64_t step = (int64_t)gridDim.x * RB * U;
  struct Step {
    V u[U], w[U], p[U], s[U], r[U];
    TU x[U];
    D d[U];
    igai::SeamPart<V> sp[U];
  };
  auto ldw = [&](int64_t i, igai::SeamPart<V> &q) {
    if constexpr (SEAM)
      return igai::seam_load(w, sg, i, q);
    else
      return w[i];
  };
  auto load = [&](int64_t b, Step &q) {
    if (b + (int64_t)(U - 1) * RB < n) {  // whole step in range: unconditional loads
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int64_t i = b + (int64_t)k * RB;
        q.u[k] = u[i];
        q.w[k] = ldw(i, q.sp[k]);
        q.p[k] = p[i];
        q.s[k] = sv[i];
        q.x[k] = x[i];
        q.r[k] = r[i];
        q.d[k] = JAC ? dinv[i] : D(1);
      }
    } else {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int64_t i = b + (int64_t)k * RB;
        const bool in = i < n;
        q.u[k] = in ? u[i] : V(0);
        q.sp[k].a1 = q.sp[k].a2 = q.sp[k].a3 = V(0);
        q.w[k] = in ? ldw(i, q.sp[k]) : V(0);
        q.p[k] = in ? p[i] : V(0);
        q.s[k] = in ? sv[i] : V(0);
        q.x[k] = in ? x[i] : TU(0);
        q.r[k] = in ? r[i] : V(0);
        q.d[k] = (JAC && in) ? dinv[i] : D(1);
      }
    }
  };
  // first step's loads before the folds, each next step's before the current
  // one is processed
  int64_t b0 = (int64_t)blockIdx.x * RB * U + threadIdx.x;
  Step q;
  if (b0 < n) load(b0, q);
  double f[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < nb; i += RB) {
    f[0] += cur[CG_RR + i];
    f[1] += cur[CG_RZ + i];
    f[2] += cur[CG_US + i];
  }
  f[3] = fold_share(tdot, ntiles);
  block_sum_n<4>(f);
  const double rr = f[0], g = f[1], us = f[2], uw = f[3];
  const int it = (int)cur[CG_IT];
  int st = 0;
  if (rr <= sc[S_TOL2]) st = 1;
  else if (!finite(rr) || rr > sc[S_DIV2]) st = 3;
  else if (it >= ic[I_MAXIT]) st = 2;
  else if (rr <= sc[S_UPD2]) st = 4;
  const double b = g / cur[CG_RZO];
  const double pap = uw + 2.0 * b * us + b * b * cur[CG_PAP];
  const double a = g / pap;
  if (!st && (!finite(a) || pap == 0.0)) st = 3;
  if (st) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc[S_RR] = rr;
      ic[I_ITER] = it;
      ic[I_DONE] = st;
    }
    return;
  }
  double v[3] = {0.0, 0.0, 0.0};
  while (b0 < n) {
    const int64_t bn = b0 + step;
    Step q2;
    if (bn < n) load(bn, q2);
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t i = b0 + (int64_t)k * RB;
      if (i >= n) continue;
      const V pn = (V)((double)q.u[k] + b * (double)q.p[k]);
      const V wk = SEAM ? igai::seam_value(q.w[k], q.sp[k]) : q.w[k];
      const V sn = (V)((double)wk + b * (double)q.s[k]);
      const V rn = (V)((double)q.r[k] - a * (double)sn);
      const V un = JAC ? (V)((double)rn * (double)q.d[k]) : rn;
      p[i] = pn;
      sv[i] = sn;
      x[i] = (TU)((double)q.x[k] + a * (double)pn);
      r[i] = rn;
      u[i] = un;
      if (zero_w) w[i] = V(0);
      v[0] += (double)rn * (double)rn;
      v[1] += (double)rn * (double)un;
      v[2] += (double)un * (double)sn;
    }
    q = q2;
    b0 = bn;
  }
  block_sum_n<3>(v);
  if (threadIdx.x == 0) {
    nw[CG_RR + blockIdx.x] = v[0];
    nw[CG_RZ + blockIdx.x] = v[1];
    nw[CG_US + blockIdx.x] = v[2];
    if (blockIdx.x == 0) {
      nw[CG_RZO] = g;
      nw[CG_PAP] = pap;
      nw[CG_IT] = (double)(it + 1);
      sc[S_RR] = rr;
      sc[S_RZ] = g;
      sc[S_ALPHA] = a;
      sc[S_BETA] = b;
      ic[I_ITER] = it + 1;
    }
  }
}

// reliable update of the fused FP32 recursion: r32 = r64 (true residual),
// u32 = M r32, NEW partials of r.r, r.u and u.s (p, s and the NEW scalars of
// the last completed iteration are kept)
__global__ void k_cgf_ru_reset(const double *__restrict__ r64, float *__restrict__ r32,
                               float *__restrict__ u32, const float *__restrict__ s32,
                               const float *__restrict__ dinv, int64_t n, double *cgs) {
  double v[3] = {0.0, 0.0, 0.0};
  GRID_LOOP(i, n) {
    const float ri = (float)r64[i];
    const float ui = dinv ? (float)((double)ri * (double)dinv[i]) : ri;
    r32[i] = ri;
    u32[i] = ui;
    v[0] += (double)ri * (double)ri;
    v[1] += (double)ri * (double)ui;
    v[2] += (double)ui * (double)s32[i];
  }
  block_sum_n<3>(v);
  if (threadIdx.x == 0) {
    double *nw = cgs + CG_L;
    nw[CG_RR + blockIdx.x] = v[0];
    nw[CG_RZ + blockIdx.x] = v[1];
    nw[CG_US + blockIdx.x] = v[2];
  }
}

__global__ void k_cgf_ru_fin(double *sc, int *ic, double upd2) {
  if (threadIdx.x == 0) {
    sc[S_UPD2] = upd2;
    ic[I_DONE] = 0;
  }
}

// ---------------------------------------------------------------- outer (FP64)
// r = f - Au; partials |r|^2 and |f|^2
__global__ void k_residual(const double *f, const double *Au, double *r, int64_t n,
                           double *part) {
  double rr = 0.0, ff = 0.0;
  GRID_LOOP(i, n) {
    const double ri = f[i] - Au[i];
    r[i] = ri;
    rr += ri * ri;
    ff += f[i] * f[i];
  }
  put_partial(part + P_RR * RG, rr);
  put_partial(part + P_RZ * RG, ff);
}

__global__ void k_fold2(const double *part, int nb, double *out) {
  const double a = fold(part + P_RR * RG, nb);
  const double b = fold(part + P_RZ * RG, nb);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

__global__ void k_inv(double *d, int64_t n) {
  GRID_LOOP(i, n) d[i] = 1.0 / d[i];
}

__global__ void k_dot(const double *a, const double *b, int64_t n, double *part) {
  double s = 0.0;
  GRID_LOOP(i, n) s += a[i] * b[i];
  put_partial(part + P_RR * RG, s);
}

__global__ void k_fold1(const double *part, int nb, double *out) {
  const double a = fold(part + P_RR * RG, nb);
  if (threadIdx.x == 0) out[0] = a;
}

// ---------------------------------------------------------------- FEM preconditioner
// CSR SpMV y = K x (one thread per row; the Q1 rows hold at most 9 entries)
template <typename V>
__global__ void k_spmv(const int *rowptr, const int *col, const V *val, const V *x, V *y,
                       int nrows, const int *skip, const int *skip2) {
  if ((skip && *skip) || (skip2 && *skip2)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  double s = 0.0;
  for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += (double)val[k] * (double)x[col[k]];
  y[i] = (V)s;
}

// z = (L U)^-1 r for the ILU(0) factors in CSR (L unit lower, U upper with
// its diagonal; the factors share the matrix's pattern), one workgroup:
// rows of a level depend only on rows of earlier levels, a barrier between
// levels.  Forward levels fptr/frows, backward bptr/brows.
template <typename V>
__global__ __launch_bounds__(1024) void k_ilu_solve(
    const int *__restrict__ rp, const int *__restrict__ ci, const V *__restrict__ lu,
    const int *__restrict__ diag, const int *__restrict__ fptr, const int *__restrict__ frows,
    int nfl, const int *__restrict__ bptr, const int *__restrict__ brows, int nbl,
    const V *__restrict__ r, V *z, const int *skip) {
  if (skip && *skip) return;
  for (int l = 0; l < nfl; ++l) {
    for (int k = fptr[l] + (int)threadIdx.x; k < fptr[l + 1]; k += (int)blockDim.x) {
      const int i = frows[k];
      double s = (double)r[i];
      for (int e = rp[i]; e < diag[i]; ++e) s -= (double)lu[e] * (double)z[ci[e]];
      z[i] = (V)s;
    }
    __syncthreads();
  }
  for (int l = 0; l < nbl; ++l) {
    for (int k = bptr[l] + (int)threadIdx.x; k < bptr[l + 1]; k += (int)blockDim.x) {
      const int i = brows[k];
      double s = (double)z[i];
      for (int e = diag[i] + 1; e < rp[i + 1]; ++e) s -= (double)lu[e] * (double)z[ci[e]];
      z[i] = (V)(s / (double)lu[diag[i]]);
    }
    __syncthreads();
  }
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

// Chebyshev-Jacobi preconditioner z = p_m(D^-1 K) D^-1 r on the Q1 matrix
// (Chebyshev semi-iteration on [lmin, lmax] from z = 0; the coefficients
// are host constants, so the preconditioner is linear and capturable).
template <typename V>
__global__ void k_cheb_init(const V *__restrict__ r, V *__restrict__ z, V *__restrict__ d,
                            const double *__restrict__ dinv, double c, int64_t n,
                            const int *skip) {
  if (skip && *skip) return;
  GRID_LOOP(i, n) {
    const V v = (V)(c * (double)r[i] * dinv[i]);
    d[i] = v;
    z[i] = v;
  }
}

template <typename V>
__global__ void k_cheb_step(const V *__restrict__ r, const V *__restrict__ q, V *__restrict__ z,
                            V *__restrict__ d, const double *__restrict__ dinv, double c1,
                            double c2, int64_t n, const int *skip) {
  if (skip && *skip) return;
  GRID_LOOP(i, n) {
    const double res = dinv[i] * ((double)r[i] - (double)q[i]);
    const V dn = (V)(c1 * (double)d[i] + c2 * res);
    d[i] = dn;
    z[i] = (V)((double)z[i] + (double)dn);
  }
}

// per-workgroup maximum over rows of sum_j |K_ij| / K_ii (Gershgorin bound
// of D^-1 K)
__global__ void k_gersh(const int *rowptr, const int *col, const double *val, int nrows,
                        double *part) {
  double m = 0.0;
  for (int i = blockIdx.x * RB + threadIdx.x; i < nrows; i += gridDim.x * RB) {
    double s = 0.0, dg = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      s += fabs(val[k]);
      if (col[k] == i) dg = fabs(val[k]);
    }
    if (dg > 0.0) m = fmax(m, s / dg);
  }
  __shared__ double sm[RB];
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int w = RB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

template <typename V>
__global__ void k_cast(const double *a, V *b, int64_t n) {
  GRID_LOOP(i, n) b[i] = (V)a[i];
}

// inner CG on the FEM matrix solving K z = r (z0 = 0, Jacobi, PETSc's default
// preconditioned-norm test |B r_k| <= max(rtol |B r_0|, abstol))
template <typename V>
__global__ void k_pc_init(const V *r, V *zz, V *rin, V *w, V *pp, const double *dinv, int64_t n,
                          double *part, const int *outer_done) {
  if (*outer_done) return;
  double rz = 0.0, ww = 0.0;
  GRID_LOOP(i, n) {
    const double ri = (double)r[i];
    zz[i] = V(0);
    rin[i] = r[i];
    const V wi = (V)(ri * dinv[i]);
    w[i] = wi;
    pp[i] = wi;
    rz += ri * (double)wi;
    ww += (double)wi * (double)wi;
  }
  put_partial(part + P_RR * RG, ww);
  put_partial(part + P_RZ * RG, rz);
}

__global__ void k_pc_init_fin(const double *part, int nb, double *sc, int *ic, double rtol2,
                              double atol2, int maxit, const int *outer_done) {
  if (*outer_done) return;
  const double ww = fold(part + P_RR * RG, nb);
  const double rz = fold(part + P_RZ * RG, nb);
  if (threadIdx.x == 0) {
    sc[S_RZ] = rz;
    sc[S_RR0] = ww;
    sc[S_TOL2] = fmax(rtol2 * ww, atol2);
    sc[S_DIV2] = 1e300;
    ic[I_ITER] = 0;
    ic[I_NEXT] = 0;
    ic[I_MAXIT] = maxit;
    ic[I_DONE] = (ww <= sc[S_TOL2]) ? 1 : 0;
  }
  for (int i = threadIdx.x; i < RG; i += blockDim.x) ic[I_TOK + i] = 0;
}

inline int nblocks(int64_t n) {
  int64_t b = (n + RB - 1) / RB;
  if (b > RG) b = RG;
  if (b < 1) b = 1;
  return (int)b;
}

// ILU(0) in natural ordering (PETSc PCILU, levels 0): for each row i and each
// k < i in its pattern, l_ik = a_ik / u_kk, then a_ij -= l_ik u_kj for the
// j > k of row i that row k also holds.  Columns sorted per row.  Returns the
// position of each diagonal in diag; false on a missing or zero pivot.
bool ilu0_factor(int n, const std::vector<int> &rp, const std::vector<int> &ci,
                 std::vector<double> &a, std::vector<int> &diag) {
  diag.assign(n, -1);
  for (int i = 0; i < n; ++i)
    for (int e = rp[i]; e < rp[i + 1]; ++e)
      if (ci[e] == i) diag[i] = e;
  for (int i = 0; i < n; ++i) {
    if (diag[i] < 0) return false;
    for (int e = rp[i]; e < diag[i]; ++e) {
      const int k = ci[e];
      const double ukk = a[diag[k]];
      if (ukk == 0.0) return false;
      const double l = a[e] / ukk;
      a[e] = l;
      // row i entries right of k against row k entries right of its diagonal
      int f = diag[k] + 1;
      for (int g = e + 1; g < rp[i + 1] && f < rp[k + 1];) {
        if (ci[g] == ci[f]) {
          a[g] -= l * a[f];
          ++g;
          ++f;
        } else if (ci[g] < ci[f]) {
          ++g;
        } else {
          ++f;
        }
      }
    }
    if (a[diag[i]] == 0.0) return false;
  }
  return true;
}

// level sets of the triangular solves: forward, a row after the rows left of
// its diagonal; backward, after the rows right of it
void ilu0_levels(int n, const std::vector<int> &rp, const std::vector<int> &ci,
                 const std::vector<int> &diag, bool forward, std::vector<int> &ptr,
                 std::vector<int> &rows) {
  std::vector<int> lev(n, 0);
  int nl = 0;
  if (forward) {
    for (int i = 0; i < n; ++i) {
      int l = 0;
      for (int e = rp[i]; e < diag[i]; ++e) l = std::max(l, lev[ci[e]] + 1);
      lev[i] = l;
      nl = std::max(nl, l + 1);
    }
  } else {
    for (int i = n - 1; i >= 0; --i) {
      int l = 0;
      for (int e = diag[i] + 1; e < rp[i + 1]; ++e) l = std::max(l, lev[ci[e]] + 1);
      lev[i] = l;
      nl = std::max(nl, l + 1);
    }
  }
  ptr.assign(nl + 1, 0);
  for (int i = 0; i < n; ++i) ++ptr[lev[i] + 1];
  for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
  rows.assign(n, 0);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
}

}  // namespace

struct iga_solver {
  iga_op *op = nullptr;
  iga_solver_opts o{};
  int64_t n = 0;
  int nb = 1;
  int device = 0;
  // FP64 state
  double *r = nullptr, *p = nullptr, *Ap = nullptr, *z = nullptr, *rold = nullptr;
  double *Au = nullptr, *dinv = nullptr;  // Jacobi of A (inverse diagonal)
  float *dinv32 = nullptr;                // the same, rounded, for the FP32 inner PCG
  double *x64 = nullptr;                  // FP64 PCG iterate (owned: graphs capture it)
  // FP32 inner state
  double *d64 = nullptr;  // inner iterate (FP64: see k_cg_init)
  float *r32 = nullptr, *p32 = nullptr, *Ap32 = nullptr, *z32 = nullptr, *rold32 = nullptr;
  // scalars: [0, S_NS) main CG, [S_NS, 2 S_NS) inner PC CG, then outer slots
  double *sc = nullptr;
  int *ic = nullptr;
  double *part = nullptr, *part_in = nullptr;
  // FEM preconditioner (CSR of the reference's Q1 matrix)
  int fem_rows = 0;
  const int *rowptr = nullptr, *col = nullptr;
  const double *val64 = nullptr;
  float *val32 = nullptr;
  double *fdinv = nullptr;
  double *rin = nullptr, *w = nullptr, *pp = nullptr, *q = nullptr;
  float *rin32 = nullptr, *w32 = nullptr, *pp32 = nullptr, *q32 = nullptr;
  double cheb_lmax = 0.0;  // Gershgorin bound of D^-1 K (FEM_CHEB)
  igai::MgHier *mg = nullptr;  // multigrid hierarchy of the Q1 matrix (FEM_MG)
  // ILU(0) factors of K (FEM_ILU): values in K's pattern, diagonal
  // positions, level sets of the two triangular solves
  double *ilu64 = nullptr;
  float *ilu32 = nullptr;
  int *ilu_diag = nullptr, *ilu_fptr = nullptr, *ilu_frows = nullptr, *ilu_bptr = nullptr,
      *ilu_brows = nullptr;
  int ilu_nfl = 0, ilu_nbl = 0;
  std::vector<void *> allocs;
  int gen = -1;  // operator generation the Jacobi diagonal was built for
  // graphs of one chunk of iterations (FP64 PCG, FP32 inner PCG)
  hipGraphExec_t g64[2] = {}, g32[2] = {};  // one per flag slot (run_chunks)
  // fused iteration: double-buffered recursion state, per-tile u.Au partials,
  // graphs of one chunk
  double *cgs = nullptr, *tdot = nullptr;
  int ntiles64 = 0, ntiles32 = 0;
  hipGraphExec_t gf64[2] = {}, gf32[2] = {};
  // block solve (k columns of nfunc): FP64 iterates, FP32 inner vectors,
  // per-column scalars / flags / partial slots, graph of one chunk
  int bk = 0;
  double *bx64 = nullptr;
  double *bd64 = nullptr;  // FP32 recursions' iterates, held in FP64 (see k_cg_init)
  float *br32 = nullptr, *bp32 = nullptr, *bAp32 = nullptr, *bz32 = nullptr;
  double *bsc = nullptr, *bpart = nullptr;
  int *bic = nullptr;
  std::vector<void *> ballocs;
  hipGraphExec_t gblk = nullptr;
  int gblk_k = 0, gblk_cfg = -1;
  // pinned host copies of the done / iteration flags after two chunks in
  // flight, and the events that mark them complete (run_chunks)
  int *hflag = nullptr;
  double *hsc = nullptr;  // [2][S_NS] the main CG's scalars after those chunks
  int *dflag = nullptr;   // the same two buffers' device addresses (k_chunk_flags)
  double *dsc = nullptr;
  hipEvent_t fev[2] = {nullptr, nullptr};
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

// operator precision of the FP64 iteration or of the FP32 inner iteration
int apply_dt(const iga_solver *s, bool f64) {
  return f64 ? IGA_F64 : (s->o.inner_f16_geometry ? IGA_F32_G16 : IGA_F32);
}

#define LAUNCH(kern, grid, ...)                                              \
  do {                                                                       \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(RB), 0, st, __VA_ARGS__);      \
  } while (0)

// one application of the FEM preconditioner z = K^-1 r (inner CG + Jacobi)
template <typename V>
int fem_pc(iga_solver *s, const V *r, V *z, V *rin, V *w, V *pp, V *q, const V *val,
           hipStream_t st) {
  double *sc = s->sc + S_NS;
  int *ic = s->ic + I_NI;
  const int *od = s->ic + I_DONE;
  const int64_t n = s->n;
  const int nb = s->nb;
  const double rt = s->o.pc_inner_rtol, at = s->o.pc_inner_abstol;
  LAUNCH((k_pc_init<V>), nb, r, z, rin, w, pp, s->fdinv, n, s->part_in, od);
  hipLaunchKernelGGL(k_pc_init_fin, dim3(1), dim3(RB), 0, st, s->part_in, nb, sc, ic, rt * rt,
                     at * at, s->o.pc_inner_maxit, od);
  const int bs = 256;
  for (int it = 0; it < s->o.pc_inner_maxit; ++it) {
    hipLaunchKernelGGL((k_spmv<V>), dim3((s->fem_rows + bs - 1) / bs), dim3(bs), 0, st, s->rowptr,
                       s->col, val, pp, q, s->fem_rows, ic + I_DONE, od);
    LAUNCH((k_cg_pap<V>), nb, pp, q, n, s->part_in, sc, ic, od, (int64_t)0);
    LAUNCH((k_cg_update<V, double, true, false>), nb, z, rin, pp, q, w, (V *)nullptr, s->fdinv,
           1, n, s->part_in, sc, ic, nb, od, (int64_t)0, (const double *)nullptr, 0,
           igai::SeamGeom{});
    LAUNCH((k_cg_pupdate<V>), nb, pp, w, n, s->part_in, sc, ic, nb, 0, od, (int64_t)0,
           (V *)nullptr);
  }
  return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
}

// z = p_m(D^-1 K) D^-1 r: degree pc_cheb_degree, interval [lmax / 30, lmax]
// (Saad, Iterative Methods for Sparse Linear Systems, Alg. 12.1)
template <typename V>
int cheb_pc(iga_solver *s, const V *r, V *z, V *d, V *q, const V *val, hipStream_t st) {
  const int *od = s->ic + I_DONE;
  const int64_t n = s->n;
  const int nb = s->nb;
  const double lmax = s->cheb_lmax, lmin = lmax / 30.0;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  LAUNCH((k_cheb_init<V>), nb, r, z, d, s->fdinv, 1.0 / theta, n, od);
  const int bs = 256;
  for (int k = 1; k < s->o.pc_cheb_degree; ++k) {
    hipLaunchKernelGGL((k_spmv<V>), dim3((s->fem_rows + bs - 1) / bs), dim3(bs), 0, st, s->rowptr,
                       s->col, val, z, q, s->fem_rows, od, (const int *)nullptr);
    const double rho1 = 1.0 / (2.0 * sigma - rho);
    LAUNCH((k_cheb_step<V>), nb, r, q, z, d, s->fdinv, rho1 * rho, 2.0 * rho1 / delta, n, od);
    rho = rho1;
  }
  return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
}

template <typename V>
int apply_pc(iga_solver *s, bool f64, const V *r, V *z, hipStream_t st) {
  if (s->o.pc == IGA_PC_FEM_MG)
    return igai::mg_apply<V>(s->mg, r, z, st, s->ic + I_DONE, nullptr);
  if (s->o.pc == IGA_PC_FEM_ILU) {
    if (f64)
      hipLaunchKernelGGL((k_ilu_solve<double>), dim3(1), dim3(1024), 0, st, s->rowptr, s->col,
                         s->ilu64, s->ilu_diag, s->ilu_fptr, s->ilu_frows, s->ilu_nfl,
                         s->ilu_bptr, s->ilu_brows, s->ilu_nbl, (const double *)r, (double *)z,
                         s->ic + I_DONE);
    else
      hipLaunchKernelGGL((k_ilu_solve<float>), dim3(1), dim3(1024), 0, st, s->rowptr, s->col,
                         s->ilu32, s->ilu_diag, s->ilu_fptr, s->ilu_frows, s->ilu_nfl,
                         s->ilu_bptr, s->ilu_brows, s->ilu_nbl, (const float *)r, (float *)z,
                         s->ic + I_DONE);
    return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
  }
  if (s->o.pc == IGA_PC_FEM_CHEB) {
    if (f64)
      return cheb_pc<double>(s, (const double *)r, (double *)z, s->pp, s->q, s->val64, st);
    return cheb_pc<float>(s, (const float *)r, (float *)z, s->pp32, s->q32, s->val32, st);
  }
  if (f64)
    return fem_pc<double>(s, (const double *)r, (double *)z, s->rin, s->w, s->pp, s->q, s->val64,
                          st);
  return fem_pc<float>(s, (const float *)r, (float *)z, s->rin32, s->w32, s->pp32, s->q32,
                       s->val32, st);
}

// One CG iteration on vectors of type V (FP64 main solve or FP32 inner solve).
template <typename V, typename TU>
int cg_iteration(iga_solver *s, bool f64, TU *u, V *r, V *p, V *Ap, V *z, V *rold,
                 hipStream_t st) {
  const int64_t n = s->n;
  const int nb = s->nb;
  const bool gather = s->o.assembly == IGA_ASSEMBLY_INVLM_GATHER;
  const bool ebe = s->o.kernel == IGA_KERNEL_EBE;
  const bool loc = local_pc(s);
  // preconditioners that are not a fixed linear operator (the inner CG) or
  // not symmetric (ILU) take the flexible (Polak-Ribiere) beta
  const bool flex = !loc && s->o.pc != IGA_PC_FEM_MG;
  const int dt = apply_dt(s, f64);
  // p.Ap accumulated by the apply kernel itself when it can, else k_cg_pap
  const bool fdot = s->o.fused == 1 && s->tdot && igai::cg_fusable_dt(s->op, dt, ebe);
  // gather form with the fused p.Ap: k_cg_update sums the seam functions of
  // Ap itself (no seam kernel, no pass over y)
  const bool fseam = fdot && gather;
  const igai::CgHook hook{s->tdot, nullptr, nullptr, 0, fseam};
  const double *tdot = fdot ? s->tdot : nullptr;
  const int ntiles = f64 ? s->ntiles64 : s->ntiles32;
  // atomic form: Ap was zeroed by the previous k_cg_pupdate (or cg_run)
  int rc = igai::apply_raw_dt(s->op, dt, ebe, gather, p, Ap, 0, 1, st, s->ic + I_DONE,
                              /*y_zeroed=*/!gather, fdot ? &hook : nullptr);
  if (rc) return rc;
  if (!fdot)
    LAUNCH((k_cg_pap<V>), nb, p, Ap, n, s->part, s->sc, s->ic, (const int *)nullptr,
           (int64_t)0);
  const V *dinv = s->o.pc != IGA_PC_JACOBI
                      ? nullptr
                      : (f64 ? (const V *)(const void *)s->dinv : (const V *)(const void *)s->dinv32);
  const igai::SeamGeom sg = fseam ? igai::seam_geom(s->op, f64) : igai::SeamGeom{};
  if (loc && fseam)
    LAUNCH((k_cg_update<V, V, true, false, TU, true>), nb, u, r, p, Ap, z, (V *)nullptr, dinv, 0, n,
           s->part, s->sc, s->ic, nb, (const int *)nullptr, (int64_t)0, tdot, ntiles, sg);
  else if (loc)
    LAUNCH((k_cg_update<V, V, true, false, TU>), nb, u, r, p, Ap, z, (V *)nullptr, dinv, 0, n, s->part,
           s->sc, s->ic, nb, (const int *)nullptr, (int64_t)0, tdot, ntiles, sg);
  else if (fseam && flex)
    LAUNCH((k_cg_update<V, V, false, true, TU, true>), nb, u, r, p, Ap, z, rold, (const V *)nullptr,
           0, n, s->part, s->sc, s->ic, nb, (const int *)nullptr, (int64_t)0, tdot, ntiles, sg);
  else if (flex)
    LAUNCH((k_cg_update<V, V, false, true, TU>), nb, u, r, p, Ap, z, rold, (const V *)nullptr, 0, n,
           s->part, s->sc, s->ic, nb, (const int *)nullptr, (int64_t)0, tdot, ntiles, sg);
  else if (fseam)
    LAUNCH((k_cg_update<V, V, false, false, TU, true>), nb, u, r, p, Ap, z, (V *)nullptr,
           (const V *)nullptr, 0, n, s->part, s->sc, s->ic, nb, (const int *)nullptr, (int64_t)0,
           tdot, ntiles, sg);
  else
    LAUNCH((k_cg_update<V, V, false, false, TU>), nb, u, r, p, Ap, z, (V *)nullptr,
           (const V *)nullptr, 0, n, s->part, s->sc, s->ic, nb, (const int *)nullptr, (int64_t)0,
           tdot, ntiles, sg);
  if (!loc) {
    rc = apply_pc<V>(s, f64, r, z, st);
    if (rc) return rc;
    LAUNCH((k_cg_rz<V>), nb, r, z, flex ? rold : (const V *)nullptr, n, s->part, s->ic);
  }
  LAUNCH((k_cg_pupdate<V>), nb, p, z, n, s->part, s->sc, s->ic, nb, flex ? 1 : 0,
         (const int *)nullptr, (int64_t)0, gather ? (V *)nullptr : Ap);
  return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
}

// Enqueue chunks (launch(nit): nit iterations) until the device's done flag
// is set.  The flags and the scalars after a chunk are written to pinned
// memory (k_chunk_flags, the chunk's last kernel: launch(nit, slot) enqueues
// it) and read while the next chunk runs, so the device does not idle for
// the host round trip between chunks.  Every iteration enqueued past the stop
// still dispatches its kernels (they return at once, a few microseconds each
// with the dispatch gaps), so near the stop the run switches to short
// chunks: the squared residual's decay rate since the run's first chunk
// predicts the iteration at which it meets the stop threshold (tolerance or
// reliable-update level), and
//  - a stop predicted inside the chunk in flight enqueues nothing more (the
//    host waits for that chunk's flags),
//  - a stop predicted inside the next chunk enqueues a short one.
// A wrong prediction costs one host round trip or a few skipped iterations;
// the iterations themselves are the same.
// The done / iteration flags and the scalars after a chunk, written straight
// into the pinned host copies (coherent host memory: the stores cross to the
// host, and the system-scope fence orders them before the kernel's end,
// which the chunk's event follows).  One kernel instead of two
// device-to-host copies, each a runtime copy kernel with its own dispatch;
// inside a chunk's graph it follows the last iteration without the idle
// time a copy after the graph waits.
__global__ void k_chunk_flags(const int *ic, const double *sc, int *hflag, double *hsc) {
  const int t = threadIdx.x;
  if (t < 2) hflag[t] = ic[t];
  if (t < S_NS) hsc[t] = sc[t];
  __threadfence_system();
}

int chunk_flags(iga_solver *s, int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_chunk_flags, dim3(1), dim3(64), 0, st, s->ic, s->sc, s->dflag + 2 * slot,
                     s->dsc + S_NS * slot);
  return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
}

// Capture body(stream, slot) once per flag slot into gexec[slot].
template <typename F>
int capture_chunks(hipGraphExec_t *gexec, F &&body) {
  for (int slot = 0; slot < 2; ++slot) {
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return IGA_ERR_HIP;
    hipGraph_t g;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipStreamDestroy(cs);
      return IGA_ERR_HIP;
    }
    const int rc = body(cs, slot);
    hipError_t e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc || e != hipSuccess) return rc ? rc : IGA_ERR_HIP;
    e = hipGraphInstantiate(&gexec[slot], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      gexec[slot] = nullptr;
      return IGA_ERR_HIP;
    }
  }
  return IGA_OK;
}

template <typename L>
int run_chunks(iga_solver *s, hipStream_t st, L &&launch, int big, int small, int *iters,
               int *done) {
  int nit_of[2] = {0, 0};
  auto issue = [&](int slot, int nit) -> int {
    const int rc = launch(nit, slot);
    if (rc) return rc;
    if (hipEventRecord(s->fev[slot], st) != hipSuccess) return IGA_ERR_HIP;
    nit_of[slot] = nit;
    return IGA_OK;
  };
  // chunks in flight, oldest first (at most two)
  int q[2], nq = 0, next_slot = 0;
  auto push = [&](int nit) -> int {
    const int slot = next_slot;
    next_slot ^= 1;
    const int rc = issue(slot, nit);
    if (rc) return rc;
    q[nq++] = slot;
    return IGA_OK;
  };
  int rc = push(big);
  if (rc) return rc;
  bool have0 = false, wait_only = false;
  int it0 = 0, ahead = big;
  double lrr0 = 0.0;
  for (;;) {
    if (nq == 1 && !wait_only) {
      rc = push(ahead);
      if (rc) return rc;
    }
    const int cur = q[0];
    if (hipEventSynchronize(s->fev[cur]) != hipSuccess) return IGA_ERR_HIP;
    q[0] = q[1];
    --nq;
    const int *h = s->hflag + 2 * cur;
    if (h[I_DONE]) {
      *iters = h[I_ITER];
      *done = h[I_DONE];
      return IGA_OK;
    }
    const double *hs = s->hsc + S_NS * cur;
    const double rr = hs[S_RR], stop = std::fmax(hs[S_TOL2], hs[S_UPD2]);
    const int it = h[I_ITER];
    const int inflight = nq ? nit_of[q[0]] : 0;
    double need = 1e30;  // predicted iterations from it to the stop
    if (rr > 0.0 && stop > 0.0 && std::isfinite(rr)) {
      const double lrr = std::log(rr);
      if (!have0) {
        have0 = true;
        it0 = it;
        lrr0 = lrr;
      } else if (it > it0 && lrr < lrr0) {
        const double rate = (lrr - lrr0) / (it - it0);  // log decrease per iteration (< 0)
        need = (std::log(stop) - lrr) / rate;
      }
    }
    wait_only = need <= inflight;
    ahead = need <= inflight + big ? small : big;
    if (nq == 0) {  // nothing in flight: the next chunk goes now
      rc = push(wait_only ? small : ahead);
      if (rc) return rc;
      wait_only = true;
    }
  }
}

int chunk_big(const iga_solver *s) { return local_pc(s) ? CHUNK : CHUNK_PC; }
int chunk_small(const iga_solver *s) { return local_pc(s) ? CHUNK_SHORT : 1; }

template <typename V, typename TU>
int cg_chunk(iga_solver *s, bool f64, TU *u, V *r, V *p, V *Ap, V *z, V *rold, hipStream_t st,
             int nit) {
  for (int k = 0; k < nit; ++k) {
    int rc = cg_iteration<V, TU>(s, f64, u, r, p, Ap, z, rold, st);
    if (rc) return rc;
  }
  return IGA_OK;
}

template <typename V, typename TU>
int cg_resume(iga_solver *s, bool f64, TU *u, V *r, V *p, V *Ap, V *z, V *rold, hipStream_t st,
              hipGraphExec_t *gexec, int *iters, int *done);

// Run a CG to completion: init (k_cg_init + k_cg_init_fin) must already have
// been enqueued on st.
template <typename V, typename TU>
int cg_run(iga_solver *s, bool f64, TU *u, V *r, V *p, V *Ap, V *z, V *rold, hipStream_t st,
           hipGraphExec_t *gexec, int *iters, int *done) {
  const int64_t n = s->n;
  if (!local_pc(s)) {
    int rc = apply_pc<V>(s, f64, r, z, st);
    if (rc) return rc;
    LAUNCH((k_cg_rz0<V>), s->nb, r, z, p, n, s->part, s->ic);
    hipLaunchKernelGGL(k_cg_rz0_fin, dim3(1), dim3(RB), 0, st, s->part, s->nb, s->sc, s->ic);
  }
  // scatter-add form: the first apply adds into a zeroed Ap, later ones into
  // the Ap that k_cg_pupdate zeroed
  if (s->o.assembly != IGA_ASSEMBLY_INVLM_GATHER &&
      hipMemsetAsync(Ap, 0, (size_t)n * sizeof(V), st) != hipSuccess)
    return IGA_ERR_HIP;
  if (s->o.use_graph && !gexec[0]) {
    // capture one chunk and its flags per slot; the graphs read every
    // scalar from device memory
    const int rc = capture_chunks(gexec, [&](hipStream_t cs, int slot) {
      const int r2 = cg_chunk<V, TU>(s, f64, u, r, p, Ap, z, rold, cs, chunk_big(s));
      return r2 ? r2 : chunk_flags(s, slot, cs);
    });
    if (rc) return rc;
  }
  return cg_resume<V, TU>(s, f64, u, r, p, Ap, z, rold, st, gexec, iters, done);
}

// Continue a CG whose device state was restarted by k_ru_fin.
template <typename V, typename TU>
int cg_resume(iga_solver *s, bool f64, TU *u, V *r, V *p, V *Ap, V *z, V *rold, hipStream_t st,
              hipGraphExec_t *gexec, int *iters, int *done) {
  const int big = chunk_big(s);
  return run_chunks(
      s, st,
      [&](int nit, int slot) -> int {
        if (nit == big && s->o.use_graph && gexec[slot])
          return hipGraphLaunch(gexec[slot], st) == hipSuccess ? IGA_OK : IGA_ERR_HIP;
        const int rc = cg_chunk<V, TU>(s, f64, u, r, p, Ap, z, rold, st, nit);
        return rc ? rc : chunk_flags(s, slot, st);
      },
      big, chunk_small(s), iters, done);
}

// ---- fused iteration (local preconditioners, single solves)
bool use_fused(const iga_solver *s, bool f64) {
  return s->o.fused == 2 && local_pc(s) && s->cgs &&
         igai::cg_fusable_dt(s->op, apply_dt(s, f64), s->o.kernel == IGA_KERNEL_EBE);
}

// one fused iteration: apply and k_cgf_update (which, in the gather form,
// also sums w's seam functions); u, w, p, sv: the recursion's M r, A u,
// direction and A p
template <typename V, typename TU>
int cgf_iteration(iga_solver *s, bool f64, TU *x, V *r, V *p, V *sv, V *u, V *w,
                  hipStream_t st) {
  const bool gather = s->o.assembly == IGA_ASSEMBLY_INVLM_GATHER;
  const bool ebe = s->o.kernel == IGA_KERNEL_EBE;
  igai::CgHook hook{s->tdot, s->cgs + CG_L, s->cgs, CG_L, gather};
  // scatter-add form: w was cleared by the previous update (or cgf_run)
  int rc = igai::apply_raw_dt(s->op, apply_dt(s, f64), ebe, gather, u, w, 0, 1, st,
                              s->ic + I_DONE, /*y_zeroed=*/!gather, &hook);
  if (rc) return rc;
  const int ntiles = f64 ? s->ntiles64 : s->ntiles32;
  const V *dinv = s->o.pc == IGA_PC_JACOBI
                      ? (f64 ? (const V *)(const void *)s->dinv : (const V *)(const void *)s->dinv32)
                      : nullptr;
  const igai::SeamGeom sg = gather ? igai::seam_geom(s->op, f64) : igai::SeamGeom{};
  const int zw = gather ? 0 : 1;
  if (dinv && gather)
    LAUNCH((k_cgf_update<V, V, true, TU, true>), s->nb, x, r, p, sv, u, w, dinv, zw, s->n, s->cgs,
           s->tdot, ntiles, s->sc, s->ic, s->nb, sg);
  else if (dinv)
    LAUNCH((k_cgf_update<V, V, true, TU>), s->nb, x, r, p, sv, u, w, dinv, zw, s->n, s->cgs,
           s->tdot, ntiles, s->sc, s->ic, s->nb, sg);
  else if (gather)
    LAUNCH((k_cgf_update<V, V, false, TU, true>), s->nb, x, r, p, sv, u, w, dinv, zw, s->n,
           s->cgs, s->tdot, ntiles, s->sc, s->ic, s->nb, sg);
  else
    LAUNCH((k_cgf_update<V, V, false, TU>), s->nb, x, r, p, sv, u, w, dinv, zw, s->n, s->cgs,
           s->tdot, ntiles, s->sc, s->ic, s->nb, sg);
  return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
}

template <typename V, typename TU>
int cgf_chunk(iga_solver *s, bool f64, TU *x, V *r, V *p, V *sv, V *u, V *w, hipStream_t st,
              int nit) {
  for (int k = 0; k < nit; ++k) {
    int rc = cgf_iteration<V, TU>(s, f64, x, r, p, sv, u, w, st);
    if (rc) return rc;
  }
  return IGA_OK;
}

// Run (or, after k_cgf_ru_reset + k_cgf_ru_fin, continue) a fused recursion
// until its done flag is set; the init kernels must already be enqueued.
template <typename V, typename TU>
int cgf_run(iga_solver *s, bool f64, TU *x, V *r, V *p, V *sv, V *u, V *w, hipStream_t st,
            hipGraphExec_t *gexec, int *iters, int *done) {
  if (s->o.assembly != IGA_ASSEMBLY_INVLM_GATHER &&
      hipMemsetAsync(w, 0, (size_t)s->n * sizeof(V), st) != hipSuccess)
    return IGA_ERR_HIP;
  if (s->o.use_graph && !gexec[0]) {
    const int rc = capture_chunks(gexec, [&](hipStream_t cs, int slot) {
      const int r2 = cgf_chunk<V, TU>(s, f64, x, r, p, sv, u, w, cs, CHUNK);
      return r2 ? r2 : chunk_flags(s, slot, cs);
    });
    if (rc) return rc;
  }
  return run_chunks(
      s, st,
      [&](int nit, int slot) -> int {
        if (nit == CHUNK && s->o.use_graph)
          return hipGraphLaunch(gexec[slot], st) == hipSuccess ? IGA_OK : IGA_ERR_HIP;
        const int rc = cgf_chunk<V, TU>(s, f64, x, r, p, sv, u, w, st, nit);
        return rc ? rc : chunk_flags(s, slot, st);
      },
      CHUNK, CHUNK_SHORT, iters, done);
}

// |f - A u|^2 and |f|^2 into h (host), r = f - A u
int outer_residual(iga_solver *s, const double *f, const double *u, double *r, double *h,
                   hipStream_t st) {
  const bool gather = s->o.assembly == IGA_ASSEMBLY_INVLM_GATHER;
  const bool ebe = s->o.kernel == IGA_KERNEL_EBE;
  int rc = igai::apply_raw(s->op, true, ebe, gather, u, s->Au, 0, 1, st, nullptr);
  if (rc) return rc;
  double *osc = s->sc + 2 * S_NS;
  LAUNCH(k_residual, s->nb, f, s->Au, r, s->n, s->part);
  hipLaunchKernelGGL(k_fold2, dim3(1), dim3(RB), 0, st, s->part, s->nb, osc);
  if (hipMemcpyAsync(h, osc, 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return IGA_ERR_HIP;
  return IGA_OK;
}

// ---------------------------------------------------------------- block solve
// k independent systems, each an FP32 PCG with FP64 reliable updates (as
// PCG_F32_REFINE), advanced in lockstep: one multi-RHS apply per iteration
// and the CG kernels with one grid row (blockIdx.y) per column.  A column
// that asks for a reliable update, converges or stops is frozen by its own
// done flag until the host has handled it between chunks.

int block_buffers(iga_solver *s, int k) {
  if (k <= s->bk) return IGA_OK;
  for (void *p : s->ballocs) (void)hipFree(p);
  s->ballocs.clear();
  if (s->gblk) {
    (void)hipGraphExecDestroy(s->gblk);
    s->gblk = nullptr;
  }
  s->bk = 0;
  const size_t nk = (size_t)s->n * k;
  auto al = [&](void **p, size_t bytes) {
    if (hipMalloc(p, bytes ? bytes : 256) != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      return IGA_ERR_OUT_OF_MEMORY;
    }
    s->ballocs.push_back(*p);
    return IGA_OK;
  };
  int rc = 0;
  rc = rc ? rc : al((void **)&s->bx64, nk * 8);
  rc = rc ? rc : al((void **)&s->bd64, nk * 8);
  rc = rc ? rc : al((void **)&s->br32, nk * 4);
  rc = rc ? rc : al((void **)&s->bp32, nk * 4);
  rc = rc ? rc : al((void **)&s->bAp32, nk * 4);
  rc = rc ? rc : al((void **)&s->bz32, nk * 4);
  rc = rc ? rc : al((void **)&s->bsc, (size_t)k * S_NS * 8);
  rc = rc ? rc : al((void **)&s->bic, (size_t)k * I_NI * 4);
  rc = rc ? rc : al((void **)&s->bpart, (size_t)k * P_NP * RG * 8);
  if (rc) return rc;
  s->bk = k;
  return IGA_OK;
}

int bcg_chunk(iga_solver *s, int k, hipStream_t st) {
  const int64_t n = s->n;
  const int nb = s->nb;
  const float *dinv = s->o.pc == IGA_PC_JACOBI ? s->dinv32 : nullptr;
  const dim3 grid(nb, k);
  for (int it = 0; it < CHUNK; ++it) {
    int rc = igai::apply_block_raw(s->op, k, (int)n, s->bp32, s->bAp32, st);
    if (rc) return rc;
    hipLaunchKernelGGL((k_cg_pap<float>), grid, dim3(RB), 0, st, s->bp32, s->bAp32, n, s->bpart,
                       s->bsc, s->bic, (const int *)nullptr, n);
    hipLaunchKernelGGL((k_cg_update<float, float, true, false, double>), grid, dim3(RB), 0, st, s->bd64,
                       s->br32, s->bp32, s->bAp32, s->bz32, (float *)nullptr, dinv, 0, n, s->bpart,
                       s->bsc, s->bic, nb, (const int *)nullptr, n, (const double *)nullptr, 0,
                       igai::SeamGeom{});
    hipLaunchKernelGGL((k_cg_pupdate<float>), grid, dim3(RB), 0, st, s->bp32, s->bz32, n,
                       s->bpart, s->bsc, s->bic, nb, 0, (const int *)nullptr, n,
                       (float *)nullptr);
  }
  return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
}

int bcg_run_chunk(iga_solver *s, int k, hipStream_t st) {
  int cfg = -1;
  (void)iga_op_get_block_config(s->op, k, &cfg);
  if (!s->o.use_graph) return bcg_chunk(s, k, st);
  if (s->gblk && (s->gblk_k != k || s->gblk_cfg != cfg)) {
    (void)hipGraphExecDestroy(s->gblk);
    s->gblk = nullptr;
  }
  if (!s->gblk) {
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return IGA_ERR_HIP;
    hipGraph_t g;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipStreamDestroy(cs);
      return IGA_ERR_HIP;
    }
    int rc = bcg_chunk(s, k, cs);
    hipError_t e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc || e != hipSuccess) return rc ? rc : IGA_ERR_HIP;
    e = hipGraphInstantiate(&s->gblk, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      s->gblk = nullptr;
      return IGA_ERR_HIP;
    }
    s->gblk_k = k;
    s->gblk_cfg = cfg;
  }
  return hipGraphLaunch(s->gblk, st) == hipSuccess ? IGA_OK : IGA_ERR_HIP;
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
  o->inner_rtol = 0.0;  // automatic: see iga_solver_create
  o->inner_maxit = 100000;
  o->pc_inner_maxit = 20;
  o->pc_inner_rtol = 1e-2;
  o->pc_inner_abstol = 0.0;
  o->pc_cheb_degree = 8;
  o->use_graph = 1;
  o->fused = 1;
}

int iga_solver_destroy(iga_solver *s) {
  if (!s) return IGA_OK;
  (void)hipSetDevice(s->device);
  for (int k = 0; k < 2; ++k) {
    if (s->g64[k]) (void)hipGraphExecDestroy(s->g64[k]);
    if (s->g32[k]) (void)hipGraphExecDestroy(s->g32[k]);
    if (s->gf64[k]) (void)hipGraphExecDestroy(s->gf64[k]);
    if (s->gf32[k]) (void)hipGraphExecDestroy(s->gf32[k]);
  }
  if (s->gblk) (void)hipGraphExecDestroy(s->gblk);
  for (hipEvent_t e : s->fev)
    if (e) (void)hipEventDestroy(e);
  if (s->hflag) (void)hipHostFree(s->hflag);
  if (s->hsc) (void)hipHostFree(s->hsc);
  igai::mg_destroy(s->mg);
  for (void *p : s->allocs) (void)hipFree(p);
  for (void *p : s->ballocs) (void)hipFree(p);
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
  if (o.pc != IGA_PC_NONE && o.pc != IGA_PC_JACOBI && o.pc != IGA_PC_FEM_CG &&
      o.pc != IGA_PC_FEM_CHEB && o.pc != IGA_PC_FEM_ILU && o.pc != IGA_PC_FEM_MG)
    return IGA_ERR_INVALID_ARGUMENT;
  const bool fem_pc_kind = o.pc == IGA_PC_FEM_CG || o.pc == IGA_PC_FEM_CHEB ||
                           o.pc == IGA_PC_FEM_ILU || o.pc == IGA_PC_FEM_MG;
  if (fem_pc_kind && !fem) return IGA_ERR_INVALID_ARGUMENT;
  if (o.pc == IGA_PC_FEM_CHEB && o.pc_cheb_degree < 1) return IGA_ERR_INVALID_ARGUMENT;
  if (o.kind == IGA_SOLVE_PCG_F32_REFINE && !op->G32) return IGA_ERR_UNSUPPORTED;
  // automatic inner tolerance: with Jacobi (or none) longer FP32 recursions
  // lose iterations to rounding, so reliable updates come after each tenfold
  // reduction; with a FEM preconditioner the iterations are few and costly
  // and each update also costs a preconditioner application and the launches
  // past the stop, so they come after each 1e4-fold reduction (C2: 15.4
  // against 18.5 ms with the V-cycle, 63 iterations either way)
  if (o.inner_rtol == 0.0) o.inner_rtol = (o.pc == IGA_PC_NONE || o.pc == IGA_PC_JACOBI) ? 0.1 : 1e-4;
  if (!(o.rtol >= 0) || !(o.abstol >= 0) || o.maxit < 1 || !(o.inner_rtol > 0) ||
      o.inner_maxit < 1 || (o.pc == IGA_PC_FEM_CG && o.pc_inner_maxit < 1) ||
      !(o.pc_inner_abstol >= 0))
    return IGA_ERR_INVALID_ARGUMENT;
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  iga_solver *s = new (std::nothrow) iga_solver;
  if (!s) return IGA_ERR_OUT_OF_MEMORY;
  s->op = op;
  s->o = o;
  s->n = op->nfunc;
  s->nb = nblocks(s->n);
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
  SA(s->ic, 2 * I_NI * 4);
  SA(s->part, (size_t)P_NP * RG * 8);
  SA(s->part_in, (size_t)P_NP * RG * 8);
  if (o.kind == IGA_SOLVE_PCG_F32_REFINE) {
    SA(s->dinv32, n * 4);
    SA(s->d64, n * 8);
    SA(s->r32, n * 4);
    SA(s->p32, n * 4);
    SA(s->Ap32, n * 4);
    SA(s->z32, n * 4);
    SA(s->rold32, n * 4);
  }
  if (hipHostMalloc((void **)&s->hflag, 4 * sizeof(int), hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void **)&s->hsc, 2 * S_NS * sizeof(double), hipHostMallocCoherent) !=
          hipSuccess ||
      hipHostGetDevicePointer((void **)&s->dflag, s->hflag, 0) != hipSuccess ||
      hipHostGetDevicePointer((void **)&s->dsc, s->hsc, 0) != hipSuccess ||
      hipEventCreateWithFlags(&s->fev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->fev[1], hipEventDisableTiming) != hipSuccess)
    return fail(IGA_ERR_HIP);
  if (o.fused < 0 || o.fused > 2) return fail(IGA_ERR_INVALID_ARGUMENT);
  if (o.inner_f16_geometry &&
      (o.kind != IGA_SOLVE_PCG_F32_REFINE || !op->G16 || o.kernel == IGA_KERNEL_EBE))
    return fail(IGA_ERR_UNSUPPORTED);
  if (o.fused) {
    s->ntiles64 = igai::apply_tiles(op, true);
    s->ntiles32 = igai::apply_tiles(op, false);
    SA(s->tdot, (size_t)std::max(s->ntiles64, s->ntiles32) * 8);
    if (o.fused == 2 && !fem_pc_kind) SA(s->cgs, (size_t)2 * CG_L * 8);
  }
  if (hipMemset(s->ic, 0, 2 * I_NI * 4) != hipSuccess || hipMemset(s->sc, 0, 64 * 8) != hipSuccess)
    return fail(IGA_ERR_HIP);
  if (fem_pc_kind) {
    iga_csr_view v;
    if (iga_fem_view(fem, &v) != IGA_OK) return fail(IGA_ERR_INVALID_ARGUMENT);
    if (v.nrows != s->n) return fail(IGA_ERR_BAD_MESH);
    s->fem_rows = v.nrows;
    s->rowptr = v.rowptr;
    s->col = v.col;
    s->val64 = v.val;
    if (o.pc == IGA_PC_FEM_MG) {
      // the hierarchy is built on the host from a copy of the CSR
      std::vector<int> hrp(v.nrows + 1), hci(v.nnz);
      std::vector<double> hv(v.nnz);
      if (hipMemcpy(hrp.data(), v.rowptr, (v.nrows + 1) * 4, hipMemcpyDeviceToHost) !=
              hipSuccess ||
          hipMemcpy(hci.data(), v.col, v.nnz * 4, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(hv.data(), v.val, v.nnz * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(IGA_ERR_HIP);
      const bool f64 = o.kind == IGA_SOLVE_PCG_F64;
      rc = igai::mg_create(hrp.data(), hci.data(), hv.data(), op->nfx, op->nfy, f64, !f64,
                           &s->mg);
      if (rc) return fail(rc);
    }
  }
  if (fem_pc_kind && o.pc != IGA_PC_FEM_MG) {
    iga_csr_view v;
    if (iga_fem_view(fem, &v) != IGA_OK) return fail(IGA_ERR_INVALID_ARGUMENT);
    SA(s->fdinv, n * 8);
    SA(s->rin, n * 8);
    SA(s->w, n * 8);
    SA(s->pp, n * 8);
    SA(s->q, n * 8);
    const int bs = 256;
    hipLaunchKernelGGL(k_csr_diag_inv, dim3((v.nrows + bs - 1) / bs), dim3(bs), 0, nullptr,
                       v.rowptr, v.col, v.val, s->fdinv, v.nrows);
    if (o.kind == IGA_SOLVE_PCG_F32_REFINE) {
      SA(s->val32, (size_t)v.nnz * 4);
      SA(s->rin32, n * 4);
      SA(s->w32, n * 4);
      SA(s->pp32, n * 4);
      SA(s->q32, n * 4);
      hipLaunchKernelGGL((k_cast<float>), dim3(nblocks(v.nnz)), dim3(RB), 0, nullptr, v.val,
                         s->val32, (int64_t)v.nnz);
    }
    if (o.pc == IGA_PC_FEM_CHEB) {
      std::vector<double> hp(RG);
      hipLaunchKernelGGL(k_gersh, dim3(RG), dim3(RB), 0, nullptr, v.rowptr, v.col, v.val,
                         v.nrows, s->part);
      if (hipMemcpy(hp.data(), s->part, RG * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(IGA_ERR_HIP);
      for (double x : hp) s->cheb_lmax = std::max(s->cheb_lmax, x);
      if (!(s->cheb_lmax > 0.0)) return fail(IGA_ERR_BAD_MESH);
    }
    if (o.pc == IGA_PC_FEM_ILU) {
      const int nr = v.nrows;
      std::vector<int> hrp(nr + 1), hci(v.nnz), hdiag, fptr, frows, bptr, brows;
      std::vector<double> hlu(v.nnz);
      if (hipMemcpy(hrp.data(), v.rowptr, (nr + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(hci.data(), v.col, v.nnz * 4, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(hlu.data(), v.val, v.nnz * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(IGA_ERR_HIP);
      if (!ilu0_factor(nr, hrp, hci, hlu, hdiag)) return fail(IGA_ERR_BAD_MESH);
      ilu0_levels(nr, hrp, hci, hdiag, true, fptr, frows);
      ilu0_levels(nr, hrp, hci, hdiag, false, bptr, brows);
      s->ilu_nfl = (int)fptr.size() - 1;
      s->ilu_nbl = (int)bptr.size() - 1;
      SA(s->ilu64, (size_t)v.nnz * 8);
      SA(s->ilu_diag, (size_t)nr * 4);
      SA(s->ilu_fptr, fptr.size() * 4);
      SA(s->ilu_frows, (size_t)nr * 4);
      SA(s->ilu_bptr, bptr.size() * 4);
      SA(s->ilu_brows, (size_t)nr * 4);
      if (hipMemcpy(s->ilu64, hlu.data(), v.nnz * 8, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(s->ilu_diag, hdiag.data(), nr * 4, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(s->ilu_fptr, fptr.data(), fptr.size() * 4, hipMemcpyHostToDevice) !=
              hipSuccess ||
          hipMemcpy(s->ilu_frows, frows.data(), nr * 4, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(s->ilu_bptr, bptr.data(), bptr.size() * 4, hipMemcpyHostToDevice) !=
              hipSuccess ||
          hipMemcpy(s->ilu_brows, brows.data(), nr * 4, hipMemcpyHostToDevice) != hipSuccess)
        return fail(IGA_ERR_HIP);
      if (o.kind == IGA_SOLVE_PCG_F32_REFINE) {
        SA(s->ilu32, (size_t)v.nnz * 4);
        hipLaunchKernelGGL((k_cast<float>), dim3(nblocks(v.nnz)), dim3(RB), 0, nullptr, s->ilu64,
                           s->ilu32, (int64_t)v.nnz);
      }
    }
  }
#undef SA
  if (hipDeviceSynchronize() != hipSuccess) return fail(IGA_ERR_HIP);
  *out = s;
  return IGA_OK;
}

int iga_solver_apply_pc(iga_solver *s, int dt, const void *r, void *z, void *stream) {
  if (!s || !r || !z || r == z || (dt != IGA_F64 && dt != IGA_F32))
    return IGA_ERR_INVALID_ARGUMENT;
  if (local_pc(s)) return IGA_ERR_UNSUPPORTED;
  const bool f64 = dt == IGA_F64;
  // the FP64 solve keeps the FP64 preconditioner, the FP32 inner solve the FP32 one
  if (f64 != (s->o.kind == IGA_SOLVE_PCG_F64)) return IGA_ERR_UNSUPPORTED;
  int rc = igai::set_device(s->device);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  // the preconditioner kernels return at once while the solver's done flag
  // is set (it stays set after a solve): clear it
  hipLaunchKernelGGL(k_ru_clear, dim3(1), dim3(64), 0, st, s->ic);
  if (f64) return apply_pc<double>(s, true, (const double *)r, (double *)z, st);
  return apply_pc<float>(s, false, (const float *)r, (float *)z, st);
}

int iga_solver_solve(iga_solver *s, const double *f, double *u, void *stream,
                     iga_solve_stats *stats) {
  if (!s || !f || !u || f == u) return IGA_ERR_INVALID_ARGUMENT;
  int rc = igai::set_device(s->device);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = s->n;
  const int nb = s->nb;
  const iga_solver_opts &o = s->o;
  iga_solve_stats stt;
  std::memset(&stt, 0, sizeof(stt));
  const bool loc = local_pc(s);
  const double *dinv = o.pc == IGA_PC_JACOBI ? s->dinv : nullptr;
  const float *dinv32 = o.pc == IGA_PC_JACOBI ? s->dinv32 : nullptr;
  if (s->gen != s->op->gen) {
    // Jacobi of A for the operator's current G_q
    iga_op *op = s->op;
    if (op->ops->diag(s->dinv, op->G64, op->tabx64, op->taby64, op->wts64, op->dirmask, op->nex,
                      op->ney, op->nfx, op->nfy, op->L64, st) != hipSuccess)
      return IGA_ERR_HIP;
    LAUNCH(k_inv, nb, s->dinv, n);
    if (s->dinv32) LAUNCH((k_cast<float>), nb, s->dinv, s->dinv32, n);
    s->gen = op->gen;
  }
  // a right-hand side with a NaN or an infinity: an error, not a divergence
  // (the reference would iterate on it until divtol or maxit)
  {
    LAUNCH(k_dot, nb, f, f, n, s->part);
    hipLaunchKernelGGL(k_fold1, dim3(1), dim3(RB), 0, st, s->part, nb, s->sc + 2 * S_NS + 8);
    double ff = 0.0;
    if (hipMemcpyAsync(&ff, s->sc + 2 * S_NS + 8, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return IGA_ERR_HIP;
    if (!std::isfinite(ff)) {
      stt.status = IGA_ERR_NOT_FINITE;
      if (stats) *stats = stt;
      return IGA_ERR_NOT_FINITE;
    }
  }
  const auto t0 = std::chrono::steady_clock::now();
  int status = IGA_OK;
  if (o.kind == IGA_SOLVE_PCG_F64) {
    int iters = 0, done = 0;
    if (use_fused(s, true)) {
      // rold serves as s = A p of the fused recursion
      LAUNCH((k_cgf_init<double, double, double>), nb, f, 1.0, s->x64, s->r, s->z, s->p, s->rold,
             dinv, n, s->cgs);
      hipLaunchKernelGGL(k_cgf_init_fin, dim3(1), dim3(RB), 0, st, s->cgs, nb, s->sc, s->ic,
                         o.rtol * o.rtol, o.abstol * o.abstol, o.divtol * o.divtol, o.maxit, 0.0);
      rc = cgf_run<double, double>(s, true, s->x64, s->r, s->p, s->rold, s->z, s->Ap, st, s->gf64,
                           &iters, &done);
    } else {
      LAUNCH((k_cg_init<double, double, double>), nb, f, 1.0, s->x64, s->r, s->z, s->p, dinv,
             loc ? 1 : 0, n, s->part);
      hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(RB), 0, st, s->part, nb, s->sc, s->ic,
                         o.rtol * o.rtol, o.abstol * o.abstol, o.divtol * o.divtol, o.maxit,
                         0.0);
      rc = cg_run<double, double>(s, true, s->x64, s->r, s->p, s->Ap, s->z, s->rold, st, s->g64, &iters,
                          &done);
    }
    if (rc) return rc;
    if (hipMemcpyAsync(u, s->x64, n * 8, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return IGA_ERR_HIP;
    stt.iterations = iters;
    stt.outer_iterations = 1;
    status = done == 1 ? IGA_OK : (done == 2 ? IGA_ERR_NOT_CONVERGED : IGA_ERR_DIVERGED);
  } else {
    // FP32 PCG with reliable updates: whenever its recursive residual has
    // dropped by RU_FACTOR (or looks converged), the FP32 iterate is added
    // to the FP64 solution, the true residual f - A u is formed with the FP64
    // operator, and the recursion restarts from it keeping the search
    // direction -- FP64 defect correction without losing the Krylov rate.
    const double RU2 = o.inner_rtol * o.inner_rtol;
    if (hipMemsetAsync(s->x64, 0, n * 8, st) != hipSuccess) return IGA_ERR_HIP;
    double h[2];
    rc = outer_residual(s, f, s->x64, s->r, h, st);
    if (rc) return rc;
    const double ff = h[1];
    double rr = h[0];
    const double tol = std::fmax(o.rtol * std::sqrt(ff), o.abstol);
    const double tol2 = tol * tol;
    int outer = 0, iters = 0, done = 0;
    status = IGA_ERR_NOT_CONVERGED;
    if (!(rr == rr)) {
      status = IGA_ERR_DIVERGED;
    } else if (rr <= tol2) {
      status = IGA_OK;
    } else if (use_fused(s, false)) {
      // rold32 serves as s = A p of the fused recursion
      LAUNCH((k_cgf_init<float, double, float, double>), nb, s->r, 1.0, s->d64, s->r32, s->z32, s->p32,
             s->rold32, dinv32, n, s->cgs);
      hipLaunchKernelGGL(k_cgf_init_fin, dim3(1), dim3(RB), 0, st, s->cgs, nb, s->sc, s->ic,
                         tol2 / rr, 0.0, o.divtol * o.divtol * ff / rr, o.maxit, RU2);
      for (;;) {
        rc = cgf_run<float, double>(s, false, s->d64, s->r32, s->p32, s->rold32, s->z32, s->Ap32, st,
                            s->gf32, &iters, &done);
        if (rc) return rc;
        if (done == 3) { status = IGA_ERR_DIVERGED; break; }
        // reliable update
        ++outer;
        LAUNCH(k_ru_accum<double>, nb, s->x64, s->d64, n);
        rc = outer_residual(s, f, s->x64, s->r, h, st);
        if (rc) return rc;
        rr = h[0];
        const int oc = igai::outer_outcome(rr, ff, tol2, o.divtol, done == 2);
        if (oc != igai::OUTCOME_CONTINUE) { status = oc; break; }
        LAUNCH(k_cgf_ru_reset, nb, s->r, s->r32, s->z32, s->rold32, dinv32, n, s->cgs);
        hipLaunchKernelGGL(k_cgf_ru_fin, dim3(1), dim3(64), 0, st, s->sc, s->ic, RU2 * rr);
      }
    } else {
      LAUNCH((k_cg_init<float, double, float, double>), nb, s->r, 1.0, s->d64, s->r32, s->z32, s->p32,
             dinv32,
             loc ? 1 : 0, n, s->part);
      hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(RB), 0, st, s->part, nb, s->sc, s->ic,
                         tol2 / rr, 0.0, o.divtol * o.divtol * ff / rr, o.maxit, RU2);
      bool first = true;
      for (;;) {
        if (first) {
          rc = cg_run<float, double>(s, false, s->d64, s->r32, s->p32, s->Ap32, s->z32, s->rold32, st,
                             s->g32, &iters, &done);
          first = false;
        } else {
          rc = cg_resume<float, double>(s, false, s->d64, s->r32, s->p32, s->Ap32, s->z32, s->rold32, st,
                                s->g32, &iters, &done);
        }
        if (rc) return rc;
        if (done == 3) { status = IGA_ERR_DIVERGED; break; }
        // reliable update
        ++outer;
        LAUNCH(k_ru_accum<double>, nb, s->x64, s->d64, n);
        rc = outer_residual(s, f, s->x64, s->r, h, st);
        if (rc) return rc;
        rr = h[0];
        const int oc = igai::outer_outcome(rr, ff, tol2, o.divtol, done == 2);
        if (oc != igai::OUTCOME_CONTINUE) { status = oc; break; }
        LAUNCH(k_ru_reset, nb, s->r, s->r32, s->z32, dinv32, loc ? 1 : 0, n, s->part);
        if (!loc) {
          hipLaunchKernelGGL(k_ru_clear, dim3(1), dim3(64), 0, st, s->ic);
          rc = apply_pc<float>(s, false, s->r32, s->z32, st);
          if (rc) return rc;
          LAUNCH(k_ru_rz, nb, s->r32, s->z32, n, s->part);
        }
        hipLaunchKernelGGL(k_ru_fin, dim3(1), dim3(RB), 0, st, s->part, nb, s->sc, s->ic,
                           RU2 * rr);
      }
    }
    if (hipMemcpyAsync(u, s->x64, n * 8, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return IGA_ERR_HIP;
    stt.iterations = iters;
    stt.outer_iterations = outer;
  }
  // final FP64 residual (into Ap as scratch)
  double h[2];
  rc = outer_residual(s, f, u, s->Ap, h, st);
  if (rc) return rc;
  const auto t1 = std::chrono::steady_clock::now();
  stt.rel_residual = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
  stt.solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  stt.status = status;
  if (stats) *stats = stt;
  return status;
}

int iga_solver_solve_block(iga_solver *s, int k, int ld, const double *F, double *U,
                           void *stream, iga_solve_stats *stats) {
  if (!s || k < 1 || !F || !U || ld < s->n) return IGA_ERR_INVALID_ARGUMENT;
  const iga_solver_opts &o = s->o;
  if (o.kind != IGA_SOLVE_PCG_F32_REFINE || !local_pc(s) || s->op->rational == 2 ||
      !s->op->G32)
    return IGA_ERR_UNSUPPORTED;
  int rc = igai::set_device(s->device);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = s->n;
  const int nb = s->nb;
  rc = block_buffers(s, k);
  if (rc) return rc;
  const float *dinv32 = o.pc == IGA_PC_JACOBI ? s->dinv32 : nullptr;
  if (s->gen != s->op->gen) {
    iga_op *op = s->op;
    if (op->ops->diag(s->dinv, op->G64, op->tabx64, op->taby64, op->wts64, op->dirmask, op->nex,
                      op->ney, op->nfx, op->nfy, op->L64, st) != hipSuccess)
      return IGA_ERR_HIP;
    LAUNCH(k_inv, nb, s->dinv, n);
    LAUNCH((k_cast<float>), nb, s->dinv, s->dinv32, n);
    s->gen = op->gen;
  }
  if (hipStreamSynchronize(st) != hipSuccess) return IGA_ERR_HIP;
  const auto t0 = std::chrono::steady_clock::now();
  const double RU2 = o.inner_rtol * o.inner_rtol;
  std::vector<double> ff(k), tol2(k), rrf(k);
  std::vector<int> status(k, IGA_ERR_NOT_CONVERGED), outer(k, 0), iters(k, 0), fin(k, 0);
  auto colp = [&](int j) { return s->bpart + (size_t)j * P_NP * RG; };
  auto colsc = [&](int j) { return s->bsc + (size_t)j * S_NS; };
  auto colic = [&](int j) { return s->bic + (size_t)j * I_NI; };
  if (hipMemsetAsync(s->bx64, 0, (size_t)k * n * 8, st) != hipSuccess) return IGA_ERR_HIP;
  int running = 0;
  for (int j = 0; j < k; ++j) {
    const double *fj = F + (size_t)j * ld;
    double h[2];
    rc = outer_residual(s, fj, s->bx64 + (size_t)j * n, s->r, h, st);
    if (rc) return rc;
    ff[j] = h[1];
    rrf[j] = h[0];
    const double tol = std::fmax(o.rtol * std::sqrt(ff[j]), o.abstol);
    tol2[j] = tol * tol;
    LAUNCH((k_cg_init<float, double, float, double>), nb, s->r, 1.0, s->bd64 + (size_t)j * n,
           s->br32 + (size_t)j * n, s->bz32 + (size_t)j * n, s->bp32 + (size_t)j * n, dinv32, 1,
           n, colp(j));
    const double rr = h[0];
    const bool bad_f = !std::isfinite(ff[j]);
    const bool done = bad_f || !(rr == rr) || rr <= tol2[j];
    hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(RB), 0, st, colp(j), nb, colsc(j), colic(j),
                       done ? 1.0 : tol2[j] / rr, 0.0, o.divtol * o.divtol * ff[j] / rr,
                       o.maxit, RU2);
    if (bad_f) {
      status[j] = IGA_ERR_NOT_FINITE;
      fin[j] = 1;
    } else if (!(rr == rr)) {
      status[j] = IGA_ERR_DIVERGED;
      fin[j] = 1;
    } else if (rr <= tol2[j]) {
      status[j] = IGA_OK;
      fin[j] = 1;
    } else {
      ++running;
    }
  }
  // columns final at the start carry a nonzero done flag (k_cg_init_fin with
  // rtol2 = 1, or a non-finite residual): the chunk kernels skip them
  std::vector<int> hic((size_t)k * I_NI);
  while (running) {
    rc = bcg_run_chunk(s, k, st);
    if (rc) return rc;
    if (hipMemcpyAsync(hic.data(), s->bic, hic.size() * 4, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return IGA_ERR_HIP;
    running = 0;
    for (int j = 0; j < k; ++j) {
      if (fin[j]) continue;
      const int code = hic[(size_t)j * I_NI + I_DONE];
      iters[j] = hic[(size_t)j * I_NI + I_ITER];
      if (code == 0) {
        ++running;
        continue;
      }
      if (code == 3) {
        status[j] = IGA_ERR_DIVERGED;
        fin[j] = 1;
        continue;
      }
      // reliable update of column j with the FP64 operator
      ++outer[j];
      double *xj = s->bx64 + (size_t)j * n;
      LAUNCH(k_ru_accum<double>, nb, xj, s->bd64 + (size_t)j * n, n);
      double h[2];
      rc = outer_residual(s, F + (size_t)j * ld, xj, s->r, h, st);
      if (rc) return rc;
      rrf[j] = h[0];
      const int oc = igai::outer_outcome(h[0], ff[j], tol2[j], o.divtol, code == 2);
      if (oc != igai::OUTCOME_CONTINUE) {
        status[j] = oc;
        fin[j] = 1;
      } else {
        LAUNCH(k_ru_reset, nb, s->r, s->br32 + (size_t)j * n, s->bz32 + (size_t)j * n, dinv32, 1,
               n, colp(j));
        hipLaunchKernelGGL(k_ru_fin, dim3(1), dim3(RB), 0, st, colp(j), nb, colsc(j), colic(j),
                           RU2 * h[0]);
        ++running;
      }
    }
  }
  for (int j = 0; j < k; ++j)
    if (hipMemcpyAsync(U + (size_t)j * ld, s->bx64 + (size_t)j * n, n * 8,
                       hipMemcpyDeviceToDevice, st) != hipSuccess)
      return IGA_ERR_HIP;
  if (hipStreamSynchronize(st) != hipSuccess) return IGA_ERR_HIP;
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  int first_bad = IGA_OK;
  for (int j = 0; j < k; ++j) {
    if (stats) {
      stats[j].iterations = iters[j];
      stats[j].outer_iterations = outer[j];
      stats[j].rel_residual = ff[j] > 0 ? std::sqrt(rrf[j] / ff[j]) : std::sqrt(rrf[j]);
      stats[j].solve_ms = ms;
      stats[j].status = status[j];
    }
    if (status[j] != IGA_OK && first_bad == IGA_OK) first_bad = status[j];
  }
  return first_bad;
}

int iga_dot_f64(const double *a, const double *b, int64_t n, double *out_host, void *stream) {
  if (!a || !b || !out_host || n < 0) return IGA_ERR_INVALID_ARGUMENT;
  hipStream_t st = (hipStream_t)stream;
  double *buf = nullptr;
  if (hipMallocAsync((void **)&buf, (size_t)(P_NP * RG + 8) * 8, st) != hipSuccess)
    return IGA_ERR_OUT_OF_MEMORY;
  const int nb = nblocks(n);
  LAUNCH(k_dot, nb, a, b, n, buf);
  hipLaunchKernelGGL(k_fold1, dim3(1), dim3(RB), 0, st, buf, nb, buf + P_NP * RG);
  double r = 0.0;
  if (hipMemcpyAsync(&r, buf + P_NP * RG, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return IGA_ERR_HIP;
  (void)hipFreeAsync(buf, st);
  *out_host = r;
  return IGA_OK;
}

}  // extern "C"
