// iga_op.hip -- operator set-up and the C entry points of the apply.
//
// Replaces GlobalAssemblyMF (.cu) / MatrixFreeMatMultCUDA of the reference
// (matrix_free_gpu/src/GlobalAssemblyMF.cu:141-242, matrix_free_gpu/src/
// mult.cu:338-558): the mesh is validated and uploaded once, the 1D basis
// tables (C_e . Bernstein, with tensor-product NURBS weights folded in) and
// the per-quadrature-point factors G_q are built once, and each apply is two
// asynchronous launches with no allocation and no synchronisation.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "block.hpp"
#include "op_internal.hpp"

namespace igak {
const Ops *ops_p1();
const Ops *ops_p2();
const Ops *ops_p3();
const Ops *ops_p4();
const Ops *ops_p5();
const Ops *ops_p6();
const Ops *ops_for_degree(int p) {
  switch (p) {
    case 1: return ops_p1();
    case 2: return ops_p2();
    case 3: return ops_p3();
    case 4: return ops_p4();
    case 5: return ops_p5();
    case 6: return ops_p6();
  }
  return nullptr;
}
}  // namespace igak

namespace {

// Gauss-Legendre on [0,1] by Newton iteration on P_n (weights sum to one).
void gauss01(int n, double *x, double *w) {
  const double pi = 3.14159265358979323846264338327950288;
  for (int i = 0; i < n; ++i) {
    double t = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const double pk = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
      }
      if (n == 1) { p0 = 1.0; p1 = t; }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      const double dt = p1 / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-17) break;
    }
    double p0 = 1.0, p1 = t;
    for (int k = 2; k <= n; ++k) {
      const double pk = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = pk;
    }
    if (n == 1) { p0 = 1.0; p1 = t; }
    dp = n * (t * p1 - p0) / (t * t - 1.0);
    x[n - 1 - i] = 0.5 * t + 0.5;
    w[n - 1 - i] = 1.0 / ((1.0 - t * t) * dp * dp);
  }
}

// Bernstein values and derivatives of degree p at xi (de Casteljau-style
// recurrence on the degree).
void bernstein(int p, double xi, double *B, double *dB) {
  double lo[16] = {1.0};
  for (int d = 1; d < p; ++d) {
    double nx[16];
    for (int k = 0; k <= d; ++k)
      nx[k] = (k < d ? (1.0 - xi) * lo[k] : 0.0) + (k > 0 ? xi * lo[k - 1] : 0.0);
    for (int k = 0; k <= d; ++k) lo[k] = nx[k];
  }
  // lo = degree p-1 basis (p >= 1)
  for (int k = 0; k <= p; ++k) {
    const double a = (k < p) ? lo[k] : 0.0, b = (k > 0) ? lo[k - 1] : 0.0;
    B[k] = (1.0 - xi) * a + xi * b;
    dB[k] = p * (b - a);
  }
}

// 1D tables of one direction: for each element, N[i][a] and dN/dxi[i][a] at
// the Gauss points (reference coordinate on [0,1]), rational when w1 != null
// (w1 = the direction's factor of tensor-product weights).
std::vector<double> basis_tables(int p, int nq, int nelem, const double *C,
                                 const double *w1, const double *gx) {
  const int nb = p + 1, tab = nq * nb;
  std::vector<double> out((size_t)nelem * 2 * tab);
  double B[16], dB[16];
  for (int e = 0; e < nelem; ++e) {
    const double *Ce = C + (size_t)e * nb * nb;
    double *o = &out[(size_t)e * 2 * tab];
    for (int i = 0; i < nq; ++i) {
      bernstein(p, gx[i], B, dB);
      double N[16], D[16], W = 0.0, dW = 0.0;
      for (int a = 0; a < nb; ++a) {
        double s = 0.0, d = 0.0;
        for (int k = 0; k < nb; ++k) {
          s += Ce[a * nb + k] * B[k];
          d += Ce[a * nb + k] * dB[k];
        }
        const double wa = w1 ? w1[e + a] : 1.0;
        N[a] = wa * s;
        D[a] = wa * d;
        W += N[a];
        dW += D[a];
      }
      for (int a = 0; a < nb; ++a) {
        double R = N[a], dR = D[a];
        if (w1) {
          R = N[a] / W;
          dR = (D[a] - dW * R) / W;
        }
        o[i * nb + a] = R;
        o[tab + i * nb + a] = dR;
      }
    }
  }
  return out;
}

int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace

namespace igai {

int set_device(int dev) {
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return IGA_ERR_HIP;
  if (cur != dev && hipSetDevice(dev) != hipSuccess) return IGA_ERR_HIP;
  return IGA_OK;
}

int dev_alloc(iga_op *op, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return e == hipErrorOutOfMemory ? IGA_ERR_OUT_OF_MEMORY : IGA_ERR_HIP;
  }
  op->allocs.push_back(*p);
  op->bytes += (int64_t)bytes;
  return IGA_OK;
}

bool cg_fusable(const iga_op *op, bool f64, bool ebe) {
  // the persistent variant (1, sum-factorised, no per-function weights) has
  // no dot output
  return op->rational == 2 || ebe || (f64 ? op->variant64 : op->variant32) != 1;
}

bool cg_fusable_dt(const iga_op *op, int dt, bool ebe) {
  return dt == IGA_F32_G16 || cg_fusable(op, dt == IGA_F64, ebe);
}

int apply_tiles(const iga_op *op, bool f64) {
  const igak::Layout &L = f64 ? op->L64 : op->L32;
  return L.ntx * L.nty;
}

namespace {
constexpr int FOLD_U = 4;  // functions per thread of k_slot_fold
// The gather's seam pass: y[i] = ((y[i] + s1[i]) + s2[i]) + s3[i] over the
// function slots (seam_load, seam_value), stored only where a slot holds a
// partial (the other functions' y is already whole).  Coalesced; out-of-range
// lanes load the last function so that the wave votes in seam_load see them.
template <typename T>
__global__ __launch_bounds__(256) void k_slot_fold(T *__restrict__ y, SeamGeom g,
                                                   const int *skip) {
  if (skip && *skip) return;
  const int64_t b = (int64_t)blockIdx.x * 256 * FOLD_U + threadIdx.x;
  T a0[FOLD_U];
  SeamPart<T> sp[FOLD_U];
#pragma unroll
  for (int k = 0; k < FOLD_U; ++k) {
    const int64_t i = b + (int64_t)k * 256;
    a0[k] = seam_load(y, g, i < g.n ? i : g.n - 1, sp[k]);
  }
#pragma unroll
  for (int k = 0; k < FOLD_U; ++k) {
    const int64_t i = b + (int64_t)k * 256;
    if (i < g.n && (sp[k].a1 != T(0) || sp[k].a2 != T(0) || sp[k].a3 != T(0)))
      y[i] = seam_value(a0[k], sp[k]);
  }
}
}  // namespace

int apply_raw(iga_op *op, bool f64, bool ebe, bool gather, const void *x, void *y,
              int mask_in, int mask_out, hipStream_t s, const int *skip, bool y_zeroed,
              const CgHook *hook) {
  return apply_raw_dt(op, f64 ? IGA_F64 : IGA_F32, ebe, gather, x, y, mask_in, mask_out, s, skip,
                      y_zeroed, hook);
}

int apply_raw_dt(iga_op *op, int dt, bool ebe, bool gather, const void *x, void *y,
                 int mask_in, int mask_out, hipStream_t s, const int *skip, bool y_zeroed,
                 const CgHook *hook) {
  const bool f64 = dt == IGA_F64, g16 = dt == IGA_F32_G16;
  const igak::Layout &L = f64 ? op->L64 : op->L32;
  if (g16 && (!op->G16 || ebe)) return IGA_ERR_UNSUPPORTED;
  if (hook && !cg_fusable_dt(op, dt, ebe)) return IGA_ERR_UNSUPPORTED;
  igak::ApplyArgs a = {};
  a.gscale = op->g16_scale;
  if (hook) {
    a.dot = hook->dot;
    a.hsrc = hook->hsrc;
    a.hdst = hook->hdst;
    a.hn = hook->hn;
  }
  a.x = x;
  a.y = y;
  a.G = f64 ? (const void *)op->G64 : (g16 ? (const void *)op->G16 : (const void *)op->G32);
  a.tabx = f64 ? (const void *)op->tabx64 : (const void *)op->tabx32;
  a.taby = f64 ? (const void *)op->taby64 : (const void *)op->taby32;
  a.dirmask = op->dirmask;
  a.wf = f64 ? (const void *)op->wts64 : (const void *)op->wts32;
  a.nex = op->nex;
  a.ney = op->ney;
  a.nfx = op->nfx;
  a.nfy = op->nfy;
  a.ntx = L.ntx;
  a.nty = L.nty;
  a.mode = gather ? 1 : 0;
  a.slots = f64 ? (void *)op->slots64 : (void *)op->slots32;
  a.mask_in = mask_in;
  a.mask_out = mask_out;
  a.skip = skip;
  a.rev = op->rev_next;
  a.ntg = dt == IGA_F32 && op->stream32;
  op->rev_next ^= 1;
  if (!a.G) return IGA_ERR_UNSUPPORTED;
  const int ntiles = L.ntx * L.nty;
  const int tile = f64 ? op->tile64 : op->tile32;
  if (!gather && !y_zeroed) {
    if (f64)
      IGA_HIP_TRY(op->ops->seam64((double *)y, op->nex, op->nfx, op->nfy, L, skip, s));
    else
      IGA_HIP_TRY(op->ops->seam32((float *)y, op->nex, op->nfx, op->nfy, L, skip, s));
  }
  const bool prof = op->prof_on && op->prof_used < op->prof_cap;
  if (prof) IGA_HIP_TRY(hipEventRecord(op->prof_ev[2 * op->prof_used], s));
  if (g16)
    IGA_HIP_TRY(op->ops->apply16(a, tile, ntiles, op->variant32, s));
  else if (a.wf)
    IGA_HIP_TRY((f64 ? op->ops->rat64 : op->ops->rat32)(a, ebe, tile, ntiles, s));
  else
    IGA_HIP_TRY((f64 ? op->ops->apply64 : op->ops->apply32)(
        a, ebe, tile, ntiles, f64 ? op->variant64 : op->variant32, s));
  if (prof) {
    IGA_HIP_TRY(hipEventRecord(op->prof_ev[2 * op->prof_used + 1], s));
    ++op->prof_used;
  }
  // the gather's seam sums, unless the consumer of y forms them (CG fusion)
  if (gather && !(hook && hook->defer_seam)) {
    const SeamGeom g = seam_geom(op, f64);
    const unsigned nb = (unsigned)((op->nfunc + FOLD_U * 256 - 1) / (FOLD_U * 256));
    if (f64)
      hipLaunchKernelGGL(k_slot_fold<double>, dim3(nb), dim3(256), 0, s, (double *)y, g, skip);
    else
      hipLaunchKernelGGL(k_slot_fold<float>, dim3(nb), dim3(256), 0, s, (float *)y, g, skip);
    IGA_HIP_TRY(hipGetLastError());
  }
  return IGA_OK;
}

SeamGeom seam_geom(const iga_op *op, bool f64) {
  const igak::Layout &L = f64 ? op->L64 : op->L32;
  SeamGeom g{};
  g.slots = f64 ? (const void *)op->slots64 : (const void *)op->slots32;
  g.n = op->nfunc;
  g.nfx = op->nfx;
  g.ney = op->ney;
  g.p = op->p;
  g.sy = __builtin_ctz((unsigned)L.ty);
  g.inv_nfx = 1.0 / op->nfx;
  return g;
}

}  // namespace igai

namespace {
// largest |v| over n floats, as the bit pattern of a non-negative float
// (ordered like unsigned integers): atomicMax per thread
__global__ void k_absmax_bits(const float *v, int64_t n, unsigned *out) {
  float m = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(v[i]));
  if (m != m) m = __builtin_inff();
  atomicMax(out, __float_as_uint(m));
}

// The FP16 copy of G32 (IGA_F32_G16): G32 / 2^e, e chosen so that the
// largest magnitude is at most 2^14 (range margin below 65504; values 2^28
// times smaller stay normal).  op->G16 and op->g16max must exist.  Reads the
// maximum back to the host, so it synchronises the stream.
int build_g16(iga_op *op, hipStream_t s) {
  const igak::Layout &L = op->L32;
  const size_t n = (size_t)L.ntx * L.nty * op->nc * op->nqp * L.tx * L.ty;
  IGA_HIP_TRY(hipMemsetAsync(op->g16max, 0, sizeof(unsigned), s));
  hipLaunchKernelGGL(k_absmax_bits, dim3(1024), dim3(256), 0, s, op->G32, (int64_t)n,
                     op->g16max);
  unsigned hb = 0;
  IGA_HIP_TRY(hipMemcpyAsync(&hb, op->g16max, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  IGA_HIP_TRY(hipStreamSynchronize(s));
  float gmax;
  std::memcpy(&gmax, &hb, sizeof(float));
  if (!std::isfinite(gmax)) return IGA_ERR_NOT_FINITE;
  int e = 0;
  if (gmax > 0.f) {
    (void)std::frexp(gmax, &e);  // gmax < 2^e
    e -= 14;
  }
  op->g16_scale = std::ldexp(1.0f, e);
  IGA_HIP_TRY(op->ops->g16conv(op->G32, op->G16, L, std::ldexp(1.0f, -e), s));
  return IGA_OK;
}
}  // namespace

extern "C" {

const char *iga_status_string(int s) {
  switch (s) {
    case IGA_OK: return "ok";
    case IGA_ERR_INVALID_ARGUMENT: return "invalid argument";
    case IGA_ERR_INVALID_DEGREE: return "invalid degree (p, q must lie in [1,8])";
    case IGA_ERR_INVALID_QUADRATURE: return "invalid quadrature (1..10 points per direction)";
    case IGA_ERR_TOO_FEW_ELEMENTS: return "too few elements (need nelem >= p + 2)";
    case IGA_ERR_BAD_MESH: return "inconsistent mesh (IEN / ID / Dir / extraction)";
    case IGA_ERR_UNSUPPORTED: return "no kernel for this configuration";
    case IGA_ERR_HIP: return "HIP runtime error";
    case IGA_ERR_OUT_OF_MEMORY: return "out of device memory";
    case IGA_ERR_NOT_FINITE: return "non-finite value in an input vector";
    case IGA_ERR_DIVERGED: return "solver diverged";
    case IGA_ERR_NOT_CONVERGED: return "solver reached its iteration limit";
    case IGA_ERR_SIZE_LIMIT: return "request exceeds a fixed capacity";
  }
  return "unknown status";
}

void iga_op_opts_default(iga_op_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->want_f64 = 1;
  o->want_f32 = 1;
  o->variant_f64 = -1;
  o->variant_f32 = -1;
  o->stream_g_f32 = -1;
}

int iga_op_destroy(iga_op *op) {
  if (!op) return IGA_OK;
  int cur = -1;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(op->device);
  for (void *p : op->allocs) (void)hipFree(p);
  for (hipEvent_t e : op->prof_ev) (void)hipEventDestroy(e);
  if (cur >= 0) (void)hipSetDevice(cur);
  delete op;
  return IGA_OK;
}

int iga_op_create(const iga_mesh_desc *m, const iga_op_opts *o_in, iga_op **out) {
  if (!m || !out) return IGA_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  iga_op_opts o;
  if (o_in) o = *o_in; else iga_op_opts_default(&o);
  const int p = m->p;
  if (p < 1 || p > 8 || m->q < 1 || m->q > 8) return IGA_ERR_INVALID_DEGREE;
  if (m->nqp_x < 1 || m->nqp_x > 10 || m->nqp_y < 1 || m->nqp_y > 10)
    return IGA_ERR_INVALID_QUADRATURE;
  if (m->nelem_x < p + 2 || m->nelem_y < m->q + 2) return IGA_ERR_TOO_FEW_ELEMENTS;
  if (!m->cp || !m->id || !m->extraction_x || !m->extraction_y || !m->elem_size_x ||
      !m->elem_size_y || (m->ndir > 0 && !m->dir) || m->ndir < 0)
    return IGA_ERR_INVALID_ARGUMENT;
  if (m->q != p || m->nqp_x != p + 1 || m->nqp_y != p + 1) return IGA_ERR_UNSUPPORTED;
  const igak::Ops *ops = igak::ops_for_degree(p);
  if (!ops) return IGA_ERR_UNSUPPORTED;

  const int nex = m->nelem_x, ney = m->nelem_y, nfx = nex + p, nfy = ney + p;
  const int nfunc = nfx * nfy, nloc = (p + 1) * (p + 1), nq = p + 1;
  // connectivity: the tensor-product IEN of matrix_free/src/IENGenerator.cpp:68-103
  if (m->ien) {
    for (int ey = 0; ey < ney; ++ey)
      for (int ex = 0; ex < nex; ++ex) {
        const int *row = m->ien + ((size_t)ey * nex + ex) * nloc;
        for (int b = 0; b <= p; ++b)
          for (int a = 0; a <= p; ++a) {
            const int v = row[b * (p + 1) + a];
            if (v < 0 || v >= nfunc) return IGA_ERR_BAD_MESH;
            if (v != (ey + b) * nfx + ex + a) return IGA_ERR_UNSUPPORTED;
          }
      }
  }
  // equation numbers: the serial partition maps survivors to themselves
  std::vector<uint8_t> mask(nfunc, 0);
  for (int A = 0; A < nfunc; ++A) {
    const int v = m->id[A];
    if (v < -1 || v >= nfunc) return IGA_ERR_BAD_MESH;
    if (v >= 0 && v != A) return IGA_ERR_UNSUPPORTED;
    mask[A] = (v < 0);
  }
  {
    std::vector<uint8_t> seen(nfunc, 0);
    int cnt = 0;
    for (int d = 0; d < m->ndir; ++d) {
      const int A = m->dir[d];
      if (A < 0 || A >= nfunc || !mask[A]) return IGA_ERR_BAD_MESH;
      if (!seen[A]) { seen[A] = 1; ++cnt; }
    }
    int nmask = 0;
    for (int A = 0; A < nfunc; ++A) nmask += mask[A];
    if (cnt != nmask) return IGA_ERR_BAD_MESH;
  }
  bool ring = true;
  for (int j = 0; j < nfy && ring; ++j)
    for (int i = 0; i < nfx; ++i) {
      const bool r = i == 0 || j == 0 || i == nfx - 1 || j == nfy - 1;
      if (r != (mask[(size_t)j * nfx + i] != 0)) { ring = false; break; }
    }
  for (int e = 0; e < nex; ++e)
    if (!(m->elem_size_x[e] > 0.0)) return IGA_ERR_BAD_MESH;
  for (int e = 0; e < ney; ++e)
    if (!(m->elem_size_y[e] > 0.0)) return IGA_ERR_BAD_MESH;
  for (size_t k = 0; k < (size_t)2 * nfunc; ++k)
    if (!std::isfinite(m->cp[k])) return IGA_ERR_BAD_MESH;
  // weights: unit, or tensor-product w[j*nfx+i] = wx[i]*wy[j]
  std::vector<double> wx, wy;
  int rational = 0;
  if (m->weights) {
    bool unit = true;
    for (int A = 0; A < nfunc; ++A) {
      if (!(m->weights[A] > 0.0)) return IGA_ERR_BAD_MESH;
      if (m->weights[A] != 1.0) unit = false;
    }
    if (!unit) {
      wx.resize(nfx);
      wy.resize(nfy);
      const double w00 = m->weights[0];
      for (int i = 0; i < nfx; ++i) wx[i] = m->weights[i];
      for (int j = 0; j < nfy; ++j) wy[j] = m->weights[(size_t)j * nfx] / w00;
      rational = 1;
      for (int j = 0; j < nfy && rational == 1; ++j)
        for (int i = 0; i < nfx; ++i) {
          const double w = m->weights[(size_t)j * nfx + i];
          if (std::fabs(w - wx[i] * wy[j]) > 1e-13 * w) {
            rational = 2;  // not a tensor product: per-function weights
            break;
          }
        }
    }
  }

  iga_op *op = new (std::nothrow) iga_op;
  if (!op) return IGA_ERR_OUT_OF_MEMORY;
  op->device = o.device;
  op->p = p;
  op->nq = nq;
  op->nex = nex;
  op->ney = ney;
  op->nfx = nfx;
  op->nfy = nfy;
  op->nfunc = nfunc;
  op->nelem = nex * ney;
  op->nqp = nq * nq;
  op->nloc = nloc;
  op->ndir = m->ndir;
  op->rational = rational;
  op->nc = rational == 2 ? 5 : 3;
  op->ring = ring;
  op->ops = ops;
  op->want64 = o.want_f64 != 0;
  op->want32 = o.want_f32 != 0;
  // defaults measured on MI355X (quarter annulus 1024^2 / 512^2, p = 2..5):
  // one workgroup per 16x16 tile, except two threads per element on 16x8
  // tiles where the per-thread state is large -- FP32 at p >= 5 and FP64 at
  // p = 3 -- and for FP32 at p = 3 (the same kernel time as one thread per
  // element, 34.9 against 35.0 us at 1024^2, and FP32 refinement solves 2 %
  // faster, 94.7 against 96.5 ms); an explicitly chosen tile other than
  // 16x8 keeps variant 0
  auto dflt_variant = [&](bool f64) {
    const int tx = f64 ? o.tile_x_f64 : o.tile_x_f32, ty = f64 ? o.tile_y_f64 : o.tile_y_f32;
    if ((tx || ty) && !(tx == 16 && ty == 8)) return 0;
    if (f64) return (p == 3 && rational != 2) ? 3 : 0;
    return (p >= 5 || (p == 3 && rational != 2)) ? 3 : 0;
  };
  op->variant64 = o.variant_f64 < 0 ? dflt_variant(true) : o.variant_f64;
  op->variant32 = o.variant_f32 < 0 ? dflt_variant(false) : o.variant_f32;
  if (o.variant_f64 < -1 || op->variant64 > 3 || o.variant_f32 < -1 || op->variant32 > 5 ||
      o.stream_g_f32 < -1 || o.stream_g_f32 > 1) {
    delete op;
    return IGA_ERR_INVALID_ARGUMENT;
  }
  auto fail = [&](int rc) {
    iga_op_destroy(op);
    return rc;
  };
  int rc = igai::set_device(op->device);
  if (rc) return fail(rc);

  // tiles
  // the split-element variant (3) runs on 16x8 tiles, and only it does
  auto pick = [&](int tx, int ty, int dflt, int variant, int &id, igak::Layout &L) {
    if (variant == 3) dflt = igak::TILE_16x8;
    id = (tx || ty) ? igak::tile_id(tx, ty) : dflt;
    if (id < 0 || !igak::tile_compiled(p, id)) return false;
    if ((id == igak::TILE_16x8) != (variant == 3)) return false;
    igak::tile_shape(id, L.tx, L.ty);
    if (L.tx < p + 1 || L.ty < p + 1) return false;
    L.ntx = ceil_div(nex, L.tx);
    L.nty = ceil_div(ney, L.ty);
    return true;
  };
  // default tile 16x16 (16x8 for the split-element variant)
  if (!pick(o.tile_x_f64, o.tile_y_f64, igak::TILE_16x16, op->variant64, op->tile64, op->L64) ||
      !pick(o.tile_x_f32, o.tile_y_f32, igak::TILE_16x16, op->variant32, op->tile32, op->L32))
    return fail(IGA_ERR_UNSUPPORTED);
  if (rational == 2) {
    // general rational kernels: variants 0 and 3, few tile shapes
    for (int id : {op->tile64, op->tile32})
      if (!igak::rat_compiled(p, id)) return fail(IGA_ERR_UNSUPPORTED);
    if (o.variant_f64 == 1 || o.variant_f64 == 2 || o.variant_f32 == 1 || o.variant_f32 == 2)
      return fail(IGA_ERR_UNSUPPORTED);
  }
  if ((op->variant32 == 4 && (rational == 2 || !igak::mfma_compiled(p, op->tile32))) ||
      (op->variant32 == 5 && (rational == 2 || !igak::lds_compiled(p, op->tile32, false))) ||
      (op->variant64 == 1 && !igak::pf_compiled(p, op->tile64)) ||
      (op->variant32 == 1 && !igak::pf_compiled(p, op->tile32)) || op->variant64 == 2 ||
      (op->variant32 == 2 && !igak::w4_compiled(p, op->tile32, false)))
    return fail(IGA_ERR_UNSUPPORTED);

  // tables and Gauss rule (host FP64)
  std::vector<double> gx(nq), gw(nq);
  gauss01(nq, gx.data(), gw.data());
  op->h_wq = gw;
  op->h_tabx = basis_tables(p, nq, nex, m->extraction_x, rational == 1 ? wx.data() : nullptr,
                            gx.data());
  op->h_taby = basis_tables(p, nq, ney, m->extraction_y, rational == 1 ? wy.data() : nullptr,
                            gx.data());
  const size_t tabn_x = op->h_tabx.size(), tabn_y = op->h_taby.size();
  std::vector<float> tx32(op->h_tabx.begin(), op->h_tabx.end());
  std::vector<float> ty32(op->h_taby.begin(), op->h_taby.end());

  hipStream_t s = nullptr;
#define ALLOC(ptr, bytes)                                          \
  do {                                                             \
    int r_ = igai::dev_alloc(op, (void **)&(ptr), (bytes));        \
    if (r_) return fail(r_);                                       \
  } while (0)
#define H2D(dst, src, bytes)                                                   \
  do {                                                                         \
    if (hipMemcpy((dst), (src), (bytes), hipMemcpyHostToDevice) != hipSuccess) \
      return fail(IGA_ERR_HIP);                                                \
  } while (0)
  ALLOC(op->tabx64, tabn_x * 8);
  ALLOC(op->taby64, tabn_y * 8);
  ALLOC(op->tabx32, tabn_x * 4);
  ALLOC(op->taby32, tabn_y * 4);
  ALLOC(op->cp, (size_t)2 * nfunc * 8);
  ALLOC(op->wq, (size_t)nq * 8);
  H2D(op->tabx64, op->h_tabx.data(), tabn_x * 8);
  H2D(op->taby64, op->h_taby.data(), tabn_y * 8);
  H2D(op->tabx32, tx32.data(), tabn_x * 4);
  H2D(op->taby32, ty32.data(), tabn_y * 4);
  H2D(op->cp, m->cp, (size_t)2 * nfunc * 8);
  H2D(op->wq, gw.data(), (size_t)nq * 8);
  if (rational == 2) {
    std::vector<float> w32(m->weights, m->weights + nfunc);
    ALLOC(op->wts64, (size_t)nfunc * 8);
    ALLOC(op->wts32, (size_t)nfunc * 4);
    H2D(op->wts64, m->weights, (size_t)nfunc * 8);
    H2D(op->wts32, w32.data(), (size_t)nfunc * 4);
  }
  if (!ring) {
    ALLOC(op->dirmask, (size_t)nfunc);
    H2D(op->dirmask, mask.data(), (size_t)nfunc);
  }
  if (m->kappa) {
    ALLOC(op->kappa, (size_t)op->nelem * op->nqp * 8);
    H2D(op->kappa, m->kappa, (size_t)op->nelem * op->nqp * 8);
  }
  const int nqp = op->nqp;
  // G_q: FP64 always (the backward pass and the refinement residual need it)
  {
    const igak::Layout &L = op->L64;
    const size_t n = (size_t)L.ntx * L.nty * op->nc * nqp * L.tx * L.ty;
    ALLOC(op->G64, n * 8);
    if (hipSuccess != ops->geom64(op->G64, op->tabx64, op->taby64, op->cp, op->wts64, op->wq,
                                  op->kappa, nex, ney, nfx, L, s))
      return fail(IGA_ERR_HIP);
    if (op->kappa) {
      ALLOC(op->G0, n * 8);
      if (hipSuccess != ops->geom64(op->G0, op->tabx64, op->taby64, op->cp, op->wts64, op->wq,
                                    nullptr, nex, ney, nfx, L, s))
        return fail(IGA_ERR_HIP);
    } else {
      op->G0 = op->G64;
    }
    // gather slots, zero where no tile ever writes
    ALLOC(op->slots64, (size_t)3 * op->nfunc * 8);
    if (hipMemsetAsync(op->slots64, 0, (size_t)3 * op->nfunc * 8, s) != hipSuccess)
      return fail(IGA_ERR_HIP);
  }
  if (op->want32) {
    const igak::Layout &L = op->L32;
    const size_t n = (size_t)L.ntx * L.nty * op->nc * nqp * L.tx * L.ty;
    ALLOC(op->G32, n * 4);
    // streamed G_q loads once the factors exceed four times the device's
    // 256 MB last-level cache: then no part of them survives from one apply
    // to the next, and normal loads only evict x and the slots (2048^2
    // elements, p = 5 general rational, 3.0 GB: 650 against 670 us per
    // apply; at 0.4-0.9 GB normal loads are as fast or faster)
    constexpr size_t kStreamBytes = (size_t)1 << 30;
    op->stream32 = op->tile32 == igak::TILE_16x8 &&
                   (o.stream_g_f32 < 0 ? n * 4 >= kStreamBytes : o.stream_g_f32 != 0);
    if (hipSuccess != ops->geom32(op->G32, op->tabx64, op->taby64, op->cp, op->wts64, op->wq,
                                  op->kappa, nex, ney, nfx, L, s))
      return fail(IGA_ERR_HIP);
    ALLOC(op->slots32, (size_t)3 * op->nfunc * 4);
    if (hipMemsetAsync(op->slots32, 0, (size_t)3 * op->nfunc * 4, s) != hipSuccess)
      return fail(IGA_ERR_HIP);
    if (o.f16_geometry) {
      if (rational == 2 || !igak::g16_compiled(op->tile32) || !ops->apply16)
        return fail(IGA_ERR_UNSUPPORTED);
      ALLOC(op->g16max, sizeof(unsigned));
      ALLOC(op->G16, n * 2);
      const int rc16 = build_g16(op, s);
      if (rc16) return fail(rc16);
    }
  }
  ALLOC(op->tdot, (size_t)std::max(op->L64.ntx * op->L64.nty, op->L32.ntx * op->L32.nty) * 8);
  if (hipDeviceSynchronize() != hipSuccess) return fail(IGA_ERR_HIP);
#undef ALLOC
#undef H2D
  op->setup_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = op;
  return IGA_OK;
}

int iga_op_get_info(const iga_op *op, iga_op_info *info) {
  if (!op || !info) return IGA_ERR_INVALID_ARGUMENT;
  info->nfunc = op->nfunc;
  info->nfx = op->nfx;
  info->nfy = op->nfy;
  info->nelem = op->nelem;
  info->nqp = op->nqp;
  info->nloc = op->nloc;
  info->ndir = op->ndir;
  info->rational = op->rational;
  info->tile_x_f64 = op->L64.tx;
  info->tile_y_f64 = op->L64.ty;
  info->tile_x_f32 = op->L32.tx;
  info->tile_y_f32 = op->L32.ty;
  info->device_bytes = op->bytes;
  info->setup_ms = op->setup_ms;
  info->variant_f64 = op->variant64;
  info->variant_f32 = op->variant32;
  info->f16_geometry = op->G16 != nullptr;
  info->f16_scale = op->g16_scale;
  info->stream_g_f32 = op->stream32;
  return IGA_OK;
}

int iga_op_apply(iga_op *op, iga_dtype dt, iga_kernel kern, iga_assembly asmb, const void *x,
                 void *y, void *stream) {
  if (!op || !x || !y || x == y) return IGA_ERR_INVALID_ARGUMENT;
  if ((dt != IGA_F64 && dt != IGA_F32 && dt != IGA_F32_G16) ||
      (kern != IGA_KERNEL_SF && kern != IGA_KERNEL_EBE) ||
      (asmb != IGA_ASSEMBLY_IEN_ATOMIC && asmb != IGA_ASSEMBLY_INVLM_GATHER))
    return IGA_ERR_INVALID_ARGUMENT;
  if (dt == IGA_F32 && !op->G32) return IGA_ERR_UNSUPPORTED;
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  return igai::apply_raw_dt(op, dt, kern == IGA_KERNEL_EBE, asmb == IGA_ASSEMBLY_INVLM_GATHER, x,
                            y, 0, 1, (hipStream_t)stream, nullptr);
}

namespace {

// one workgroup: out = sum of the n per-tile partials (fixed order)
__global__ void k_fold_tiles(const double *part, int n, double *out) {
  __shared__ double sm[256];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v += part[i];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sm[threadIdx.x] += sm[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

}  // namespace

int iga_op_apply_dot(iga_op *op, iga_dtype dt, iga_kernel kern, iga_assembly asmb,
                     const void *x, void *y, double *xy, void *stream) {
  if (!op || !x || !y || !xy || x == y) return IGA_ERR_INVALID_ARGUMENT;
  if ((dt != IGA_F64 && dt != IGA_F32 && dt != IGA_F32_G16) ||
      (kern != IGA_KERNEL_SF && kern != IGA_KERNEL_EBE) ||
      (asmb != IGA_ASSEMBLY_IEN_ATOMIC && asmb != IGA_ASSEMBLY_INVLM_GATHER))
    return IGA_ERR_INVALID_ARGUMENT;
  const bool f64 = dt == IGA_F64, ebe = kern == IGA_KERNEL_EBE;
  if (dt == IGA_F32 && !op->G32) return IGA_ERR_UNSUPPORTED;
  if (!igai::cg_fusable_dt(op, dt, ebe)) return IGA_ERR_UNSUPPORTED;
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const igai::CgHook hook{op->tdot, nullptr, nullptr, 0};
  rc = igai::apply_raw_dt(op, dt, ebe, asmb == IGA_ASSEMBLY_INVLM_GATHER, x, y, 0, 1, s, nullptr,
                          false, &hook);
  if (rc) return rc;
  hipLaunchKernelGGL(k_fold_tiles, dim3(1), dim3(256), 0, s, op->tdot,
                     igai::apply_tiles(op, f64), xy);
  return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
}

namespace {

bool block_cfg_ok(const iga_op *op, int cfg, int k) {
  // compiled for the degree; a chunk wider than k leaves most threads idle
  if (!igak::block_cfg_compiled(op->p, cfg)) return false;
  const igak::BlockCfg c = igak::block_cfg(cfg);
  return c.kr <= 8 || c.kr <= k;  // (the MFMA form has 16-wide chunks: k >= 16)
}

int block_default(const iga_op *op, int k) {
  // without tuning: 16x16 tiles, the widest rhs chunk that k fills
  (void)op;
  return k >= 16 ? 2 : (k >= 8 ? 1 : 0);
}

int block_launch(iga_op *op, int cfg, int k, int ld, const float *X, float *Y, hipStream_t s) {
  igak::BlockArgs a;
  a.X = X;
  a.Y = Y;
  a.ld = ld;
  a.k = k;
  a.G = op->G32;
  a.tabx = op->tabx32;
  a.taby = op->taby32;
  a.dirmask = op->dirmask;
  a.nex = op->nex;
  a.ney = op->ney;
  a.nfx = op->nfx;
  a.nfy = op->nfy;
  a.L = op->L32;
  a.nbx = a.nby = a.nch = 0;
  IGA_HIP_TRY(op->ops->block(a, cfg, s));
  return IGA_OK;
}

}  // namespace

int iga_op_apply_block(iga_op *op, int k, int ld, const float *X, float *Y, void *stream) {
  if (!op || !X || !Y || k < 1 || ld < op->nfunc || X == Y) return IGA_ERR_INVALID_ARGUMENT;
  if (!op->G32 || op->rational == 2) return IGA_ERR_UNSUPPORTED;
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  return igai::apply_block_raw(op, k, ld, X, Y, (hipStream_t)stream);
}

int iga_op_tune_block(iga_op *op, int k, int ld, const float *X, float *Y, void *stream,
                      int reps, int *best_cfg, double *best_ms) {
  if (!op || !X || !Y || k < 1 || ld < op->nfunc || reps < 1) return IGA_ERR_INVALID_ARGUMENT;
  if (!op->G32 || op->rational == 2) return IGA_ERR_UNSUPPORTED;
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0, e1;
  IGA_HIP_TRY(hipEventCreate(&e0));
  IGA_HIP_TRY(hipEventCreate(&e1));
  int best = -1;
  float bms = 1e30f;
  for (int cfg = 0; cfg < igak::NBLOCKCFG; ++cfg) {
    if (!block_cfg_ok(op, cfg, k)) continue;
    rc = block_launch(op, cfg, k, ld, X, Y, s);  // warm-up
    if (rc) break;
    IGA_HIP_TRY(hipEventRecord(e0, s));
    for (int r = 0; r < reps && !rc; ++r) rc = block_launch(op, cfg, k, ld, X, Y, s);
    IGA_HIP_TRY(hipEventRecord(e1, s));
    IGA_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    IGA_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    ms /= reps;
    if (ms < bms) {
      bms = ms;
      best = cfg;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (best < 0) return IGA_ERR_UNSUPPORTED;
  bool found = false;
  for (auto &kv : op->block_tuned)
    if (kv.first == k) {
      kv.second = best;
      found = true;
    }
  if (!found) op->block_tuned.push_back({k, best});
  if (best_cfg) *best_cfg = best;
  if (best_ms) *best_ms = bms;
  return IGA_OK;
}

int iga_block_config_count(void) { return igak::NBLOCKCFG; }

int iga_block_config_desc(int cfg, int *tile_x, int *tile_y, int *rhs_chunk) {
  if (cfg < 0 || cfg >= igak::NBLOCKCFG) return IGA_ERR_INVALID_ARGUMENT;
  const igak::BlockCfg c = igak::block_cfg(cfg);
  if (tile_x) *tile_x = c.txb;
  if (tile_y) *tile_y = c.tyb;
  if (rhs_chunk) *rhs_chunk = c.kr;
  return IGA_OK;
}

int iga_block_config_matrix_core(int cfg, int *matrix_core) {
  if (cfg < 0 || cfg >= igak::NBLOCKCFG || !matrix_core) return IGA_ERR_INVALID_ARGUMENT;
  *matrix_core = igak::block_cfg(cfg).mfma;
  return IGA_OK;
}

int iga_op_set_block_config(iga_op *op, int k, int cfg) {
  if (!op || k < 1 || cfg < -1 || cfg >= igak::NBLOCKCFG) return IGA_ERR_INVALID_ARGUMENT;
  if (cfg >= 0 && !igak::block_cfg_compiled(op->p, cfg)) return IGA_ERR_UNSUPPORTED;
  for (size_t i = 0; i < op->block_tuned.size(); ++i)
    if (op->block_tuned[i].first == k) {
      if (cfg < 0)
        op->block_tuned.erase(op->block_tuned.begin() + i);
      else
        op->block_tuned[i].second = cfg;
      return IGA_OK;
    }
  if (cfg >= 0) op->block_tuned.push_back({k, cfg});
  return IGA_OK;
}

int iga_op_get_block_config(const iga_op *op, int k, int *cfg) {
  if (!op || k < 1 || !cfg) return IGA_ERR_INVALID_ARGUMENT;
  *cfg = block_default(op, k);
  for (auto &kv : op->block_tuned)
    if (kv.first == k) *cfg = kv.second;
  return IGA_OK;
}

int iga_op_time_kernel(iga_op *op, int dt, int kern, int asmb, const void *x, void *y,
                       int nlaunch, void *stream, double *avg_ms) {
  if (!op || !x || !y || x == y || nlaunch < 1 || !avg_ms) return IGA_ERR_INVALID_ARGUMENT;
  if (dt != IGA_F64 && dt != IGA_F32 && dt != IGA_F32_G16) return IGA_ERR_INVALID_ARGUMENT;
  const bool f64 = dt == IGA_F64, g16 = dt == IGA_F32_G16;
  if ((f64 && !op->G64) || (dt == IGA_F32 && !op->G32)) return IGA_ERR_UNSUPPORTED;
  if (g16 && (!op->G16 || kern == IGA_KERNEL_EBE)) return IGA_ERR_UNSUPPORTED;
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const igak::Layout &L = f64 ? op->L64 : op->L32;
  igak::ApplyArgs a = {};
  a.x = x;
  a.y = y;
  a.gscale = op->g16_scale;
  a.G = f64 ? (const void *)op->G64 : (g16 ? (const void *)op->G16 : (const void *)op->G32);
  a.tabx = f64 ? (const void *)op->tabx64 : (const void *)op->tabx32;
  a.taby = f64 ? (const void *)op->taby64 : (const void *)op->taby32;
  a.dirmask = op->dirmask;
  a.slots = f64 ? (void *)op->slots64 : (void *)op->slots32;
  a.wf = f64 ? (const void *)op->wts64 : (const void *)op->wts32;
  a.nex = op->nex;
  a.ney = op->ney;
  a.nfx = op->nfx;
  a.nfy = op->nfy;
  a.ntx = L.ntx;
  a.nty = L.nty;
  a.mode = asmb == IGA_ASSEMBLY_INVLM_GATHER ? 1 : 0;
  a.mask_in = 0;
  a.mask_out = 1;
  a.skip = nullptr;
  a.ntg = dt == IGA_F32 && op->stream32;
  const bool ebe = kern == IGA_KERNEL_EBE;
  const int ntiles = L.ntx * L.nty;
  const int tile = f64 ? op->tile64 : op->tile32;
  hipEvent_t e0, e1;
  IGA_HIP_TRY(hipEventCreate(&e0));
  IGA_HIP_TRY(hipEventCreate(&e1));
  auto launch = [&]() {
    if (g16) return op->ops->apply16(a, tile, ntiles, op->variant32, s);
    return a.wf ? (f64 ? op->ops->rat64 : op->ops->rat32)(a, ebe, tile, ntiles, s)
                : (f64 ? op->ops->apply64 : op->ops->apply32)(
                      a, ebe, tile, ntiles, f64 ? op->variant64 : op->variant32, s);
  };
  hipError_t e = launch();  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int i = 0; i < nlaunch && e == hipSuccess; ++i) e = launch();
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return IGA_ERR_HIP;
  *avg_ms = (double)ms / nlaunch;
  return IGA_OK;
}

int iga_op_profile(iga_op *op, int enable, int max_calls) {
  if (!op || (enable && max_calls < 1)) return IGA_ERR_INVALID_ARGUMENT;
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  if (!enable) {
    op->prof_on = false;
    return IGA_OK;
  }
  while ((int)op->prof_ev.size() < 2 * max_calls) {
    hipEvent_t e;
    IGA_HIP_TRY(hipEventCreate(&e));
    op->prof_ev.push_back(e);
  }
  op->prof_cap = max_calls;
  op->prof_used = 0;
  op->prof_on = true;
  return IGA_OK;
}

int iga_op_profile_read(iga_op *op, int *count, double *total_ms, double *min_ms,
                        double *max_ms) {
  if (!op || !count || !total_ms) return IGA_ERR_INVALID_ARGUMENT;
  double tot = 0.0, mn = 1e300, mx = 0.0;
  for (int i = 0; i < op->prof_used; ++i) {
    IGA_HIP_TRY(hipEventSynchronize(op->prof_ev[2 * i + 1]));
    float ms = 0.f;
    IGA_HIP_TRY(hipEventElapsedTime(&ms, op->prof_ev[2 * i], op->prof_ev[2 * i + 1]));
    tot += ms;
    mn = ms < mn ? ms : mn;
    mx = ms > mx ? ms : mx;
  }
  *count = op->prof_used;
  *total_ms = tot;
  if (min_ms) *min_ms = op->prof_used ? mn : 0.0;
  if (max_ms) *max_ms = mx;
  return IGA_OK;
}

namespace {
int load_common(iga_op *op, int kind, double value, const double *fq, double *F,
                void *stream) {
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  if (!op->fJ) {
    rc = igai::dev_alloc(op, (void **)&op->fJ, (size_t)op->nelem * op->nqp * 8);
    if (rc) return rc;
  }
  IGA_HIP_TRY(op->ops->load(F, op->fJ, op->tabx64, op->taby64, op->cp, op->wts64, op->wq,
                            op->dirmask, op->nex, op->ney, op->nfx, op->nfy, kind, value, fq,
                            (hipStream_t)stream));
  return IGA_OK;
}
}  // namespace

int iga_op_load(iga_op *op, int kind, double value, double *F, void *stream) {
  if (!op || !F || kind < 0 || kind > 2) return IGA_ERR_INVALID_ARGUMENT;
  return load_common(op, kind, value, nullptr, F, stream);
}

int iga_op_load_qp(iga_op *op, const double *fq, double *F, void *stream) {
  if (!op || !fq || !F) return IGA_ERR_INVALID_ARGUMENT;
  return load_common(op, 0, 0.0, fq, F, stream);
}

int iga_op_qp_coords(iga_op *op, double *xq, void *stream) {
  if (!op || !xq) return IGA_ERR_INVALID_ARGUMENT;
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  IGA_HIP_TRY(op->ops->qp_coords(xq, op->tabx64, op->taby64, op->cp, op->wts64, op->nex,
                                 op->ney, op->nfx, (hipStream_t)stream));
  return IGA_OK;
}

int iga_op_set_kappa(iga_op *op, const double *kappa, void *stream) {
  if (!op || !kappa) return IGA_ERR_INVALID_ARGUMENT;
  if (!op->G64) return IGA_ERR_UNSUPPORTED;  // kappa is applied to the FP64 factors
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const igak::Layout &L = op->L64;
  const size_t n = (size_t)L.ntx * L.nty * op->nc * op->nqp * L.tx * L.ty;
  if (!op->kappa) {
    rc = igai::dev_alloc(op, (void **)&op->kappa, (size_t)op->nelem * op->nqp * 8);
    if (rc) return rc;
  }
  if (op->G0 == op->G64) {
    // keep the geometry-only factors for the backward pass
    rc = igai::dev_alloc(op, (void **)&op->G0, n * 8);
    if (rc) return rc;
    IGA_HIP_TRY(hipMemcpyAsync(op->G0, op->G64, n * 8, hipMemcpyDeviceToDevice, s));
  }
  IGA_HIP_TRY(hipMemcpyAsync(op->kappa, kappa, (size_t)op->nelem * op->nqp * 8,
                             hipMemcpyDeviceToDevice, s));
  IGA_HIP_TRY(op->ops->geom64(op->G64, op->tabx64, op->taby64, op->cp, op->wts64, op->wq,
                              op->kappa, op->nex, op->ney, op->nfx, op->L64, s));
  if (op->G32)
    IGA_HIP_TRY(op->ops->geom32(op->G32, op->tabx64, op->taby64, op->cp, op->wts64, op->wq,
                                op->kappa, op->nex, op->ney, op->nfx, op->L32, s));
  ++op->gen;
  // the FP16 copy follows G32, with its power-of-two scale chosen again
  if (op->G16) return build_g16(op, s);
  return IGA_OK;
}

int iga_op_apply_backward(iga_op *op, const double *x, const double *g, double *grad_x,
                          double *grad_kappa, void *stream) {
  if (!op || !g || (!grad_x && !grad_kappa) || (grad_kappa && !x)) return IGA_ERR_INVALID_ARGUMENT;
  if (grad_x == g) return IGA_ERR_INVALID_ARGUMENT;
  int rc = igai::set_device(op->device);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (grad_x) {
    // A^T g = K~ P_int g: gather P_int g, keep every row
    rc = igai::apply_raw(op, true, false, true, g, grad_x, 1, 0, s, nullptr);
    if (rc) return rc;
  }
  if (grad_kappa)
    IGA_HIP_TRY(op->ops->kappa_grad(grad_kappa, x, g, op->G0, op->tabx64, op->taby64,
                                    op->wts64, op->dirmask, op->nex, op->ney, op->nfx, op->nfy, op->L64,
                                    s));
  return IGA_OK;
}

}  // extern "C"

namespace igai {
int apply_block_raw(iga_op *op, int k, int ld, const float *X, float *Y, hipStream_t s) {
  if (!op->G32 || op->rational == 2) return IGA_ERR_UNSUPPORTED;
  int cfg = -1;
  for (auto &kv : op->block_tuned)
    if (kv.first == k) cfg = kv.second;
  if (cfg < 0) cfg = block_default(op, k);
  return block_launch(op, cfg, k, ld, X, Y, s);
}
}  // namespace igai

