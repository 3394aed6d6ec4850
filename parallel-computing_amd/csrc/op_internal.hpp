// op_internal.hpp -- the operator's device state, shared by the translation
// units of the library (apply, solver, block apply).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "iga_mf.h"
#include "launch.hpp"

#define IGA_HIP_TRY(expr)                                  \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) {                                \
      (void)hipGetLastError();                             \
      return e_ == hipErrorOutOfMemory ? IGA_ERR_OUT_OF_MEMORY : IGA_ERR_HIP; \
    }                                                      \
  } while (0)

struct iga_op {
  // direction of the next apply's tile sweep (alternates: see k_apply)
  int rev_next = 0;
  int device = 0;
  int p = 0, nq = 0, nex = 0, ney = 0, nfx = 0, nfy = 0;
  int nfunc = 0, nelem = 0, nqp = 0, nloc = 0, ndir = 0;
  int rational = 0;                 // 0 polynomial, 1 tensor-product weights, 2 general
  int nc = 3;                       // G_q components (5 on general rational meshes)
  bool ring = true;                 // Dirichlet set is the boundary ring
  const igak::Ops *ops = nullptr;
  uint8_t *dirmask = nullptr;       // device, only when !ring
  double *tabx64 = nullptr, *taby64 = nullptr;
  float *tabx32 = nullptr, *taby32 = nullptr;
  double *cp = nullptr;             // device [nfunc][2]
  double *wts64 = nullptr;          // per-function weights, general rational meshes only
  float *wts32 = nullptr;
  double *wq = nullptr;             // device Gauss weights on [0,1]
  igak::Layout L64{}, L32{};
  int tile64 = 0, tile32 = 0;       // igak::TileId
  int variant64 = 0, variant32 = 0; // 0 per-tile workgroups, 1 persistent pipelined
  double *G64 = nullptr;            // kappa-weighted
  double *G0 = nullptr;             // geometry only (== G64 when kappa unset)
  float *G32 = nullptr;
  // the FP32 split-element SF kernels load G32 non-temporally (their
  // streaming instantiation; see iga_op_opts.stream_g_f32)
  bool stream32 = false;
  _Float16 *G16 = nullptr;          // FP16 copy of G32 / g16_scale (f16_geometry)
  float g16_scale = 1.0f;
  unsigned *g16max = nullptr;       // device word: bits of max |G32| (FP16 scale)
  double *kappa = nullptr;          // device [E][nqp] or nullptr
  // gather assembly: seam partials of the x-, y- and diagonal neighbour tile,
  // [3][nfunc] per precision, zero wherever no tile writes (see k_apply)
  double *slots64 = nullptr;
  float *slots32 = nullptr;
  double *fJ = nullptr;             // load-vector scratch [E][nqp]
  double *tdot = nullptr;           // per-tile x.y partials (iga_op_apply_dot)
  bool want64 = true, want32 = true;
  int gen = 0;                      // bumped whenever G_q changes (set_kappa)
  int64_t bytes = 0;
  double setup_ms = 0.0;
  std::vector<void *> allocs;
  // block apply: tuned configuration per number of right-hand sides
  std::vector<std::pair<int, int>> block_tuned;  // (k, cfg)
  // k_apply timing (iga_op_profile)
  bool prof_on = false;
  int prof_cap = 0, prof_used = 0;
  std::vector<hipEvent_t> prof_ev;
  // host copies used by the solver set-up
  std::vector<double> h_tabx, h_taby, h_wq;
};

namespace igai {
// geometric multigrid on the Q1 FEM matrix (mg.hip): rowptr / col / val a
// host copy of the nx x ny grid's CSR in natural numbering
struct MgHier;
int mg_create(const int *rowptr, const int *col, const double *val, int nx, int ny, bool want64,
              bool want32, MgHier **out);
void mg_destroy(MgHier *h);
int64_t mg_bytes(const MgHier *h);
int mg_levels(const MgHier *h);
// z = one V-cycle applied to r (an approximation of K^-1 r); every kernel
// returns at once while *s1 or *s2 is set
template <typename V>
int mg_apply(MgHier *h, const V *r, V *z, hipStream_t st, const int *s1, const int *s2);

int dev_alloc(iga_op *op, void **p, size_t bytes);
int set_device(int dev);
// CG fusion of an apply (see igak::ApplyArgs): per-tile partials of x.y into
// dot[ntiles], and the hand-off copy of hn doubles hsrc -> hdst
// defer_seam (gather assembly): seam partials into function slots, summed
// by the consumer of y (seam_load, seam_value)
struct CgHook {
  double *dot;
  const double *hsrc;
  double *hdst;
  int hn;
  // gather assembly: leave the seam partials in the operator's slot arrays
  // for the consumer of y (seam_load, seam_value), with no seam pass
  bool defer_seam;
};
// What a consumer of a defer_seam gather apply needs to sum the seam
// functions itself: the slot arrays and the tile rows of that precision.
struct SeamGeom {
  const void *slots;  // [3][n]: x-, y- and diagonal neighbour partials
  int64_t n;          // functions
  int nfx, ney, p;
  int sy;             // tiles of 2^sy element rows
  double inv_nfx;
};
SeamGeom seam_geom(const iga_op *op, bool f64);

// The further partials of function i (see seam_load)
template <typename V>
struct SeamPart {
  V a1, a2, a3;
};

// Loads of y at function i after a defer_seam gather apply: y[i] (the whole
// value, or the lowest touching tile's partial) and the three slot partials,
// which are zero where the function has no such neighbour tile.  All loads
// are coalesced; the y- and diagonal-neighbour slots are loaded only by a
// wave holding a function on a horizontal tile seam (wave vote: most waves
// of a row-major sweep hold none, and their slots are zero).
template <typename V>
__device__ __forceinline__ V seam_load(const V *__restrict__ y, const SeamGeom &g, int64_t i,
                                       SeamPart<V> &s) {
  const V *S = static_cast<const V *>(g.slots);
  int gy = (int)((double)i * g.inv_nfx);  // within one of i / nfx
  const int gx = (int)(i - (int64_t)gy * g.nfx);
  gy += gx >= g.nfx ? 1 : (gx < 0 ? -1 : 0);
  const bool shy = (max(gy - g.p, 0) >> g.sy) != (min(gy, g.ney - 1) >> g.sy);
  const V a0 = y[i];
  s.a1 = S[i];
  s.a2 = s.a3 = V(0);
  if (__any(shy)) {
    s.a2 = S[g.n + i];
    s.a3 = S[2 * g.n + i];
  }
  return a0;
}

// y at function i from what seam_load returned: the sum in k_seam's order
// (lowest tile, x-, y-, diagonal neighbour; an absent partial adds +0)
template <typename V>
__device__ __forceinline__ V seam_value(V a0, const SeamPart<V> &s) {
  return ((a0 + s.a1) + s.a2) + s.a3;
}
// whether the apply kernel configured for this precision supports CgHook
bool cg_fusable(const iga_op *op, bool f64, bool ebe);
bool cg_fusable_dt(const iga_op *op, int dt, bool ebe);
// tiles of the apply layout of this precision (entries of CgHook::dot)
int apply_tiles(const iga_op *op, bool f64);

// y = A x with an explicit precision (used by the solver).  y_zeroed: y is
// already zero (the scatter-add form then skips its seam-zeroing kernel).
int apply_raw(iga_op *op, bool f64, bool ebe, bool gather, const void *x, void *y,
              int mask_in, int mask_out, hipStream_t s, const int *skip,
              bool y_zeroed = false, const CgHook *hook = nullptr);
// the same with an iga_dtype (IGA_F32_G16: FP32 arithmetic on the FP16 G_q)
int apply_raw_dt(iga_op *op, int dt, bool ebe, bool gather, const void *x, void *y,
                 int mask_in, int mask_out, hipStream_t s, const int *skip,
                 bool y_zeroed = false, const CgHook *hook = nullptr);
// Y = A X for k FP32 columns (leading dimension ld), the configuration
// tuned or defaulted for k; no device switch (callable under graph capture)
int apply_block_raw(iga_op *op, int k, int ld, const float *X, float *Y, hipStream_t s);
}  // namespace igai
