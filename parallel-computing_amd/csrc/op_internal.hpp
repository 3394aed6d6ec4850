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
  _Float16 *G16 = nullptr;          // FP16 copy of G32 / g16_scale (f16_geometry)
  float g16_scale = 1.0f;
  double *kappa = nullptr;          // device [E][nqp] or nullptr
  double *halo64 = nullptr;
  float *halo32 = nullptr;
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
int dev_alloc(iga_op *op, void **p, size_t bytes);
int set_device(int dev);
// CG fusion of an apply (see igak::ApplyArgs): per-tile partials of x.y into
// dot[ntiles], and the hand-off copy of hn doubles hsrc -> hdst
struct CgHook {
  double *dot;
  const double *hsrc;
  double *hdst;
  int hn;
};
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
