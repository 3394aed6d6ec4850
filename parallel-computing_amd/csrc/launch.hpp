// launch.hpp -- host-side launchers of the templated kernels, one set per
// degree P (instantiated in kern_p<P>.hip so the degrees compile in
// parallel).  Quadrature is the reference's p+1 Gauss points per direction.
#pragma once

#include "kernels.hpp"

namespace igak {

// element tiles with compiled kernels
enum TileId {
  TILE_32x32 = 0,
  TILE_32x16 = 1,
  TILE_16x16 = 2,
  TILE_64x8 = 3,
  TILE_16x8 = 4,  // split-element variant (3) and its EBE companion only
  NTILE = 5
};
inline void tile_shape(int id, int &tx, int &ty) {
  static const int s[NTILE][2] = {{32, 32}, {32, 16}, {16, 16}, {64, 8}, {16, 8}};
  tx = s[id][0];
  ty = s[id][1];
}
inline int tile_id(int tx, int ty) {
  for (int i = 0; i < NTILE; ++i) {
    int a, b;
    tile_shape(i, a, b);
    if (a == tx && b == ty) return i;
  }
  return -1;
}

// compiled (tile, variant) pairs: high degrees spill at 512+ threads
inline bool tile_compiled(int p, int id) {
  return p <= 4 || id == TILE_32x16 || id == TILE_16x16 || id == TILE_16x8;
}
inline bool pf_compiled(int p, int id) { return p <= 4 || id == TILE_16x16; }
inline bool w4_compiled(int p, int id, bool f64) { return p <= 4 && id == TILE_16x16 && !f64; }
inline bool g16_compiled(int id) { return id == TILE_16x16 || id == TILE_16x8; }
// matrix-core variant (4): degree 3, 16x16 tiles, FP32 arithmetic
inline bool mfma_compiled(int p, int id) { return p == 3 && id == TILE_16x16; }
// LDS-staged persistent variant (5): 16x16 tiles, FP32 arithmetic; G_q of a
// tile staged in LDS (FP32 G_q up to degree 3, FP16 G_q at degree 3)
constexpr bool lds_compiled_c(int p, bool g16) { return g16 ? p == 3 : p >= 1 && p <= 3; }
inline bool lds_compiled(int p, int id, bool g16) {
  return id == TILE_16x16 && lds_compiled_c(p, g16);
}
inline bool rat_compiled(int p, int id) {
  return id == TILE_16x16 || id == TILE_16x8 || (p <= 4 && id == TILE_32x16);
}

template <int P>
struct Launch {
  static constexpr int NQ = P + 1;
  // split-element kernels on thread-major geometry (gq_split) at three waves
  // per SIMD (the register budget they then fit without spilling)
  template <typename T>
  static constexpr int split_minw() {
    return gq_split(sizeof(T), 16, 8, NQ) ? 3 : 1;
  }

  template <typename T, int TX, int TY, bool EBE, bool RAT = false, int MINW = 1, int SPLIT = 1,
            typename TG = T, bool NTG = kStreamG<TG>>
  static hipError_t apply_t(const ApplyArgs &a, int ntiles, hipStream_t s) {
    hipLaunchKernelGGL((k_apply<T, P, NQ, TX, TY, EBE, RAT, MINW, SPLIT, TG, NTG>), dim3(ntiles),
                       dim3(TX * TY * SPLIT), 0, s, a);
    return hipGetLastError();
  }
  // matrix-core SF apply (variant 4, degree 3, 16x16 tiles) on G_q stored as TG
  template <typename TG>
  static hipError_t apply_mfma_t(const ApplyArgs &a, int ntiles, hipStream_t s) {
    if constexpr (P == 3) {
      hipLaunchKernelGGL((k_apply_mfma<TG>), dim3(ntiles), dim3(256), 0, s, a);
      return hipGetLastError();
    }
    return hipErrorInvalidValue;
  }
  // LDS-staged persistent SF apply (variant 5): as many workgroups as fit
  // at once, each walking every G-th tile
  template <typename TG>
  static hipError_t apply_lds_t(const ApplyArgs &a, int ntiles, hipStream_t s) {
    if constexpr (lds_compiled_c(P, sizeof(TG) == 2)) {
      static int resident = 0;
      if (!resident) {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per, reinterpret_cast<const void *>(&k_apply_lds<float, P, NQ, TG>), 256, 0) !=
                hipSuccess)
          return hipErrorUnknown;
        resident = cus * (per > 0 ? per : 1);
      }
      const int grid = ntiles < resident ? ntiles : resident;
      hipLaunchKernelGGL((k_apply_lds<float, P, NQ, TG>), dim3(grid), dim3(256), 0, s, a, ntiles);
      return hipGetLastError();
    }
    return hipErrorInvalidValue;
  }
  // FP32 SF apply on FP16-stored G_q (polynomial / tensor-weight meshes):
  // the 16x16 one-workgroup-per-tile kernel (its matrix-core form for
  // variant 4, its LDS-staged persistent form for variant 5), or the
  // split-element one on 16x8 tiles
  static hipError_t apply_g16(const ApplyArgs &a, int tile, int ntiles, int variant,
                              hipStream_t s) {
    if (variant == 4 && mfma_compiled(P, tile)) return apply_mfma_t<_Float16>(a, ntiles, s);
    if (variant == 5 && lds_compiled(P, tile, true)) return apply_lds_t<_Float16>(a, ntiles, s);
    if (tile == TILE_16x16) return apply_t<float, 16, 16, false, false, 1, 1, _Float16>(a, ntiles, s);
    if (tile == TILE_16x8)
      return apply_t<float, 16, 8, false, false, split_minw<float>(), 2, _Float16>(a, ntiles, s);
    return hipErrorInvalidValue;
  }
  static hipError_t g16_from_g32(const float *G32, _Float16 *G16, const Layout &L,
                                 float inv_scale, hipStream_t s) {
    const int NT = L.tx * L.ty;
    const int64_t n = (int64_t)L.ntx * L.nty * 3 * NQ * NQ * NT;
    int64_t nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL((k_g16_from_g32<NQ * NQ, 3>), dim3((unsigned)nb), dim3(256), 0, s, G32,
                       G16, n, NT, inv_scale, gq_split(4, L.tx, L.ty, NQ) ? 1 : 0);
    return hipGetLastError();
  }
  template <typename T, bool EBE>
  static hipError_t apply_tile(const ApplyArgs &a, int tile, int ntiles, hipStream_t s) {
    if constexpr (P <= 4) {
      if (tile == TILE_32x32) return apply_t<T, 32, 32, EBE>(a, ntiles, s);
      if (tile == TILE_64x8) return apply_t<T, 64, 8, EBE>(a, ntiles, s);
    }
    switch (tile) {
      case TILE_32x16: return apply_t<T, 32, 16, EBE>(a, ntiles, s);
      case TILE_16x16: return apply_t<T, 16, 16, EBE>(a, ntiles, s);
    }
    return hipErrorInvalidValue;
  }
  // persistent pipelined SF kernel: as many workgroups as fit at once
  template <typename T, int TX, int TY>
  static hipError_t apply_pf_t(const ApplyArgs &a, int ntiles, hipStream_t s) {
    static int resident = 0;
    if (!resident) {
      int dev = 0, cus = 0, per = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &per, reinterpret_cast<const void *>(&k_apply_pf<T, P, NQ, TX, TY>), TX * TY, 0) !=
              hipSuccess)
        return hipErrorUnknown;
      resident = cus * (per > 0 ? per : 1);
    }
    const int grid = ntiles < resident ? ntiles : resident;
    hipLaunchKernelGGL((k_apply_pf<T, P, NQ, TX, TY>), dim3(grid), dim3(TX * TY), 0, s, a, ntiles);
    return hipGetLastError();
  }
  template <typename T>
  static hipError_t apply_pf_tile(const ApplyArgs &a, int tile, int ntiles, hipStream_t s) {
    if constexpr (P <= 4) {
      switch (tile) {
        case TILE_32x32: return apply_pf_t<T, 32, 32>(a, ntiles, s);
        case TILE_32x16: return apply_pf_t<T, 32, 16>(a, ntiles, s);
        case TILE_64x8: return apply_pf_t<T, 64, 8>(a, ntiles, s);
      }
    }
    if (tile == TILE_16x16) return apply_pf_t<T, 16, 16>(a, ntiles, s);
    return hipErrorInvalidValue;
  }
  // variant: 0 one workgroup per tile, 1 persistent pipelined (SF only),
  // 2 one workgroup per tile at four waves per SIMD (FP32 SF, 16x16, p <= 4)
  // 3 two threads per element (SF; 16x8 tiles, which the EBE apply of such
  // an operator also uses), 4 matrix-core SF (FP32, 16x16, p = 3), 5
  // persistent with G_q staged in LDS (FP32 SF, 16x16, p <= 3).  One
  // precision per call, so that FP32 and FP64 kernels compile in separate
  // translation units.
  template <typename T>
  static hipError_t apply_prec(const ApplyArgs &a, bool ebe, int tile, int ntiles, int variant,
                               hipStream_t s) {
    if (tile == TILE_16x8) {
      if constexpr (sizeof(T) == 4)
        if (!ebe && a.ntg)
          return apply_t<T, 16, 8, false, false, split_minw<T>(), 2, T, true>(a, ntiles, s);
      return ebe ? apply_t<T, 16, 8, true>(a, ntiles, s)
                 : apply_t<T, 16, 8, false, false, split_minw<T>(), 2>(a, ntiles, s);
    }
    if constexpr (P <= 4 && sizeof(T) == 4)
      if (variant == 2 && !ebe && tile == TILE_16x16)
        return apply_t<float, 16, 16, false, false, 4>(a, ntiles, s);
    if constexpr (sizeof(T) == 4) {
      if (variant == 4 && !ebe && mfma_compiled(P, tile)) return apply_mfma_t<float>(a, ntiles, s);
      if (variant == 5 && !ebe && lds_compiled(P, tile, false))
        return apply_lds_t<float>(a, ntiles, s);
    }
    if (variant == 2 || variant >= 4) variant = 0;  // EBE applies of a variant-2 / 4 / 5 operator
    if (variant == 1 && !ebe) return apply_pf_tile<T>(a, tile, ntiles, s);
    return ebe ? apply_tile<T, true>(a, tile, ntiles, s) : apply_tile<T, false>(a, tile, ntiles, s);
  }

  template <typename T>
  static hipError_t seam(T *y, int nex, int nfx, int nfy, const Layout &L, const int *skip,
                         hipStream_t s) {
    const int n = (L.ntx - 1) * P * nfy + (L.nty - 1) * P * nfx;
    if (n == 0) return hipSuccess;
    const int bs = 256;
    const int sx = __builtin_ctz((unsigned)L.tx), sy = __builtin_ctz((unsigned)L.ty);
    hipLaunchKernelGGL((k_seam<T, P>), dim3((n + bs - 1) / bs), dim3(bs), 0, s, y, nex, nfx,
                       nfy, sx, sy, L.ntx, L.nty, skip);
    return hipGetLastError();
  }

  static hipError_t diag(double *d, const double *G, const double *tabx, const double *taby,
                         const double *wts, const uint8_t *mask, int nex, int ney, int nfx,
                         int nfy, const Layout &L, hipStream_t s) {
    const int bs = 256, nf = nfx * nfy;
    hipLaunchKernelGGL((k_diag<P, NQ>), dim3((nf + bs - 1) / bs), dim3(bs), 0, s, d, G, tabx,
                       taby, wts, mask, nex, ney, nfx, nfy, L.tx, L.ty, L.ntx);
    return hipGetLastError();
  }

  template <typename T>
  static hipError_t geometry(T *G, const double *tabx, const double *taby, const double *cp,
                             const double *wts, const double *wq, const double *kappa, int nex,
                             int ney, int nfx, const Layout &L, hipStream_t s) {
    const int nt = L.ntx * L.nty;
    int bs = L.tx * L.ty;
    if (bs > 256) bs = 256;
    hipLaunchKernelGGL((k_geometry<T, P, NQ>), dim3(nt), dim3(bs), 0, s, G, tabx, taby, cp, wts,
                       wq, kappa, nex, ney, nfx, L.tx, L.ty, L.ntx);
    return hipGetLastError();
  }

  static hipError_t load(double *F, double *fJ, const double *tabx, const double *taby,
                         const double *cp, const double *wts, const double *wq,
                         const uint8_t *mask, int nex, int ney, int nfx, int nfy, int kind,
                         double value, const double *fq, hipStream_t s) {
    const int bs = 256;
    const int ne = nex * ney, nf = nfx * nfy;
    hipLaunchKernelGGL((k_load_qp<P, NQ>), dim3((ne + bs - 1) / bs), dim3(bs), 0, s, fJ, tabx,
                       taby, cp, wts, wq, nex, ney, nfx, kind, value, fq);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_load_gather<P, NQ>), dim3((nf + bs - 1) / bs), dim3(bs), 0, s, F, fJ,
                       tabx, taby, wts, mask, nex, ney, nfx, nfy);
    return hipGetLastError();
  }

  static hipError_t qp_coords(double *xq, const double *tabx, const double *taby,
                              const double *cp, const double *wts, int nex, int ney, int nfx,
                              hipStream_t s) {
    const int bs = 256, ne = nex * ney;
    hipLaunchKernelGGL((k_qp_coords<P, NQ>), dim3((ne + bs - 1) / bs), dim3(bs), 0, s, xq, tabx,
                       taby, cp, wts, nex, ney, nfx);
    return hipGetLastError();
  }

  static hipError_t kappa_grad(double *out, const double *x, const double *g, const double *G0,
                               const double *tabx, const double *taby, const double *wts,
                               const uint8_t *mask, int nex, int ney, int nfx, int nfy,
                               const Layout &L, hipStream_t s) {
    const int bs = 256, ne = nex * ney;
    hipLaunchKernelGGL((k_kappa_grad<P, NQ>), dim3((ne + bs - 1) / bs), dim3(bs), 0, s, out, x,
                       g, G0, tabx, taby, wts, mask, nex, ney, nfx, nfy, L.tx, L.ty, L.ntx);
    return hipGetLastError();
  }
};

// General rational meshes (per-function weights): one workgroup per tile,
// 16x16 tiles (and 32x16 for p <= 4); instantiated in kern_p<P>_rat.hip.
template <int P>
struct LaunchRat {
  // the FP32 split-element kernel on thread-major geometry (gq_split: one
  // load per row and component) at three waves per SIMD: it then fits 168
  // VGPRs without spilling (182 VGPRs, two waves per SIMD, unbounded)
  template <typename T, int TX, int TY, bool EBE, int SPLIT = 1>
  static hipError_t run_t(const ApplyArgs &a, int ntiles, hipStream_t s) {
    constexpr int MINW = (SPLIT == 2 && gq_split(sizeof(T), TX, TY, P + 1)) ? 3 : 1;
    if constexpr (SPLIT == 2 && sizeof(T) == 4)
      if (a.ntg) {
        hipLaunchKernelGGL((k_apply<T, P, P + 1, TX, TY, EBE, true, MINW, SPLIT, T, true>),
                           dim3(ntiles), dim3(TX * TY * SPLIT), 0, s, a);
        return hipGetLastError();
      }
    hipLaunchKernelGGL((k_apply<T, P, P + 1, TX, TY, EBE, true, MINW, SPLIT>), dim3(ntiles),
                       dim3(TX * TY * SPLIT), 0, s, a);
    return hipGetLastError();
  }
  template <typename T, bool EBE>
  static hipError_t run_tile(const ApplyArgs &a, int tile, int ntiles, hipStream_t s) {
    if constexpr (P <= 4)
      if (tile == TILE_32x16) return run_t<T, 32, 16, EBE>(a, ntiles, s);
    if (tile == TILE_16x16) return run_t<T, 16, 16, EBE>(a, ntiles, s);
    if (tile == TILE_16x8) return EBE ? run_t<T, 16, 8, true>(a, ntiles, s)
                                      : run_t<T, 16, 8, false, 2>(a, ntiles, s);
    return hipErrorInvalidValue;
  }
  template <typename T>
  static hipError_t apply_prec(const ApplyArgs &a, bool ebe, int tile, int ntiles,
                               hipStream_t s) {
    return ebe ? run_tile<T, true>(a, tile, ntiles, s) : run_tile<T, false>(a, tile, ntiles, s);
  }
};

struct BlockArgs;

// per-degree entry points (kern_p<P>.hip, kern_p<P>_rat.hip, kern_p<P>_block.hip)
struct Ops {
  hipError_t (*apply64)(const ApplyArgs &, bool ebe, int tile, int ntiles, int variant,
                        hipStream_t);
  hipError_t (*apply32)(const ApplyArgs &, bool ebe, int tile, int ntiles, int variant,
                        hipStream_t);
  // zero the seam functions before a scatter-add apply
  hipError_t (*seam64)(double *, int, int, int, const Layout &, const int *, hipStream_t);
  hipError_t (*seam32)(float *, int, int, int, const Layout &, const int *, hipStream_t);
  hipError_t (*diag)(double *, const double *, const double *, const double *, const double *,
                     const uint8_t *, int, int, int, int, const Layout &, hipStream_t);
  hipError_t (*geom64)(double *, const double *, const double *, const double *,
                       const double *, const double *, const double *, int, int, int,
                       const Layout &, hipStream_t);
  hipError_t (*geom32)(float *, const double *, const double *, const double *,
                       const double *, const double *, const double *, int, int, int,
                       const Layout &, hipStream_t);
  hipError_t (*load)(double *, double *, const double *, const double *, const double *,
                     const double *, const double *, const uint8_t *, int, int, int, int, int,
                     double, const double *, hipStream_t);
  hipError_t (*kappa_grad)(double *, const double *, const double *, const double *,
                           const double *, const double *, const double *, const uint8_t *, int,
                           int, int, int, const Layout &, hipStream_t);
  hipError_t (*qp_coords)(double *, const double *, const double *, const double *,
                          const double *, int, int, int, hipStream_t);
  hipError_t (*rat64)(const ApplyArgs &, bool ebe, int tile, int ntiles, hipStream_t);
  hipError_t (*rat32)(const ApplyArgs &, bool ebe, int tile, int ntiles, hipStream_t);
  hipError_t (*block)(const BlockArgs &, int cfg, hipStream_t);
  hipError_t (*apply16)(const ApplyArgs &, int tile, int ntiles, int variant, hipStream_t);
  hipError_t (*g16conv)(const float *, _Float16 *, const Layout &, float, hipStream_t);
};

template <int P>
Ops make_ops() {
  Ops o;
  o.apply32 = &Launch<P>::template apply_prec<float>;
  o.apply64 = nullptr;
  o.seam64 = &Launch<P>::template seam<double>;
  o.seam32 = &Launch<P>::template seam<float>;
  o.geom64 = &Launch<P>::template geometry<double>;
  o.geom32 = &Launch<P>::template geometry<float>;
  o.load = &Launch<P>::load;
  o.diag = &Launch<P>::diag;
  o.kappa_grad = &Launch<P>::kappa_grad;
  o.qp_coords = &Launch<P>::qp_coords;
  o.rat64 = o.rat32 = nullptr;
  o.block = nullptr;
  o.apply16 = nullptr;
  o.g16conv = nullptr;
  return o;
}

// ops_for_degree(P) returns nullptr for degrees without compiled kernels
const Ops *ops_for_degree(int p);

}  // namespace igak

#define IGAK_DEFINE_DEGREE(P)                                                             \
  namespace igak {                                                                        \
  hipError_t apply64_p##P(const ApplyArgs &, bool, int, int, int, hipStream_t);           \
  hipError_t rat64_p##P(const ApplyArgs &, bool, int, int, hipStream_t);                  \
  hipError_t rat32_p##P(const ApplyArgs &, bool, int, int, hipStream_t);                  \
  hipError_t block_p##P(const BlockArgs &, int, hipStream_t);                             \
  hipError_t apply16_p##P(const ApplyArgs &, int, int, int, hipStream_t);                 \
  hipError_t g16conv_p##P(const float *, _Float16 *, const Layout &, float, hipStream_t); \
  const Ops *ops_p##P() {                                                                 \
    static const Ops o = [] {                                                             \
      Ops r = make_ops<P>();                                                              \
      r.apply64 = &apply64_p##P;                                                          \
      r.rat64 = &rat64_p##P;                                                              \
      r.rat32 = &rat32_p##P;                                                              \
      r.block = &block_p##P;                                                              \
      r.apply16 = &apply16_p##P;                                                          \
      r.g16conv = &g16conv_p##P;                                                          \
      return r;                                                                           \
    }();                                                                                  \
    return &o;                                                                            \
  }                                                                                       \
  }

// FP64 apply kernels of degree P (kern_p<P>_f64.hip)
#define IGAK_DEFINE_F64(P)                                                                \
  namespace igak {                                                                        \
  hipError_t apply64_p##P(const ApplyArgs &a, bool ebe, int tile, int ntiles, int variant, \
                          hipStream_t s) {                                                \
    return Launch<P>::template apply_prec<double>(a, ebe, tile, ntiles, variant, s);      \
  }                                                                                       \
  }

// FP32 apply kernels on FP16-stored G_q of degree P (kern_p<P>_g16.hip)
#define IGAK_DEFINE_G16(P)                                                                \
  namespace igak {                                                                        \
  hipError_t apply16_p##P(const ApplyArgs &a, int tile, int ntiles, int variant,          \
                          hipStream_t s) {                                                \
    return Launch<P>::apply_g16(a, tile, ntiles, variant, s);                             \
  }                                                                                       \
  hipError_t g16conv_p##P(const float *G32, _Float16 *G16, const Layout &L, float inv,    \
                          hipStream_t s) {                                                \
    return Launch<P>::g16_from_g32(G32, G16, L, inv, s);                                  \
  }                                                                                       \
  }

// general rational apply kernels of degree P, one precision per file
#define IGAK_DEFINE_RAT(P)                                                                \
  namespace igak {                                                                        \
  hipError_t rat32_p##P(const ApplyArgs &a, bool ebe, int tile, int ntiles, hipStream_t s) { \
    return LaunchRat<P>::template apply_prec<float>(a, ebe, tile, ntiles, s);             \
  }                                                                                       \
  }
#define IGAK_DEFINE_RAT64(P)                                                              \
  namespace igak {                                                                        \
  hipError_t rat64_p##P(const ApplyArgs &a, bool ebe, int tile, int ntiles, hipStream_t s) { \
    return LaunchRat<P>::template apply_prec<double>(a, ebe, tile, ntiles, s);            \
  }                                                                                       \
  }
