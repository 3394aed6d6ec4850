// block.hpp -- multi right-hand-side apply Y[:,j] = A X[:,j], FP32.
//
// A workgroup owns a 16 x TY element tile and a chunk of RC right-hand
// sides; thread = element.  The element's G_q are loaded into registers
// once and its 1D tables staged in LDS once, then the thread runs the
// sum-factorised element operator for each of the RC columns in turn, so
// the operator bytes (the bulk of a single apply's traffic) are read once
// per chunk instead of once per column.  The X footprints of the chunk are
// staged in LDS up front; each column's element vectors are overlap-added
// with the DPP row reduction of the single-vector kernel (double-buffered,
// one barrier per column) and stored: functions of this tile only directly,
// seam functions by atomic add after k_seam_block zeroed them.
#pragma once

#include "kernels.hpp"

namespace igak {

struct BlockArgs {
  const float *X;   // column-major, leading dimension ld
  float *Y;
  int ld, k;
  const float *G;   // single-vector FP32 layout (Layout L)
  const float *tabx, *taby;
  const uint8_t *dirmask;
  int nex, ney, nfx, nfy;
  Layout L;         // G layout
  int nbx, nby;     // block tiles per direction
  int nch;          // rhs chunks of RC columns
};

template <int P, int NQ, int TY, int RC>
__global__ __launch_bounds__(16 * TY) void k_apply_block(BlockArgs A) {
  constexpr int TX = 16, NB = P + 1, NT = TX * TY;
  constexpr int FX = TX + P, FY = TY + P, NFP = FX * FY;
  constexpr int NQ2 = NQ * NQ, TAB = NQ * NB;
  __shared__ float xs[RC * NFP];
  __shared__ float tbx[2 * TAB][TX];
  __shared__ float tby[2 * TAB][TY];
  __shared__ float red[2][row16_lds<P, TY>()];

  const int t = threadIdx.x, lx = t % TX, ly = t / TX;
  // the rhs chunks of one tile are consecutive logical ids, which xcd_tile
  // keeps on one XCD, so they share the tile's G_q through that XCD's L2
  const int lid = xcd_tile(blockIdx.x, gridDim.x);
  const int bt = lid / A.nch;
  const int bx = bt % A.nbx, by = bt / A.nbx;
  const int ex0 = bx * TX, ey0 = by * TY;
  const int j0 = (lid % A.nch) * RC;
  const int ex = ex0 + lx, ey = ey0 + ly;
  const bool valid = ex < A.nex && ey < A.ney;

  // G_q of this element into registers
  float g[3][NQ2];
  if (A.L.tx == TX && A.L.ty == TY) {
    load_gq<float, NQ2>(A.G, by * A.L.ntx + bx, t, NT, g);
  } else {
    const int exc = valid ? ex : 0, eyc = valid ? ey : 0;
    const int tile = (eyc / A.L.ty) * A.L.ntx + exc / A.L.tx;
    const int tt = (eyc % A.L.ty) * A.L.tx + exc % A.L.tx;
    const int NTL = A.L.tx * A.L.ty;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < NQ2; ++q) g[c][q] = A.G[gq_index<float, NQ2>(tile, c, q, tt, NTL)];
  }
  // X footprints of the chunk ([rhs][function]) and the 1D tables
  for (int i = t; i < RC * NFP; i += NT) {
    const int jr = i / NFP, kk = i % NFP;
    const int gx = ex0 + kk % FX, gy = ey0 + kk / FX;
    float v = 0.f;
    if (j0 + jr < A.k && gx < A.nfx && gy < A.nfy)
      v = A.X[(size_t)(j0 + jr) * A.ld + (size_t)gy * A.nfx + gx];
    xs[i] = v;
  }
  for (int i = t; i < 2 * TAB * TX; i += NT) {
    const int l = i % TX, c = i / TX;
    int e = ex0 + l;
    if (e > A.nex - 1) e = A.nex - 1;
    tbx[c][l] = A.tabx[(size_t)e * 2 * TAB + c];
  }
  for (int i = t; i < 2 * TAB * TY; i += NT) {
    const int l = i % TY, c = i / TY;
    int e = ey0 + l;
    if (e > A.ney - 1) e = A.ney - 1;
    tby[c][l] = A.taby[(size_t)e * 2 * TAB + c];
  }
  __syncthreads();

  ApplyArgs oa = {};  // output addressing shared with the single-vector store
  oa.dirmask = A.dirmask;
  oa.nex = A.nex;
  oa.ney = A.ney;
  oa.nfx = A.nfx;
  oa.nfy = A.nfy;
  oa.mode = 0;
  oa.mask_out = 1;
  auto gl = [&](int c, int q, int, int) { return g[c][q]; };
  // block tile with neighbours on all sides, ring Dirichlet set (see put_footprint)
  const bool interior = bx > 0 && by > 0 && bx < A.nbx - 1 && by < A.nby - 1 && !A.dirmask;
  const int nr = (A.k - j0) < RC ? (A.k - j0) : RC;
  for (int r = 0; r < nr; ++r) {
    float ye[NB][NB];
    {
      float X[NB][NB];
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int a = 0; a < NB; ++a) X[b][a] = xs[r * NFP + (ly + b) * FX + lx + a];
      element_sf<float, P, NQ, TX, TY>(X, tbx, tby, lx, ly, gl, ye);
      if (!valid) {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int a = 0; a < NB; ++a) ye[b][a] = 0.f;
      }
    }
    double dnone = 0.0;
    reduce_store_row16<float, P, TY, false>(oa, ye, red[r & 1], t, 0, ex0, ey0,
                                            A.Y + (size_t)(j0 + r) * A.ld, nullptr, nullptr,
                                            xs, dnone, interior);
  }
}

// zero the seam functions of the block tiling for all k columns
template <int P>
__global__ void k_seam_block(float *Y, int ld, int k, int nex, int ney, int nfx, int nfy,
                             int TX, int TY, int ntx, int nty) {
  const int64_t nv = (int64_t)(ntx - 1) * P * nfy;
  const int64_t nh = (int64_t)(nty - 1) * P * nfx;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv + nh) return;
  int gx, gy;
  if (i < nv) {
    const int band = (int)(i / ((int64_t)P * nfy));
    const int r = (int)(i % ((int64_t)P * nfy));
    gx = (band + 1) * TX + r % P;
    gy = r / P;
  } else {
    const int64_t jj = i - nv;
    const int band = (int)(jj / ((int64_t)P * nfx));
    const int r = (int)(jj % ((int64_t)P * nfx));
    gy = (band + 1) * TY + r / nfx;
    gx = r % nfx;
    int xlo, xhi;
    touching_tiles<P>(gx, nex, TX, xlo, xhi);
    if (xlo != xhi) return;
  }
  const size_t Ag = (size_t)gy * nfx + gx;
  for (int j = 0; j < k; ++j) Y[(size_t)j * ld + Ag] = 0.f;
}

// compiled block configurations (tile 16 x TY elements, RC rhs per workgroup)
struct BlockCfg {
  int txb, tyb, kr;
};
constexpr int NBLOCKCFG = 6;
inline BlockCfg block_cfg(int i) {
  static const BlockCfg c[NBLOCKCFG] = {{16, 16, 4}, {16, 16, 8}, {16, 16, 16},
                                        {16, 8, 4},  {16, 8, 8},  {16, 8, 16}};
  return c[i];
}
// configurations compiled for degree p (all: 128 or 256 threads)
inline bool block_cfg_compiled(int p, int i) {
  (void)p;
  return i >= 0 && i < NBLOCKCFG;
}

template <int P>
struct LaunchBlock {
  static constexpr int NQ = P + 1;
  template <int TY, int RC>
  static hipError_t run_t(BlockArgs a, hipStream_t s) {
    constexpr int TX = 16;
    a.nbx = (a.nex + TX - 1) / TX;
    a.nby = (a.ney + TY - 1) / TY;
    a.nch = (a.k + RC - 1) / RC;
    const int64_t n = (int64_t)(a.nbx - 1) * P * a.nfy + (int64_t)(a.nby - 1) * P * a.nfx;
    if (n > 0)
      hipLaunchKernelGGL((k_seam_block<P>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                         a.Y, a.ld, a.k, a.nex, a.ney, a.nfx, a.nfy, TX, TY, a.nbx, a.nby);
    hipLaunchKernelGGL((k_apply_block<P, NQ, TY, RC>), dim3(a.nbx * a.nby * a.nch),
                       dim3(TX * TY), 0, s, a);
    return hipGetLastError();
  }
  static hipError_t run(const BlockArgs &a, int cfg, hipStream_t s) {
    switch (cfg) {
      case 0: return run_t<16, 4>(a, s);
      case 1: return run_t<16, 8>(a, s);
      case 2: return run_t<16, 16>(a, s);
      case 3: return run_t<8, 4>(a, s);
      case 4: return run_t<8, 8>(a, s);
      case 5: return run_t<8, 16>(a, s);
    }
    return hipErrorInvalidValue;
  }
};

#define IGAK_DEFINE_BLOCK(P)                                              \
  namespace igak {                                                        \
  hipError_t block_p##P(const BlockArgs &a, int cfg, hipStream_t s) {     \
    return LaunchBlock<P>::run(a, cfg, s);                                \
  }                                                                       \
  }

}  // namespace igak
