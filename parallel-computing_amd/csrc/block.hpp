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
  const bool split = gq_split(4, A.L.tx, A.L.ty, NQ);  // the operator's storage order
  if (A.L.tx == TX && A.L.ty == TY && !split) {
    load_gq<float, NQ2>(A.G, by * A.L.ntx + bx, t, NT, g);
  } else {
    const int exc = valid ? ex : 0, eyc = valid ? ey : 0;
    const int tile = (eyc / A.L.ty) * A.L.ntx + exc / A.L.tx;
    const int tt = (eyc % A.L.ty) * A.L.tx + exc % A.L.tx;
    const int NTL = A.L.tx * A.L.ty;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < NQ2; ++q) g[c][q] = A.G[gq_at<float, NQ>(tile, c, q, tt, NTL, split)];
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
    reduce_store_row16<float, P, TY, false>(oa, ye, red[r & 1], t, ex0, ey0,
                                            A.Y + (size_t)(j0 + r) * A.ld, nullptr, xs, dnone,
                                            interior);
  }
}

// Matrix-core form of the same apply.  A wave takes one element and the 16
// right-hand sides of the chunk and forms the element operator as two dense
// products on v_mfma_f32_16x16x4f32 (FP32 in, FP32 accumulate), with the
// right-hand sides as the 16 columns of every product:
//   forward   U[m][r] = sum_c K[m][c] X[c][r]        m = 2 q + d (50 x 25 at p = 4)
//   flux      F[2q][r] = -(G11 U[2q] + G12 U[2q+1]), F[2q+1][r] = -(G12 U[2q] + G22 U[2q+1])
//   backward  Y[c][r] = sum_m K[m][c] F[m][r]
// where K[2q][c] = By[j][b] Dx[i][a] and K[2q+1][c] = Dy[j][b] Bx[i][a]
// (q = j NQ + i, c = b NB + a; B, D the element's 1D tables, read from LDS):
// the Kronecker-product form of the sum-factorised operator, (p+1)^2 / 2
// times its multiply-adds (2.3x at p = 4), in exchange for operand shapes
// that fill the 16 x 16 x 4 tiles (78 % of rows and 89 % of the K steps
// used at p = 4, against 31-62 % for the sum-factorised stages, whose
// extents are p + 1).  Operand layouts of the instruction: A[i][k] in lane
// 16 k + i, B[k][j] in lane 16 k + j, D[4 g + v][j] in register v of lane
// 16 g + j.  The element results are added into an LDS footprint of the
// tile with LDS atomics, then stored like the VALU kernel's.
template <int P, int NQ, int TY>
__global__ __launch_bounds__(256, 2) void k_apply_block_mfma(BlockArgs A) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  constexpr int TX = 16, NB = P + 1, NT = 256, NW = NT / 64, RC = 16;
  constexpr int FX = TX + P, FY = TY + P, NFP = FX * FY;
  constexpr int TAB = NQ * NB, NQ2 = NQ * NQ, NC = NB * NB, NM = 2 * NQ2;
  constexpr int MT = (NM + 15) / 16;  // 16-row tiles of the forward product
  constexpr int KS = (NC + 3) / 4;    // its K steps
  constexpr int CT = (NC + 15) / 16;  // 16-row tiles of the backward product
  constexpr int KS2 = (NM + 3) / 4;   // its K steps
  // LDS rows padded to an odd number of words: the 16 right-hand sides of a
  // column (lanes 16 g + r) and the table rows of one element fall in
  // different banks
  constexpr int NFPP = NFP + 1, TABP = 2 * TAB + 1;
  __shared__ float xs[RC][NFPP];
  __shared__ float ys[RC][NFPP];
  __shared__ float tbx[TX][TABP];  // [element column][table row]
  __shared__ float tby[TY][TABP];
  __shared__ float fb[NW][MT * 16][RC + 1];

  const int t = threadIdx.x, lane = t & 63, g = lane >> 4, rr = lane & 15;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int lid = xcd_tile(blockIdx.x, gridDim.x);
  const int bt = lid / A.nch;
  const int bx = bt % A.nbx, by = bt / A.nbx;
  const int ex0 = bx * TX, ey0 = by * TY;
  const int j0 = (lid % A.nch) * RC;

  for (int i = t; i < RC * NFP; i += NT) {
    const int jr = i / NFP, kk = i % NFP;
    const int gx = ex0 + kk % FX, gy = ey0 + kk / FX;
    float v = 0.f;
    if (j0 + jr < A.k && gx < A.nfx && gy < A.nfy)
      v = A.X[(size_t)(j0 + jr) * A.ld + (size_t)gy * A.nfx + gx];
    xs[jr][kk] = v;
    ys[jr][kk] = 0.f;
  }
  for (int i = t; i < 2 * TAB * TX; i += NT) {
    const int l = i / (2 * TAB), c = i % (2 * TAB);
    int e = ex0 + l;
    if (e > A.nex - 1) e = A.nex - 1;
    tbx[l][c] = A.tabx[(size_t)e * 2 * TAB + c];
  }
  for (int i = t; i < 2 * TAB * TY; i += NT) {
    const int l = i / (2 * TAB), c = i % (2 * TAB);
    int e = ey0 + l;
    if (e > A.ney - 1) e = A.ney - 1;
    tby[l][c] = A.taby[(size_t)e * 2 * TAB + c];
  }
  __syncthreads();

  // entry (m, c) of K: the y-table row (its factor is per element row), the
  // x-table row, or -1 outside the 2 NQ2 x NC block
  auto yrow = [](int m, int c) {
    if (m >= NM || c >= NC) return -1;
    const int q = m >> 1, d = m & 1, j = q / NQ, b = c / NB;
    return (d ? TAB : 0) + j * NB + b;
  };
  auto xrow = [](int m, int c) {
    if (m >= NM || c >= NC) return 0;
    const int q = m >> 1, d = m & 1, i = q % NQ, a = c % NB;
    return (d ? 0 : TAB) + i * NB + a;
  };
  // footprint offset of coefficient c of the element at (0, 0)
  auto coff = [](int c) { return (c / NB) * FX + c % NB; };

  const int NTL = A.L.tx * A.L.ty;
  const bool split = gq_split(4, A.L.tx, A.L.ty, NQ);  // the operator's storage order
  for (int ly = w; ly < TY; ly += NW) {
    const int ey = ey0 + ly;
    if (ey >= A.ney) break;
    // y-table factors of this element row: forward A[m = 16 s1 + rr][c = 4 s + g],
    // backward A[c = 16 s1 + rr][m = 4 s + g]
    float ryf[MT][KS], ryb[CT][KS2];
#pragma unroll
    for (int s1 = 0; s1 < MT; ++s1)
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int yr = yrow(16 * s1 + rr, 4 * s + g);
        ryf[s1][s] = yr < 0 ? 0.f : tby[ly][yr];
      }
#pragma unroll
    for (int s1 = 0; s1 < CT; ++s1)
#pragma unroll
      for (int s = 0; s < KS2; ++s) {
        const int yr = yrow(4 * s + g, 16 * s1 + rr);
        ryb[s1][s] = yr < 0 ? 0.f : tby[ly][yr];
      }
    for (int lx = 0; lx < TX; ++lx) {
      const int ex = ex0 + lx;
      if (ex >= A.nex) break;
      // G_q at the rows this lane holds after the forward product:
      // m = 16 s1 + 4 g + v, q = 8 s1 + 2 g + (v >> 1)
      const int tileG = (ey / A.L.ty) * A.L.ntx + ex / A.L.tx;
      const int tt = (ey % A.L.ty) * A.L.tx + ex % A.L.tx;
      float g11[MT][2], g12[MT][2], g22[MT][2];
#pragma unroll
      for (int s1 = 0; s1 < MT; ++s1)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int q = 8 * s1 + 2 * g + h;
          const bool in = q < NQ2;
          const int qq = in ? q : 0;
          g11[s1][h] = in ? A.G[gq_at<float, NQ>(tileG, 0, qq, tt, NTL, split)] : 0.f;
          g12[s1][h] = in ? A.G[gq_at<float, NQ>(tileG, 1, qq, tt, NTL, split)] : 0.f;
          g22[s1][h] = in ? A.G[gq_at<float, NQ>(tileG, 2, qq, tt, NTL, split)] : 0.f;
        }
      // X relative to the element's first coefficient (the operator
      // annihilates constants; see centre_element)
      const int fo = ly * FX + lx;
      const float x00 = xs[rr][fo];
      float xb[KS];
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int c = 4 * s + g;
        xb[s] = c < NC ? xs[rr][fo + coff(c)] - x00 : 0.f;
      }
      f4 acc[MT];
#pragma unroll
      for (int s1 = 0; s1 < MT; ++s1) acc[s1] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int s1 = 0; s1 < MT; ++s1) {
          const float a = ryf[s1][s] * tbx[lx][xrow(16 * s1 + rr, 4 * s + g)];
          acc[s1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xb[s], acc[s1], 0, 0, 0);
        }
      // flux, then F to this wave's LDS plane in [m][r] order
#pragma unroll
      for (int s1 = 0; s1 < MT; ++s1)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const float ux = acc[s1][2 * h], uy = acc[s1][2 * h + 1];
          fb[w][16 * s1 + 4 * g + 2 * h][rr] = -(g11[s1][h] * ux + g12[s1][h] * uy);
          fb[w][16 * s1 + 4 * g + 2 * h + 1][rr] = -(g12[s1][h] * ux + g22[s1][h] * uy);
        }
      __builtin_amdgcn_wave_barrier();
      f4 acc2[CT];
#pragma unroll
      for (int s1 = 0; s1 < CT; ++s1) acc2[s1] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < KS2; ++s) {
        const float fbv = fb[w][4 * s + g][rr];
#pragma unroll
        for (int s1 = 0; s1 < CT; ++s1) {
          const float a = ryb[s1][s] * tbx[lx][xrow(4 * s + g, 16 * s1 + rr)];
          acc2[s1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, fbv, acc2[s1], 0, 0, 0);
        }
      }
      __builtin_amdgcn_wave_barrier();
      // Y[c][r], c = 16 s1 + 4 g + v, into the tile footprint
#pragma unroll
      for (int s1 = 0; s1 < CT; ++s1)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = 16 * s1 + 4 * g + v;
          if (c < NC) atomicAdd(&ys[rr][fo + coff(c)], acc2[s1][v]);
        }
    }
  }
  __syncthreads();

  ApplyArgs oa = {};
  oa.dirmask = A.dirmask;
  oa.nex = A.nex;
  oa.ney = A.ney;
  oa.nfx = A.nfx;
  oa.nfy = A.nfy;
  oa.mode = 0;
  oa.mask_out = 1;
  const bool interior = bx > 0 && by > 0 && bx < A.nbx - 1 && by < A.nby - 1 && !A.dirmask;
  double dnone = 0.0;
  for (int i = t; i < RC * NFP; i += NT) {
    const int jr = i / NFP, kk = i % NFP;
    if (j0 + jr >= A.k) continue;
    put_footprint<float, P, TX, TY, false>(oa, A.Y + (size_t)(j0 + jr) * A.ld, nullptr, ex0,
                                           ey0, kk % FX, kk / FX, ys[jr][kk], 0.f, dnone,
                                           interior);
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

// compiled block configurations (tile 16 x TY elements, RC rhs per
// workgroup; mfma: k_apply_block_mfma)
struct BlockCfg {
  int txb, tyb, kr, mfma;
};
constexpr int NBLOCKCFG = 8;
constexpr int MFMA_BLOCK_PMAX = 4;  // register budget of the MFMA form
inline BlockCfg block_cfg(int i) {
  static const BlockCfg c[NBLOCKCFG] = {{16, 16, 4, 0}, {16, 16, 8, 0}, {16, 16, 16, 0},
                                        {16, 8, 4, 0},  {16, 8, 8, 0},  {16, 8, 16, 0},
                                        {16, 16, 16, 1}, {16, 8, 16, 1}};
  return c[i];
}
// configurations compiled for degree p
inline bool block_cfg_compiled(int p, int i) {
  if (i < 0 || i >= NBLOCKCFG) return false;
  return !block_cfg(i).mfma || p <= MFMA_BLOCK_PMAX;
}

template <int P>
struct LaunchBlock {
  static constexpr int NQ = P + 1;
  template <int TY, int RC, bool MF = false>
  static hipError_t run_t(BlockArgs a, hipStream_t s) {
    constexpr int TX = 16;
    a.nbx = (a.nex + TX - 1) / TX;
    a.nby = (a.ney + TY - 1) / TY;
    a.nch = (a.k + RC - 1) / RC;
    const int64_t n = (int64_t)(a.nbx - 1) * P * a.nfy + (int64_t)(a.nby - 1) * P * a.nfx;
    if (n > 0)
      hipLaunchKernelGGL((k_seam_block<P>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                         a.Y, a.ld, a.k, a.nex, a.ney, a.nfx, a.nfy, TX, TY, a.nbx, a.nby);
    if constexpr (MF) {
      if constexpr (P <= MFMA_BLOCK_PMAX)
        hipLaunchKernelGGL((k_apply_block_mfma<P, NQ, TY>), dim3(a.nbx * a.nby * a.nch),
                           dim3(256), 0, s, a);
      else
        return hipErrorInvalidValue;
    } else {
      hipLaunchKernelGGL((k_apply_block<P, NQ, TY, RC>), dim3(a.nbx * a.nby * a.nch),
                         dim3(TX * TY), 0, s, a);
    }
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
      case 6: return run_t<16, 16, true>(a, s);
      case 7: return run_t<8, 16, true>(a, s);
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
