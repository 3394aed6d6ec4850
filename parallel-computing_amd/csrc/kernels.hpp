// kernels.hpp -- gfx950 kernels of the matrix-free operator, templated on the
// scalar type T (float / double), the degree P, the Gauss points per
// direction NQ and the element tile TX x TY.
//
// Work decomposition (one workgroup = one tile of TX x TY elements, one
// element per thread):
//   1. the tile's footprint of x ((TX+P) x (TY+P) functions) and the 1D basis
//      tables of its element columns / rows are staged in LDS;
//   2. each thread streams its element's G_q (3 x NQ^2 values, tile-major SoA,
//      coalesced) and applies the element operator, sum-factorised (SF) or
//      dense (EBE);
//   3. the element vectors are reduced over the footprint in LDS in a fixed
//      order (P+1 passes, one element-row offset per pass);
//   4. functions touched by this tile only are stored; functions on a tile
//      seam are either atomically added (IEN scatter-add, y pre-zeroed on the
//      seams) or, in the inverse-LM gather, written to the slot of this
//      tile's position among the (up to four) tiles touching the function:
//      the lowest tile's partial into y, the others' into three
//      function-indexed slot arrays, zero wherever no tile writes.  The sum
//      y + slot1 + slot2 + slot3, in that fixed order, is formed by
//      k_slot_fold or by the CG update kernel that consumes y (bitwise
//      reproducible, no atomics).
// Dirichlet rows are written as zero; x is gathered for every function
// (matrix_free/src/GlobalAssemblyMF.cpp:166-205).
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>

// 1: load all of an element's G_q into registers at kernel entry;
// 0: load each quadrature row's factors where they are used
#ifndef IGAK_PREFETCH_G
#define IGAK_PREFETCH_G 1
#endif
#include <stdint.h>

namespace igak {

struct Layout {
  int tx, ty;    // tile shape (elements)
  int ntx, nty;  // tiles per direction
};

struct ApplyArgs {
  const void *x;
  void *y;
  const void *G;          // [tile][3][NQ*NQ][TX*TY]
  const void *tabx;       // [nelem_x][2][NQ][P+1] : N, dN/dxi (reference coords)
  const void *taby;       // [nelem_y][2][NQ][P+1]
  const uint8_t *dirmask; // nullptr = Dirichlet on the boundary ring
  const void *wf;         // per-function NURBS weights (general rational meshes) or nullptr
  int nex, ney, nfx, nfy, ntx, nty;
  int mode;               // 0 scatter-add, 1 gather into the function slots
  void *slots;            // mode 1: [3][nfx*nfy] partials of the x-, y- and diagonal
                          // neighbour of the lowest touching tile (the lowest writes y)
  int mask_in;            // treat x as P_int x
  int mask_out;           // zero Dirichlet rows of y
  const int *skip;        // device flag: nonzero = do nothing (converged solve)
  // CG fusion (all null / 0 for a plain apply):
  double *dot;            // [tile]: sum over the tile's non-Dirichlet footprint of
                          // x_f * (its contribution to y_f), so sum(dot) = x . y
  const double *hsrc;     // workgroup 0 copies hn doubles hsrc -> hdst (hand-off of
  double *hdst;           // the CG state between the update kernels around this one)
  int hn;
  float gscale;           // FP16-stored G_q: the power of two they were divided by
  int rev;                // k_apply: visit the tiles in reverse order
  int ntg;                // host side: the FP32 split-element kernels' streaming
                          // instantiation (NTG) loads G_q
};

__device__ __forceinline__ bool is_dirichlet(const uint8_t *mask, int A, int gx,
                                             int gy, int nfx, int nfy) {
  if (mask) return mask[A] != 0;
  return gx == 0 || gy == 0 || gx == nfx - 1 || gy == nfy - 1;
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of tiles so that neighbouring tiles (which share footprint
// rows of x) hit the same L2.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int xcd = bid & 7, li = bid >> 3;
  return xcd * q + (xcd < r ? xcd : r) + li;
}

// Tiles touching function g along one direction: elements [g-P, g] clipped.
template <int P>
__device__ __forceinline__ void touching_tiles(int g, int nel, int tile, int &lo,
                                               int &hi) {
  int e0 = g - P;
  if (e0 < 0) e0 = 0;
  int e1 = g;
  if (e1 > nel - 1) e1 = nel - 1;
  lo = e0 / tile;
  hi = e1 / tile;
}

template <typename T>
__device__ __forceinline__ void atomic_add(T *p, T v) {
  atomicAdd(p, v);
}

// the same with power-of-two tiles of size 1 << sh
template <int P>
__device__ __forceinline__ void touching_tiles_sh(int g, int nel, int sh, int &lo, int &hi) {
  int e0 = g - P;
  if (e0 < 0) e0 = 0;
  int e1 = g;
  if (e1 > nel - 1) e1 = nel - 1;
  lo = e0 >> sh;
  hi = e1 >> sh;
}

// --------------------------------------------------------------------------
// G_q storage: per tile and component c (G11, G12, G22; on general rational
// meshes also a_xi, a_eta, so NC = 5) a block of NQ2*NT values; the first
// NQ2 - NQ2 % VW points are stored as VW-wide vectors per thread
// ([chunk][t][VW], one 16-byte load each), the rest as [q][t].
// --------------------------------------------------------------------------
template <typename T, int NQ2, int NC = 3>
__device__ __forceinline__ size_t gq_index(int tile, int c, int q, int t, int NT) {
  constexpr int VW = 16 / sizeof(T);
  constexpr int NV = NQ2 / VW;
  const size_t base = ((size_t)tile * NC + c) * NQ2 * NT;
  if (q < NV * VW) return base + ((size_t)(q / VW) * NT + t) * VW + q % VW;
  return base + (size_t)NV * VW * NT + (size_t)(q - NV * VW) * NT + t;
}

// The split-element kernels (two threads per element, 16 x 8 tiles) read
// FP32 and FP16 factors in a thread-major order when the points per
// direction are even: thread t = 2 el + h of the tile holds, for every row
// j, the columns i = h NI + il (il < NI = NQ/2) of element el as NI
// consecutive values ([tile][c][j][t][il]), so a thread reads one row of one
// component with one load (with the element-major order the half rows
// straddle the 16-byte vectors, and the reads become single strided dwords).
// FP64 keeps the element-major order.
constexpr bool gq_split(int tsize, int tx, int ty, int nq) {
  return tsize <= 4 && tx == 16 && ty == 8 && nq % 2 == 0;
}

template <int NQ, int NI, int NC>
__device__ __forceinline__ size_t gq_split_index(int tile, int c, int j, int il, int t, int NT) {
  return (((size_t)tile * NC + c) * NQ + j) * (size_t)(NI * NT) + (size_t)t * NI + il;
}

// Index of point q = j NQ + i of element el (of NE) of the tile, component c.
template <typename T, int NQ, int NC = 3>
__device__ __forceinline__ size_t gq_at(int tile, int c, int q, int el, int NE, bool split) {
  if constexpr (NQ % 2 == 0) {
    if (split) {
      constexpr int NI = NQ / 2;
      const int j = q / NQ, i = q - j * NQ, h = i >= NI ? 1 : 0;
      return gq_split_index<NQ, NI, NC>(tile, c, j, i - h * NI, 2 * el + h, 2 * NE);
    }
  }
  return gq_index<T, NQ * NQ, NC>(tile, c, q, el, NE);
}

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec16<_Float16> {
  typedef _Float16 type __attribute__((ext_vector_type(8)));
};

// Loads of the geometry factors.  FP64 factors (2x the FP32 bytes: more than
// the device's 256 MB last-level cache holds at the 1024^2 p = 3 size) are
// loaded non-temporally: the stream, read once per apply, then does not
// evict the vectors the surrounding solver kernels re-read every iteration
// (the FP64 CG solve of that mesh measured 7% faster).  FP32 factors use
// normal loads by default: part of them stays in that cache from one apply
// to the next (consecutive applies visit the tiles in opposite orders), and
// streaming them measured slower (45.7 against 54.7 us per apply).  The
// split-element FP32 kernels also have a streaming instantiation (NTG), for
// factors several times larger than that cache, where nothing survives
// from one apply to the next and the stream only evicts x and the slots.
template <bool NT, typename TG>
__device__ __forceinline__ TG gload_t(const TG *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
typedef double double2_v __attribute__((ext_vector_type(2)));
template <typename TG>
constexpr bool kStreamG = std::is_same_v<TG, double> || std::is_same_v<TG, double2_v>;
template <typename TG>
__device__ __forceinline__ TG gload(const TG *p) {
  return gload_t<kStreamG<TG>>(p);
}

// G_q of element t of a tile into registers of type T (TG: storage type)
template <typename TG, int NQ2, int NC = 3, bool NTG = kStreamG<TG>, typename T>
__device__ __forceinline__ void load_gq(const TG *G, int tile, int t, int NT, T (&g)[NC][NQ2]) {
  constexpr int VW = 16 / sizeof(TG);
  constexpr int NV = NQ2 / VW;
  typedef typename Vec16<TG>::type V;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const TG *Gc = G + ((size_t)tile * NC + c) * NQ2 * NT;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const V v = gload_t<NTG>(reinterpret_cast<const V *>(Gc + ((size_t)k * NT + t) * VW));
#pragma unroll
      for (int l = 0; l < VW; ++l) g[c][k * VW + l] = (T)v[l];
    }
#pragma unroll
    for (int r = NV * VW; r < NQ2; ++r)
      g[c][r] = (T)gload_t<NTG>(Gc + (size_t)NV * VW * NT + (size_t)(r - NV * VW) * NT + t);
  }
}

// --------------------------------------------------------------------------
// element operators.  Tables live in LDS as tb[component][lane]:
// component i*NB+a is N(xi_i) of local function a, TAB + i*NB + a its
// reference derivative; `lane` is the element column (x) or row (y) in the
// tile.  X[b][a], ye[b][a]: b = row offset (eta), a = column offset (xi).
// --------------------------------------------------------------------------

// FP32 inputs of a polynomial or tensor-weight element are taken relative to
// its first value: the basis is a partition of unity, so A_e 1 = 0 and the
// element vector is unchanged, while the FP32 contractions see the local
// variation of x instead of its magnitude.  Without it the FP32 operator's
// rounding is relative to |x|, which for the smooth iterates of a CG is far
// above the variation across an element, and the FP32 inner solve needs
// about 20 % more iterations (quarter annulus, Jacobi).  Not for general
// weights (their quotient rule uses the weighted value itself) or FP64.
template <typename T, int P, bool RAT>
__device__ __forceinline__ void centre_element(const T (&X0)[P + 1][P + 1],
                                               T (&X)[P + 1][P + 1]) {
  constexpr bool C = std::is_same_v<T, float> && !RAT;
  const T c = C ? X0[0][0] : T(0);
#pragma unroll
  for (int b = 0; b <= P; ++b)
#pragma unroll
    for (int a = 0; a <= P; ++a) X[b][a] = C ? X0[b][a] - c : X0[b][a];
}

// Sum-factorised action ye = -sum_q gradR^T G_q gradR X.  The eta loop fuses
// the forward eta contraction, the quadrature-point operator and the backward
// eta contraction, so only one table row is live at a time.
// NI quadrature columns i0 .. i0 + NI - 1 (all of them by default): with
// NI < NQ two threads share an element, each contracting its half of the
// columns; the element vector is then the sum of the two threads' ye.
template <typename T, int P, int NQ, int LX, int LY, bool RAT = false, int NI = NQ,
          typename GL>
__device__ __forceinline__ void element_sf(const T (&X0)[P + 1][P + 1],
                                           const T (*tbx)[LX], const T (*tby)[LY],
                                           int lx, int ly, const GL &gl,
                                           T (&ye)[P + 1][P + 1], int i0 = 0) {
  constexpr int NB = P + 1, TAB = NQ * NB;
  T X[NB][NB];
  centre_element<T, P, RAT>(X0, X);
  T t0[NB][NI], t1[NB][NI];
#pragma unroll
  for (int il = 0; il < NI; ++il) {
    const int i = i0 + il;
    const bool in = NI == NQ || i < NQ;
    T n1[NB], d1[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      n1[a] = in ? tbx[i * NB + a][lx] : T(0);
      d1[a] = in ? tbx[TAB + i * NB + a][lx] : T(0);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      T s0 = T(0), s1 = T(0);
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        s0 += n1[a] * X[b][a];
        s1 += d1[a] * X[b][a];
      }
      t0[b][il] = s0;
      t1[b][il] = s1;
    }
  }
  T s0[NB][NI], s1[NB][NI];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int il = 0; il < NI; ++il) {
      s0[b][il] = T(0);
      s1[b][il] = T(0);
    }
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    T n2[NB], d2[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      n2[b] = tby[j * NB + b][ly];
      d2[b] = tby[TAB + j * NB + b][ly];
    }
#pragma unroll
    for (int il = 0; il < NI; ++il) {
      const int i = i0 + il;
      if (NI != NQ && i >= NQ) continue;
      T ux = T(0), uy = T(0), uv = T(0);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        ux += n2[b] * t1[b][il];
        uy += d2[b] * t0[b][il];
        if (RAT) uv += n2[b] * t0[b][il];
      }
      // G accessor: point q = j*NQ + i, also given as (j, il) so that a
      // thread's own columns can come from a statically indexed array
      const int q = j * NQ + i;
      T ax = T(0), ay = T(0);
      if (RAT) {
        // grad(U/W) W = grad U - U a with a = grad W / W (1/W^2 is in G)
        ax = gl(3, q, j, il);
        ay = gl(4, q, j, il);
        ux -= uv * ax;
        uy -= uv * ay;
      }
      const T g11 = gl(0, q, j, il), g12 = gl(1, q, j, il), g22 = gl(2, q, j, il);
      const T fx = -(g11 * ux + g12 * uy);
      const T fy = -(g12 * ux + g22 * uy);
      const T fv = RAT ? -(ax * fx + ay * fy) : T(0);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        s0[b][il] += n2[b] * fx;
        s1[b][il] += RAT ? d2[b] * fy + n2[b] * fv : d2[b] * fy;
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int a = 0; a < NB; ++a) ye[b][a] = T(0);
#pragma unroll
  for (int il = 0; il < NI; ++il) {
    const int i = i0 + il;
    const bool in = NI == NQ || i < NQ;
    T n1[NB], d1[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      n1[a] = in ? tbx[i * NB + a][lx] : T(0);
      d1[a] = in ? tbx[TAB + i * NB + a][lx] : T(0);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int a = 0; a < NB; ++a) ye[b][a] += d1[a] * s0[b][il] + n1[a] * s1[b][il];
  }
}
template <typename T, int P, int NQ, int LX, int LY, bool RAT = false, typename GL>
__device__ __forceinline__ void element_ebe(const T (&X0)[P + 1][P + 1],
                                            const T (*tbx)[LX], const T (*tby)[LY],
                                            int lx, int ly, const GL &gl,
                                            T (&ye)[P + 1][P + 1]) {
  constexpr int NB = P + 1, TAB = NQ * NB;
  T X[NB][NB];
  centre_element<T, P, RAT>(X0, X);
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int a = 0; a < NB; ++a) ye[b][a] = T(0);
#pragma unroll
  for (int j = 0; j < NQ; ++j) {
    T n2[NB], d2[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      n2[b] = tby[j * NB + b][ly];
      d2[b] = tby[TAB + j * NB + b][ly];
    }
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      T n1[NB], d1[NB];
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        n1[a] = tbx[i * NB + a][lx];
        d1[a] = tbx[TAB + i * NB + a][lx];
      }
      T ux = T(0), uy = T(0), uv = T(0);
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          ux += (d1[a] * n2[b]) * X[b][a];
          uy += (n1[a] * d2[b]) * X[b][a];
          if (RAT) uv += (n1[a] * n2[b]) * X[b][a];
        }
      const int q = j * NQ + i;
      T ax = T(0), ay = T(0);
      if (RAT) {
        ax = gl(3, q, j, i);
        ay = gl(4, q, j, i);
        ux -= uv * ax;
        uy -= uv * ay;
      }
      const T g11 = gl(0, q, j, i), g12 = gl(1, q, j, i), g22 = gl(2, q, j, i);
      const T fx = -(g11 * ux + g12 * uy);
      const T fy = -(g12 * ux + g22 * uy);
      const T fv = RAT ? -(ax * fx + ay * fy) : T(0);
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          ye[b][a] += (d1[a] * n2[b]) * fx + (n1[a] * d2[b]) * fy;
          if (RAT) ye[b][a] += (n1[a] * n2[b]) * fv;
        }
    }
  }
}

// --------------------------------------------------------------------------
// Tile staging: the x footprint and the tile's 1D table entries are loaded
// into registers first (load), the G_q loads are issued after them, and
// only then are the staged values written to LDS (store).  The memory
// counter completes in issue order, so the LDS stores wait for the x and
// table loads alone while the G_q loads stay in flight.
// --------------------------------------------------------------------------
template <typename T, int P, int NQ, int TX, int TY, bool RAT, int NTH = TX * TY>
struct TileStage {
  static constexpr int NB = P + 1, NT = NTH, FX = TX + P, FY = TY + P, NFP = FX * FY;
  static constexpr int TAB = NQ * NB, NOWN = (NFP + NT - 1) / NT;
  static constexpr int NTBX = (2 * TAB * TX + NT - 1) / NT, NTBY = (2 * TAB * TY + NT - 1) / NT;
  T xv[NOWN];
  T wv[RAT ? NOWN : 1];
  T tbx[NTBX], tby[NTBY];

  __device__ __forceinline__ void load(const ApplyArgs &A, int ex0, int ey0, int t) {
    const T *x = static_cast<const T *>(A.x);
    const T *wf = static_cast<const T *>(A.wf);
    const T *tx_ = static_cast<const T *>(A.tabx);
    const T *ty_ = static_cast<const T *>(A.taby);
#pragma unroll
    for (int o = 0; o < NOWN; ++o) {
      const int k = t + o * NT;
      xv[o] = T(0);
      if (RAT) wv[o] = T(0);
      if (k < NFP) {
        const int gx = ex0 + k % FX, gy = ey0 + k / FX;
        if (gx < A.nfx && gy < A.nfy) {
          const int Ag = gy * A.nfx + gx;
          xv[o] = x[Ag];
          if (RAT) wv[o] = wf[Ag];
        }
      }
    }
#pragma unroll
    for (int o = 0; o < NTBX; ++o) {
      const int k = t + o * NT;
      tbx[o] = T(0);
      if (k < 2 * TAB * TX) {
        int e = ex0 + k % TX;
        if (e > A.nex - 1) e = A.nex - 1;
        tbx[o] = tx_[(size_t)e * 2 * TAB + k / TX];
      }
    }
#pragma unroll
    for (int o = 0; o < NTBY; ++o) {
      const int k = t + o * NT;
      tby[o] = T(0);
      if (k < 2 * TAB * TY) {
        int e = ey0 + k % TY;
        if (e > A.ney - 1) e = A.ney - 1;
        tby[o] = ty_[(size_t)e * 2 * TAB + k / TY];
      }
    }
  }

  // LX, LY: row strides of the LDS tables (TX, TY, or padded)
  template <int LX = TX, int LY = TY>
  __device__ __forceinline__ void store(const ApplyArgs &A, int ex0, int ey0, int t, T *xs,
                                        T (*ltx)[LX], T (*lty)[LY]) const {
#pragma unroll
    for (int o = 0; o < NOWN; ++o) {
      const int k = t + o * NT;
      if (k < NFP) {
        T v = xv[o];
        if (RAT) v *= wv[o];
        if (A.mask_in) {
          const int gx = ex0 + k % FX, gy = ey0 + k / FX;
          if (gx < A.nfx && gy < A.nfy &&
              is_dirichlet(A.dirmask, gy * A.nfx + gx, gx, gy, A.nfx, A.nfy))
            v = T(0);
        }
        xs[k] = v;
      }
    }
#pragma unroll
    for (int o = 0; o < NTBX; ++o) {
      const int k = t + o * NT;
      if (k < 2 * TAB * TX) ltx[k / TX][k % TX] = tbx[o];
    }
#pragma unroll
    for (int o = 0; o < NTBY; ++o) {
      const int k = t + o * NT;
      if (k < 2 * TAB * TY) lty[k / TY][k % TY] = tby[o];
    }
  }
};

// --------------------------------------------------------------------------
// footprint reduction and store
// --------------------------------------------------------------------------
// One footprint value v at (fx, fy) of the tile at (ex0, ey0).  Functions
// only this tile touches are stored.  Seam functions are atomically added
// (mode 0, scatter-add, after k_seam zeroed them) or, in the gather (mode 1),
// written to the slot of this tile's position among the touching tiles: 0
// the lowest (into y), 1 its x-, 2 its y-, 3 its diagonal neighbour (into
// slot array k - 1).  Every touching tile writes its slot on every apply,
// Dirichlet rows included (as zeros), so a slot holds this apply's partial or
// the zero it was created with.  Dirichlet rows get 0; general rational
// meshes scale by the function's weight.
// xin: the staged input at this footprint point (x, or w x for general
// weights), accumulated into dacc as xin * v when A.dot is set.
// interior: the tile has neighbours on all four sides and the Dirichlet set
// is the boundary ring (or rows are not masked), so no footprint function
// is Dirichlet or outside the mesh and a function is shared exactly when it
// lies in the first P or the last P footprint columns / rows.
template <typename T, int P, int TX, int TY, bool RAT>
__device__ __forceinline__ void put_footprint_slots(const ApplyArgs &A, T *y, const T *wf,
                                                    int ex0, int ey0, int fx, int fy, T v,
                                                    T xin, double &dacc, bool interior) {
  const int gx = ex0 + fx, gy = ey0 + fy;
  const int Ag = gy * A.nfx + gx;
  int k;
  if (interior) {  // first P columns / rows: this tile is the upper neighbour
    k = (fx < P ? 1 : 0) + (fy < P ? 2 : 0);
  } else {
    if (gx >= A.nfx || gy >= A.nfy) return;
    int xlo, xhi, ylo, yhi;
    touching_tiles<P>(gx, A.nex, TX, xlo, xhi);
    touching_tiles<P>(gy, A.ney, TY, ylo, yhi);
    k = (ex0 / TX > xlo ? 1 : 0) + (ey0 / TY > ylo ? 2 : 0);
    if (A.mask_out && is_dirichlet(A.dirmask, Ag, gx, gy, A.nfx, A.nfy)) v = T(0), xin = T(0);
  }
  if (A.dot) dacc += (double)xin * (double)v;
  if (RAT) v *= wf[Ag];
  T *const d = k ? static_cast<T *>(A.slots) + (size_t)(k - 1) * A.nfx * A.nfy : y;
  d[Ag] = v;
}

template <typename T, int P, int TX, int TY, bool RAT>
__device__ __forceinline__ void put_footprint(const ApplyArgs &A, T *y, const T *wf, int ex0,
                                              int ey0, int fx, int fy, T v, T xin, double &dacc,
                                              bool interior) {
  if (A.mode != 0) {
    put_footprint_slots<T, P, TX, TY, RAT>(A, y, wf, ex0, ey0, fx, fy, v, xin, dacc, interior);
    return;
  }
  const int gx = ex0 + fx, gy = ey0 + fy;
  if (interior) {
    const int Ag = gy * A.nfx + gx;
    const bool shared = fx < P || fx >= TX || fy < P || fy >= TY;
    if (A.dot) dacc += (double)xin * (double)v;
    if (RAT) v *= wf[Ag];
    if (!shared)
      y[Ag] = v;
    else
      atomic_add(&y[Ag], v);
    return;
  }
  if (gx >= A.nfx || gy >= A.nfy) return;
  int xlo, xhi, ylo, yhi;
  touching_tiles<P>(gx, A.nex, TX, xlo, xhi);
  touching_tiles<P>(gy, A.ney, TY, ylo, yhi);
  const bool shared = (xlo != xhi) || (ylo != yhi);
  const int Ag = gy * A.nfx + gx;
  if (A.mask_out && is_dirichlet(A.dirmask, Ag, gx, gy, A.nfx, A.nfy)) {
    y[Ag] = T(0);
    return;
  }
  if (A.dot) dacc += (double)xin * (double)v;
  if (RAT) v *= wf[Ag];
  if (!shared)
    y[Ag] = v;
  else
    atomic_add(&y[Ag], v);
}

// Cross-lane move inside 16-lane DPP rows (zero where the source lane falls
// outside the row).  CTRL: 0x110 + n = row_shr:n, 0x100 + n = row_shl:n.
template <int CTRL>
__device__ __forceinline__ float dpp_row(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_row(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Overlap-add of the element vectors of a 16-wide tile (element (lx, ly) on
// lane lx of DPP row ly % 4).  Along x inside each row: column lx of the
// tile collects ye[b][a] of lane lx - a (row_shr:a), the P columns past the
// tile's right edge collect lanes 16 + c - a (row_shl:16-a, valid on lanes
// c < P).  Along y through LDS: rows ly - b.  One barrier; fixed order, so
// bitwise reproducible.  lds needs 2 * (P + 1) * 16 * TY values.
template <int P, int A_, typename T>
__device__ __forceinline__ void row16_x(const T (&ye)[P + 1][P + 1], T (&rin)[P + 1],
                                        T (&rwr)[P + 1]) {
  if constexpr (A_ <= P) {
#pragma unroll
    for (int b = 0; b <= P; ++b) {
      rin[b] += dpp_row<0x110 + A_>(ye[b][A_]);
      rwr[b] += dpp_row<0x100 + 16 - A_>(ye[b][A_]);
    }
    row16_x<P, A_ + 1>(ye, rin, rwr);
  }
}

// LDS values reduce_store_row16 needs for a 16 x TY tile: element-row sums
// of the tile's own footprint columns (bin) and of the P columns past its
// right edge (bwr), each with P zero rows above and below so that the
// column sums read no out-of-tile rows conditionally.
template <int P, int TY>
constexpr int row16_lds() {
  return 2 * (P + 1) * (TY + 2 * P) * 16;
}

template <typename T, int P, int TY, bool RAT>
__device__ __forceinline__ void reduce_store_row16(const ApplyArgs &A, const T (&ye)[P + 1][P + 1],
                                                   T *lds, int t, int ex0, int ey0, T *y,
                                                   const T *wf, const T *xs, double &dacc,
                                                   bool interior) {
  constexpr int FX = 16 + P;
  constexpr int NB = P + 1, NT = 16 * TY, R = (TY + 2 * P) * 16;  // values per b
  const int lx = t & 15, ly = t >> 4;
  T rin[NB], rwr[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    rin[b] = ye[b][0];
    rwr[b] = T(0);
  }
  row16_x<P, 1>(ye, rin, rwr);  // rwr is zero on lanes lx >= P
  T *bin = lds, *bwr = lds + NB * R;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    bin[b * R + (ly + P) * 16 + lx] = rin[b];
    bwr[b * R + (ly + P) * 16 + lx] = rwr[b];
  }
  // the zero rows: 2 buffers x NB x 2P rows x 16
  for (int i = t; i < 2 * NB * 2 * P * 16; i += NT) {
    const int l = i & 15, r = (i >> 4) % (2 * P), u = (i >> 4) / (2 * P);
    const int row = r < P ? r : TY + r;  // rows -P..-1 and TY..TY+P-1, shifted by P
    lds[u * R + row * 16 + l] = T(0);
  }
  __syncthreads();
  // rows 0..TY-1 of this thread's column, and rows TY..TY+P-1 on ly < P
  T v0 = T(0), w0 = T(0);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    v0 += bin[b * R + (ly - b + P) * 16 + lx];
    w0 += bwr[b * R + (ly - b + P) * 16 + lx];
  }
  auto xin = [&](int fx, int fy) { return A.dot ? xs[fy * FX + fx] : T(0); };
  put_footprint<T, P, 16, TY, RAT>(A, y, wf, ex0, ey0, lx, ly, v0, xin(lx, ly),
                                   dacc, interior);
  if (lx < P)
    put_footprint<T, P, 16, TY, RAT>(A, y, wf, ex0, ey0, 16 + lx, ly, w0,
                                     xin(16 + lx, ly), dacc, interior);
  if (ly < P) {
    T v1 = T(0), w1 = T(0);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      v1 += bin[b * R + (TY + ly - b + P) * 16 + lx];
      w1 += bwr[b * R + (TY + ly - b + P) * 16 + lx];
    }
    put_footprint<T, P, 16, TY, RAT>(A, y, wf, ex0, ey0, lx, TY + ly, v1,
                                     xin(lx, TY + ly), dacc, interior);
    if (lx < P)
      put_footprint<T, P, 16, TY, RAT>(A, y, wf, ex0, ey0, 16 + lx, TY + ly, w1,
                                       xin(16 + lx, TY + ly), dacc, interior);
  }
}

// Sum over the 64 lanes of a wave, on DPP moves (no LDS round trips): pairs
// (quad_perm 1,0,3,2), quads (2,3,0,1), 8 lanes (row_half_mirror), 16 lanes
// (row_mirror), then the four rows in order.  Fixed order: reproducible.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_row<0xB1>(v);
  v += dpp_row<0x4E>(v);
  v += dpp_row<0x141>(v);
  v += dpp_row<0x140>(v);
  auto lane = [&](int l) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
  };
  return ((lane(0) + lane(16)) + lane(32)) + lane(48);
}

// Workgroup sum of the per-thread x.y contributions into A.dot[tile]
// (fixed reduction and wave order: bitwise reproducible).
template <int NT>
__device__ __forceinline__ void store_tile_dot(const ApplyArgs &A, int tile, double v) {
  static_assert(NT % 64 == 0, "whole waves");
  __shared__ double wsum[NT / 64];
  v = wave_sum_dpp(v);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += wsum[i];
    A.dot[tile] = s;
  }
}

// CG state hand-off: workgroup 0 copies A.hn doubles (see ApplyArgs)
__device__ __forceinline__ void cg_handoff(const ApplyArgs &A, int nt) {
  if (A.hn && blockIdx.x == 0)
    for (int i = threadIdx.x; i < A.hn; i += nt) A.hdst[i] = A.hsrc[i];
}

// --------------------------------------------------------------------------
// the apply kernel
// --------------------------------------------------------------------------
// MINW: waves per SIMD the register budget must allow (1 = compiler's
// choice; 4 caps a 256-thread tile at 128 VGPRs and streams G_q on demand).
// SPLIT = 2: two threads per element (sum-factorised only), each
// contracting half of the quadrature columns; halves the per-thread state
// of high degrees and FP64.
// TG: storage type of G_q (T, or _Float16 for the FP16-geometry FP32 apply:
// values divided by A.gscale, the element result multiplied back by it).
// Row stride of the x table of a split-element kernel: its two threads of an
// element read rows NI*NB apart, which with a stride of 16 FP32 values land
// in the same LDS banks (a two-way conflict on every table read); the
// smallest stride >= TX that moves them 16 banks apart removes it.
template <typename T, int NB, int NI, int TX, int SPLIT>
constexpr int xtab_stride() {
  if (SPLIT != 2 || sizeof(T) != 4) return TX;
  for (int L = TX; L < TX + 32; ++L)
    if ((NI * NB * L) % 32 == 16) return L;
  return TX;
}

template <typename T, int P, int NQ, int TX, int TY, bool EBE, bool RAT = false, int MINW = 1,
          int SPLIT = 1, typename TG = T, bool NTG = kStreamG<TG>>
__global__ __launch_bounds__(TX *TY *SPLIT, MINW) void k_apply(ApplyArgs A) {
  static_assert(SPLIT == 1 || !EBE, "split elements are sum-factorised");
  constexpr int NB = P + 1, NE = TX * TY, NT = NE * SPLIT, FX = TX + P, FY = TY + P;
  constexpr int NQ2 = NQ * NQ, TAB = NQ * NB, NFP = FX * FY;
  constexpr bool ROW16 = TX == 16 && SPLIT == 1;
  constexpr int LXT = xtab_stride<T, NB, (NQ + SPLIT - 1) / SPLIT, TX, SPLIT>();
  // split-element tiles reduce the footprint in one pass: every element
  // vector to LDS at once ([b][a][element]), one barrier
  constexpr bool ONEPASS = SPLIT == 2;
  __shared__ T xs[NFP];
  __shared__ T tbx[2 * TAB][LXT];
  __shared__ T tby[2 * TAB][TY];
  __shared__ T bufm[ROW16 ? row16_lds<P, TY>() : (ONEPASS ? NB * NB * NE : NB * NE)];
  T(*buf)[NE] = reinterpret_cast<T(*)[NE]>(bufm);

  if (A.skip && *A.skip) return;
  cg_handoff(A, NT);
  const int t = threadIdx.x, h = t % SPLIT, el = t / SPLIT, lx = el % TX, ly = el / TX;
  const int tile0 = xcd_tile(blockIdx.x, gridDim.x);
  const int tile = A.rev ? (int)gridDim.x - 1 - tile0 : tile0;
  const int tx = tile % A.ntx, ty = tile / A.ntx;
  const int ex0 = tx * TX, ey0 = ty * TY;
  const bool valid = (ex0 + lx < A.nex) && (ey0 + ly < A.ney);
  // uniform over the workgroup: every element of the tile exists; the tile is
  // interior (see put_footprint; gather assembly only: with the scatter-add
  // form the shorter store path measured slower, 123 against 85 us for the
  // FP64 apply of p = 3 on 1024 x 1024 elements)
  const bool full = ex0 + TX <= A.nex && ey0 + TY <= A.ney;
  const bool interior = A.mode == 1 && tx > 0 && ty > 0 && tx < A.ntx - 1 && ty < A.nty - 1 &&
                        (!A.mask_out || !A.dirmask);

  // x footprint and tables first, then the geometry factors of this element
  // (held in registers unless they are many: high degree, general rational)
  TileStage<T, P, NQ, TX, TY, RAT, NT> stg;
  stg.load(A, ex0, ey0, t);
  constexpr int NC = RAT ? 5 : 3;
  constexpr int NI = (NQ + SPLIT - 1) / SPLIT;
  // whole element (one thread) or this thread's quadrature columns (split)
  constexpr bool PREF =
      IGAK_PREFETCH_G && MINW < 4 && SPLIT == 1 && NC * NQ2 * sizeof(T) <= 320;
  // (general rational meshes: at most 320 bytes -- the degree-5 FP32 kernel
  // at 360 bytes held 212 VGPRs, two waves per SIMD, and ran 862 us on the
  // 2048^2 mesh against 812 us loading each row where it is used, 166 VGPRs)
  constexpr bool PREFH =
      IGAK_PREFETCH_G && SPLIT > 1 && NC * NQ * NI * sizeof(T) <= (RAT ? 320 : 400);
  T g[PREF ? NC : 1][PREF ? NQ2 : 1];
  T gh[PREFH ? NC : 1][PREFH ? NQ : 1][PREFH ? NI : 1];
  const TG *Gt = static_cast<const TG *>(A.G);
  // thread-major factors of the split kernels (gq_split): thread t's own
  // point j NI + il of its columns; the 16x8 EBE companion reads the same
  // storage through gq_at
  constexpr bool GS = gq_split(sizeof(TG), TX, TY, NQ);
  static_assert(!GS || SPLIT == 2 || EBE, "thread-major factors: split kernels");
  if constexpr (PREF && !GS) load_gq<TG, NQ2, NC, NTG>(Gt, tile, el, NE, g);
  if constexpr (PREF && GS) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int q = 0; q < NQ2; ++q) g[c][q] = (T)gload_t<NTG>(Gt + gq_at<TG, NQ, NC>(tile, c, q, el, NE, true));
  }
  if constexpr (PREFH) {
    const int i0 = h * NI;
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int j = 0; j < NQ; ++j)
#pragma unroll
        for (int il = 0; il < NI; ++il) {
          if constexpr (GS)
            gh[c][j][il] = (T)gload_t<NTG>(Gt + gq_split_index<NQ, NI, NC>(tile, c, j, il, t, NT));
          else
            gh[c][j][il] =
                (i0 + il < NQ)
                    ? (T)gload_t<NTG>(Gt + gq_index<TG, NQ2, NC>(tile, c, j * NQ + i0 + il, el, NE))
                    : T(0);
        }
  }
  auto gl = [&](int c, int q, int j, int il) {
    if constexpr (PREF)
      return g[c][q];
    else if constexpr (PREFH)
      return gh[c][j][il];
    else if constexpr (GS && SPLIT == 2)
      return (T)gload_t<NTG>(Gt + gq_split_index<NQ, NI, NC>(tile, c, j, il, t, NT));
    else
      return (T)gload_t<NTG>(Gt + gq_at<TG, NQ, NC>(tile, c, q, el, NE, GS));
  };
  const T *wf = static_cast<const T *>(A.wf);
  stg.template store<LXT, TY>(A, ex0, ey0, t, xs, tbx, tby);
  __syncthreads();

  T ye[NB][NB];
  {
    T X[NB][NB];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int a = 0; a < NB; ++a) X[b][a] = xs[(ly + b) * FX + lx + a];
    if constexpr (EBE)
      element_ebe<T, P, NQ, LXT, TY, RAT>(X, tbx, tby, lx, ly, gl, ye);
    else
      element_sf<T, P, NQ, LXT, TY, RAT, NI>(X, tbx, tby, lx, ly, gl, ye, h * NI);
    if constexpr (SPLIT == 2) {
      // the pair (lanes 2e, 2e+1) adds its halves: quad_perm [1,0,3,2]
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int a = 0; a < NB; ++a) ye[b][a] += dpp_row<0xB1>(ye[b][a]);
    }
    if constexpr (!std::is_same_v<TG, T>) {
      const T gs = (T)A.gscale;
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int a = 0; a < NB; ++a) ye[b][a] *= gs;
    }
    if (!full && !valid) {
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int a = 0; a < NB; ++a) ye[b][a] = T(0);
    }
  }

  T *y = static_cast<T *>(A.y);
  double dacc = 0.0;
  if constexpr (ROW16) {
    reduce_store_row16<T, P, TY, RAT>(A, ye, bufm, t, ex0, ey0, y, wf, xs, dacc, interior);
    if (A.dot) store_tile_dot<NT>(A, tile, dacc);
    return;
  }
  // fixed-order reduction over the footprint: pass b adds element row offset b
  constexpr int NOWN = (NFP + NT - 1) / NT;
  T acc[NOWN];
#pragma unroll
  for (int o = 0; o < NOWN; ++o) acc[o] = T(0);
  if constexpr (ONEPASS) {
    // both threads of an element hold its summed vector: half h writes the
    // rows b of parity h; the sums then run in the order of the passes below
    T(*bb)[NB][NE] = reinterpret_cast<T(*)[NB][NE]>(bufm);
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if ((b & 1) == h) {
#pragma unroll
        for (int a = 0; a < NB; ++a) bb[b][a][el] = ye[b][a];
      }
    __syncthreads();
#pragma unroll
    for (int o = 0; o < NOWN; ++o) {
      const int k = t + o * NT;
      if (k < NFP) {
        const int fx = k % FX, fy = k / FX;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int el_y = fy - b;
          if (el_y >= 0 && el_y < TY) {
#pragma unroll
            for (int a = 0; a < NB; ++a) {
              const int el_x = fx - a;
              if (el_x >= 0 && el_x < TX) acc[o] += bb[b][a][el_y * TX + el_x];
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < (ONEPASS ? 0 : NB); ++b) {
    if (b > 0) __syncthreads();
    if (h == 0) {
#pragma unroll
      for (int a = 0; a < NB; ++a) buf[a][el] = ye[b][a];
    }
    __syncthreads();
#pragma unroll
    for (int o = 0; o < NOWN; ++o) {
      const int k = t + o * NT;
      if (k < NFP) {
        const int fx = k % FX, fy = k / FX;
        const int el_y = fy - b;
        if (el_y >= 0 && el_y < TY) {
#pragma unroll
          for (int a = 0; a < NB; ++a) {
            const int el_x = fx - a;
            if (el_x >= 0 && el_x < TX) acc[o] += buf[a][el_y * TX + el_x];
          }
        }
      }
    }
  }
#pragma unroll
  for (int o = 0; o < NOWN; ++o) {
    const int k = t + o * NT;
    if (k < NFP)
      put_footprint<T, P, TX, TY, RAT>(A, y, wf, ex0, ey0, k % FX, k / FX, acc[o],
                                       A.dot ? xs[k] : T(0), dacc, interior);
  }
  if (A.dot) store_tile_dot<NT>(A, tile, dacc);
}

// --------------------------------------------------------------------------
// Matrix-core form of the sum-factorised apply, degree 3 (4 Gauss points per
// direction), 16 x 16 tiles, FP32 arithmetic on FP32 or FP16-stored G_q
// (polynomial and tensor-weight meshes).  The element operator is eight
// 4x4x4 products per element, each taken by v_mfma_f32_4x4x1_16b_f32 (16
// independent 4x4 blocks, rank-1 update per instruction): block = element,
// so one instruction advances 16 elements (one tile row) by one k.  With the
// block layout C[r][c] in lane 4e + c, register r (A[r] from lane 4e + r,
// B[c] from lane 4e + c) every product's accumulator is the next product's
// operand without any lane exchange:
//   T1[b][i] = sum_a X[b][a] d1[i][a]     (T0 with n1)    lane i, reg b
//   Ux[j][i] = sum_b n2[j][b] T1[b][i]    (Uy: d2, T0)    lane i, reg j
//   Fx, Fy   = -G_q (Ux, Uy) at q = (j, i)                 pointwise
//   S0[i][b] = sum_j Fx[j][i] n2[j][b]    (S1: Fy, d2)    lane b, reg i
//   Ye[a][b] = sum_i d1[i][a] S0[i][b] + n1[i][a] S1[i][b] lane b, reg a
// (the element vector ye[b][a] of element e in lane 4e + b, register a).
// Wave w takes tile rows 4w .. 4w+3 in turn.  The FP32 G_q layout puts the
// four points i of row j of one element in one 16-byte vector, so a row of
// 16 elements loads as one contiguous 256-byte access per (component, j).
// The element vectors are reduced over the footprint through LDS in a fixed
// order (element row offset b, then column offset a): bitwise reproducible.
// --------------------------------------------------------------------------
// A pointer that is the same on every lane, as a value the compiler keeps in
// scalar registers (loads from it plus a per-lane offset then need no
// per-lane 64-bit address arithmetic).
typedef const __attribute__((address_space(1))) char global_cchar;
__device__ __forceinline__ global_cchar *wave_uniform(const void *p) {
  const unsigned long long v = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (global_cchar *)(((unsigned long long)hi << 32) | lo);
}

// Register budget of four waves per SIMD (128 VGPRs, no spills): 4 workgroups
// per CU.
template <typename TG>
__global__ __launch_bounds__(256, 4) void k_apply_mfma(ApplyArgs A) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  constexpr int P = 3, NQ = 4, NB = 4, TX = 16, TY = 16, NT = 256, FX = TX + P, FY = TY + P;
  constexpr int NQ2 = NQ * NQ, TAB = NQ * NB, NFP = FX * FY, NOWN = (NFP + NT - 1) / NT;
  constexpr int PS = TX * TY + 16;  // element-vector plane stride (conflict-free stores)
  // tables with rows of 20 values: the rows 4r + a that the four lanes r of
  // an element read at once fall in different LDS banks
  constexpr int LT = 20;
  __shared__ float xs[NFP];
  __shared__ float tbx[2 * TAB][LT];
  __shared__ float tby[2 * TAB][LT];
  __shared__ float yb[NB * NB * PS];

  if (A.skip && *A.skip) return;
  cg_handoff(A, NT);
  // w is wave-uniform: the G_q row bases then live in scalar registers
  const int t = threadIdx.x, lane = t & 63, e = lane >> 2, r = lane & 3;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = tile % A.ntx, ty = tile / A.ntx;
  const int ex0 = tx * TX, ey0 = ty * TY;
  const bool interior = A.mode == 1 && tx > 0 && ty > 0 && tx < A.ntx - 1 && ty < A.nty - 1 &&
                        (!A.mask_out || !A.dirmask);

  // x footprint and tables, then this lane's G_q values of its four rows:
  // point (j, r) of element e of row 4w + k.  gq_index(tile, c, 4j + r,
  // 16(4w + k) + e) is a wave-uniform base plus a per-lane offset 4e + r
  // (FP32: the four points of an element row are one 16-byte vector) or
  // 8e + r (FP16: rows j, j + 1 share an 8-value vector).
  TileStage<float, P, NQ, TX, TY, false, NT> stg;
  stg.load(A, ex0, ey0, t);
  const TG *Gt = static_cast<const TG *>(A.G);
  constexpr bool F32G = std::is_same_v<TG, float>;
  const unsigned goff = (F32G ? (unsigned)lane : (unsigned)(8 * e + r)) * (unsigned)sizeof(TG);
  float g[4][3][NQ];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        const size_t base = (size_t)(tile * 3 + c) * NQ2 * NT +
                            (F32G ? (size_t)(j * NT + (4 * w + k) * TX) * 4
                                  : (size_t)((j / 2) * NT + (4 * w + k) * TX) * 8 + (j % 2) * 4);
        g[k][c][j] = (float)*(const __attribute__((address_space(1))) TG *)(
            wave_uniform(Gt + base) + goff);
      }
  stg.template store<LT, LT>(A, ex0, ey0, t, xs, tbx, tby);
  __syncthreads();

  // x-direction tables of column e: B operands of the first products, A
  // operands of the last
  float b1n[NB], b1d[NB], a5n[NQ], a5d[NQ];
#pragma unroll
  for (int a = 0; a < NB; ++a) {
    b1n[a] = tbx[r * NB + a][e];
    b1d[a] = tbx[TAB + r * NB + a][e];
  }
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    a5n[i] = tbx[i * NB + r][e];
    a5d[i] = tbx[TAB + i * NB + r][e];
  }
  const bool col_ok = ex0 + e < A.nex;
  const float gs = F32G ? 1.0f : A.gscale;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int ly = 4 * w + k;
    float a2n[NB], a2d[NB], b4n[NQ], b4d[NQ], xa[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      a2n[b] = tby[r * NB + b][ly];
      a2d[b] = tby[TAB + r * NB + b][ly];
    }
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      b4n[j] = tby[j * NB + r][ly];
      b4d[j] = tby[TAB + j * NB + r][ly];
    }
    // row r of the element's inputs, relative to its first value (the basis
    // is a partition of unity: the element vector is unchanged, the FP32
    // contractions see the local variation of x)
    const float x0 = xs[ly * FX + e];
#pragma unroll
    for (int a = 0; a < NB; ++a) xa[a] = xs[(ly + r) * FX + e + a] - x0;
    f4 t0 = {0.f, 0.f, 0.f, 0.f}, t1 = t0;
#pragma unroll
    for (int a = 0; a < NB; ++a) {
      t0 = __builtin_amdgcn_mfma_f32_4x4x1f32(xa[a], b1n[a], t0, 0, 0, 0);
      t1 = __builtin_amdgcn_mfma_f32_4x4x1f32(xa[a], b1d[a], t1, 0, 0, 0);
    }
    f4 ux = {0.f, 0.f, 0.f, 0.f}, uy = ux;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      ux = __builtin_amdgcn_mfma_f32_4x4x1f32(a2n[b], t1[b], ux, 0, 0, 0);
      uy = __builtin_amdgcn_mfma_f32_4x4x1f32(a2d[b], t0[b], uy, 0, 0, 0);
    }
    f4 fx, fy;
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      fx[j] = -(g[k][0][j] * ux[j] + g[k][1][j] * uy[j]);
      fy[j] = -(g[k][1][j] * ux[j] + g[k][2][j] * uy[j]);
    }
    f4 s0 = {0.f, 0.f, 0.f, 0.f}, s1 = s0;
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      s0 = __builtin_amdgcn_mfma_f32_4x4x1f32(fx[j], b4n[j], s0, 0, 0, 0);
      s1 = __builtin_amdgcn_mfma_f32_4x4x1f32(fy[j], b4d[j], s1, 0, 0, 0);
    }
    f4 ye = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      ye = __builtin_amdgcn_mfma_f32_4x4x1f32(a5d[i], s0[i], ye, 0, 0, 0);
      ye = __builtin_amdgcn_mfma_f32_4x4x1f32(a5n[i], s1[i], ye, 0, 0, 0);
    }
    // ye[a] = element vector (b = r, a) of element (e, ly); elements outside
    // the mesh (partial tiles) contribute nothing
    const bool ok = col_ok && ey0 + ly < A.ney;
#pragma unroll
    for (int a = 0; a < NB; ++a) yb[(a * NB + r) * PS + ly * TX + e] = ok ? ye[a] * gs : 0.f;
  }
  __syncthreads();

  float *y = static_cast<float *>(A.y);
  double dacc = 0.0;
#pragma unroll
  for (int o = 0; o < NOWN; ++o) {
    const int k = t + o * NT;
    if (k < NFP) {
      const int fx = k % FX, fy = k / FX;
      float v = 0.f;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int ly = fy - b;
#pragma unroll
        for (int a = 0; a < NB; ++a) {
          const int lx = fx - a;
          const bool in = ly >= 0 && ly < TY && lx >= 0 && lx < TX;
          const float u = yb[(a * NB + b) * PS + (in ? ly * TX + lx : 0)];
          v += in ? u : 0.f;
        }
      }
      put_footprint<float, P, TX, TY, false>(A, y, nullptr, ex0, ey0, fx, fy, v,
                                             A.dot ? xs[k] : 0.f, dacc, interior);
    }
  }
  if (A.dot) store_tile_dot<NT>(A, tile, dacc);
}

// Persistent, software-pipelined form of k_apply (sum-factorised): a
// workgroup walks tiles tile0, tile0 + G, ... and keeps the next tile's G_q
// and x footprint loads in flight while it computes and reduces the current
// tile, so HBM streaming does not stop during the compute and barrier phases.
template <typename T, int P, int NQ, int TX, int TY>
__global__ __launch_bounds__(TX *TY) void k_apply_pf(ApplyArgs A, int ntiles) {
  constexpr int NB = P + 1, NT = TX * TY, FX = TX + P, FY = TY + P;
  constexpr int NQ2 = NQ * NQ, TAB = NQ * NB, NFP = FX * FY;
  constexpr int NOWN = (NFP + NT - 1) / NT;
  __shared__ T xs[NFP];
  __shared__ T tbx[2 * TAB][TX];
  __shared__ T tby[2 * TAB][TY];
  __shared__ T bufm[TX == 16 ? row16_lds<P, TY>() : NB * NT];
  T(*buf)[NT] = reinterpret_cast<T(*)[NT]>(bufm);

  if (A.skip && *A.skip) return;
  const int t = threadIdx.x, lx = t % TX, ly = t / TX;
  const T *Gp = static_cast<const T *>(A.G);
  T *y = static_cast<T *>(A.y);

  // this workgroup's tiles: start, start + G, ... (start XCD-contiguous)
  const int G = gridDim.x;
  int tile = xcd_tile(blockIdx.x, G);

  // next tile's x footprint and tables (first), then its G_q, in registers
  TileStage<T, P, NQ, TX, TY, false> stn;
  T gn[3][NQ2];
  auto prefetch = [&](int tl) {
    stn.load(A, (tl % A.ntx) * TX, (tl / A.ntx) * TY, t);
    load_gq<T, NQ2>(Gp, tl, t, NT, gn);
  };
  if (tile < ntiles) prefetch(tile);
  for (; tile < ntiles; tile += G) {
    const int tx = tile % A.ntx, ty = tile / A.ntx;
    const int ex0 = tx * TX, ey0 = ty * TY;
    const bool valid = (ex0 + lx < A.nex) && (ey0 + ly < A.ney);
    const bool interior = A.mode == 1 && tx > 0 && ty > 0 && tx < A.ntx - 1 &&
                          ty < A.nty - 1 && (!A.mask_out || !A.dirmask);
    __syncthreads();  // previous tile's LDS reads are done
    stn.store(A, ex0, ey0, t, xs, tbx, tby);
    T g[3][NQ2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < NQ2; ++q) g[c][q] = gn[c][q];
    const int nxt = tile + G;
    if (nxt < ntiles) prefetch(nxt);
    __syncthreads();

    T ye[NB][NB];
    {
      T X[NB][NB];
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int a = 0; a < NB; ++a) X[b][a] = xs[(ly + b) * FX + lx + a];
      auto gl = [&](int c, int q, int, int) { return g[c][q]; };
      element_sf<T, P, NQ, TX, TY>(X, tbx, tby, lx, ly, gl, ye);
      if (!valid) {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int a = 0; a < NB; ++a) ye[b][a] = T(0);
      }
    }
    double dnone = 0.0;  // A.dot is null here (no CG fusion in this variant)
    if constexpr (TX == 16) {
      reduce_store_row16<T, P, TY, false>(A, ye, bufm, t, ex0, ey0, y, nullptr, xs, dnone,
                                          interior);
      continue;
    }
    T acc[NOWN];
#pragma unroll
    for (int o = 0; o < NOWN; ++o) acc[o] = T(0);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if (b > 0) __syncthreads();
#pragma unroll
      for (int a = 0; a < NB; ++a) buf[a][t] = ye[b][a];
      __syncthreads();
#pragma unroll
      for (int o = 0; o < NOWN; ++o) {
        const int k = t + o * NT;
        if (k < NFP) {
          const int fx = k % FX, fy = k / FX;
          const int el_y = fy - b;
          if (el_y >= 0 && el_y < TY) {
#pragma unroll
            for (int a = 0; a < NB; ++a) {
              const int el_x = fx - a;
              if (el_x >= 0 && el_x < TX) acc[o] += buf[a][el_y * TX + el_x];
            }
          }
        }
      }
    }
#pragma unroll
    for (int o = 0; o < NOWN; ++o) {
      const int k = t + o * NT;
      if (k < NFP)
        put_footprint<T, P, TX, TY, false>(A, y, nullptr, ex0, ey0, k % FX, k / FX,
                                           acc[o], T(0), dnone, interior);
    }
  }
}

// Direct global-to-LDS copy (gfx950 global_load_lds): lane l of the wave
// moves SZ bytes (4 or 16) from src to lds_wave + l * SZ, where lds_wave is
// the same address on every lane of the wave.  The data passes no register;
// it is in LDS once the issuing wave's vmcnt has drained.
template <int SZ>
__device__ __forceinline__ void lds_dma(const void *src, void *lds_wave) {
  static_assert(SZ == 4 || SZ == 16, "dword or dwordx4 copies");
  const auto g = (const __attribute__((address_space(1))) void *)src;
  const auto l = (__attribute__((address_space(3))) void *)lds_wave;
  if constexpr (SZ == 16)
    __builtin_amdgcn_global_load_lds(g, l, 16, 0, 0);
  else
    __builtin_amdgcn_global_load_lds(g, l, 4, 0, 0);
}

// Persistent sum-factorised apply on 16 x 16 tiles with the geometry factors
// staged in LDS (variant 5).  While a workgroup computes and reduces one
// tile, the next tile's G_q stream from HBM straight into LDS and its x
// footprint and 1D tables into registers, so the G_q stream does not stop
// during the compute, reduction and barrier phases and costs no registers.
// Lane t of the copy moves exactly the G_q values that thread t reads back
// (the storage order of gq_index), so the staging needs no barrier of its
// own: each wave waits for its own copies (vmcnt) before reading them, and
// for its reads (lgkmcnt) before it starts the next copy into the same bytes.
// Polynomial and tensor-weight meshes; T = float on TG = float or _Float16.
template <typename T, int P, int NQ, typename TG = T>
__global__ __launch_bounds__(256, 2) void k_apply_lds(ApplyArgs A, int ntiles) {
  constexpr int TX = 16, TY = 16, NB = P + 1, NT = TX * TY, FX = TX + P, FY = TY + P;
  constexpr int NQ2 = NQ * NQ, TAB = NQ * NB, NFP = FX * FY;
  constexpr int VW = 16 / sizeof(TG), NV = NQ2 / VW, NR = NQ2 - NV * VW;
  static_assert(NR == 0 || sizeof(TG) == 4, "points past the vectors are staged as dwords");
  constexpr int CB = NQ2 * NT * sizeof(TG);  // bytes of one component of one tile
  __shared__ T xs[NFP];
  __shared__ T tbx[2 * TAB][TX];
  __shared__ T tby[2 * TAB][TY];
  __shared__ T bufm[row16_lds<P, TY>()];
  __shared__ __attribute__((aligned(16))) unsigned char gs[3 * CB];
  typedef typename Vec16<TG>::type V;

  if (A.skip && *A.skip) return;
  cg_handoff(A, NT);
  const int t = threadIdx.x, lx = t & 15, ly = t >> 4, w0 = t & ~63;
  const unsigned char *Gb = static_cast<const unsigned char *>(A.G);
  T *y = static_cast<T *>(A.y);
  const int G = gridDim.x;
  auto tile_of = [&](int s) { return A.rev ? ntiles - 1 - s : s; };
  TileStage<T, P, NQ, TX, TY, false> stn;
  auto prefetch = [&](int tl) {
    stn.load(A, (tl % A.ntx) * TX, (tl / A.ntx) * TY, t);
    const unsigned char *src = Gb + (size_t)tl * 3 * CB;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        const int o = c * CB + k * NT * 16;
        lds_dma<16>(src + o + t * 16, gs + o + w0 * 16);
      }
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int o = c * CB + NV * NT * 16 + r * NT * 4;
        lds_dma<4>(src + o + t * 4, gs + o + w0 * 4);
      }
    }
  };
  // this workgroup's steps: start, start + G, ... (starts XCD-contiguous)
  int s = xcd_tile(blockIdx.x, G);
  if (s < ntiles) prefetch(tile_of(s));
  for (; s < ntiles; s += G) {
    const int tile = tile_of(s);
    const int tx = tile % A.ntx, ty = tile / A.ntx;
    const int ex0 = tx * TX, ey0 = ty * TY;
    const bool valid = (ex0 + lx < A.nex) && (ey0 + ly < A.ney);
    const bool interior = A.mode == 1 && tx > 0 && ty > 0 && tx < A.ntx - 1 &&
                          ty < A.nty - 1 && (!A.mask_out || !A.dirmask);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this tile's G_q copies and loads
    __syncthreads();  // the previous tile's reads of xs, tables and bufm are done
    stn.store(A, ex0, ey0, t, xs, tbx, tby);
    T g[3][NQ2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        const V v = *reinterpret_cast<const V *>(gs + c * CB + k * NT * 16 + t * 16);
#pragma unroll
        for (int l = 0; l < VW; ++l) g[c][k * VW + l] = (T)v[l];
      }
#pragma unroll
      for (int r = 0; r < NR; ++r)
        g[c][NV * VW + r] =
            (T) * reinterpret_cast<const TG *>(gs + c * CB + NV * NT * 16 + r * NT * 4 + t * 4);
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // g read before gs is refilled
    if (s + G < ntiles) prefetch(tile_of(s + G));
    __syncthreads();  // xs and tables written

    T ye[NB][NB];
    {
      T X[NB][NB];
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int a = 0; a < NB; ++a) X[b][a] = xs[(ly + b) * FX + lx + a];
      auto gl = [&](int c, int q, int, int) { return g[c][q]; };
      element_sf<T, P, NQ, TX, TY>(X, tbx, tby, lx, ly, gl, ye);
      if constexpr (!std::is_same_v<TG, T>) {
        const T gsc = (T)A.gscale;
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int a = 0; a < NB; ++a) ye[b][a] *= gsc;
      }
      if (!valid) {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int a = 0; a < NB; ++a) ye[b][a] = T(0);
      }
    }
    double dacc = 0.0;
    reduce_store_row16<T, P, TY, false>(A, ye, bufm, t, ex0, ey0, y, nullptr, xs, dacc, interior);
    if (A.dot) store_tile_dot<NT>(A, tile, dacc);
  }
}

// Seam functions (touched by more than one tile): zero them before a
// scatter-add apply.  One thread per seam function, numbered row by row
// within each band of P columns right of a vertical tile boundary (so a
// wavefront covers short runs of a row), then the P rows below each
// horizontal boundary (functions not on a vertical seam).  32-bit indices;
// tile sizes are powers of two (shifts sx, sy).
template <typename T, int P>
__global__ void k_seam(T *y, int nex, int nfx, int nfy, int sx, int sy, int ntx, int nty,
                       const int *skip) {
  if (skip && *skip) return;
  const int TX = 1 << sx, TY = 1 << sy;
  const int nv = (ntx - 1) * P * nfy;
  const int nh = (nty - 1) * P * nfx;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv + nh) return;
  int gx, gy;
  if (i < nv) {
    const int pn = P * nfy;
    const int band = i / pn, r = i - band * pn;
    gy = r / P;
    gx = (band + 1) * TX + (r - gy * P);
  } else {
    const int j = i - nv;
    const int pn = P * nfx;
    const int band = j / pn, r = j - band * pn;
    const int rr = r / nfx;
    gy = (band + 1) * TY + rr;
    gx = r - rr * nfx;
    int xlo, xhi;
    touching_tiles_sh<P>(gx, nex, sx, xlo, xhi);
    if (xlo != xhi) return;  // owned by the vertical-band pass
  }
  y[gy * nfx + gx] = T(0);
}

// FP16 copy of FP32 G_q (same tile layout, 8 values per 16-byte vector):
// G16 = G32 / scale, rounded to nearest
template <int NQ2, int NC>
__global__ void k_g16_from_g32(const float *__restrict__ G32, _Float16 *__restrict__ G16,
                               int64_t n, int NT, float inv_scale, int linear) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (linear) {  // thread-major order (gq_split): the same index in both copies
      G16[i] = (_Float16)(G32[i] * inv_scale);
      continue;
    }
    const int t = (int)(i % NT);
    const int64_t r = i / NT;
    const int q = (int)(r % NQ2);
    const int64_t r2 = r / NQ2;
    const int c = (int)(r2 % NC);
    const int tile = (int)(r2 / NC);
    G16[gq_index<_Float16, NQ2, NC>(tile, c, q, t, NT)] =
        (_Float16)(G32[gq_index<float, NQ2, NC>(tile, c, q, t, NT)] * inv_scale);
  }
}

// --------------------------------------------------------------------------
// geometry precompute (FP64 arithmetic, stored as T)
// --------------------------------------------------------------------------
// G_q = kappa_q w_q det(J_q) J_q^-1 J_q^-T with J the Jacobian of the map
// from the element's reference square [0,1]^2 (the reference's parametric
// Jacobian times diag(hx,hy), matrix_free/src/ElementMF.cpp:159-206).
// On general rational meshes (per-function weights wts, plain B-spline
// tables) R_A = w_A B_A / W, and the kernels apply the operator to w o x:
// G_q is stored divided by W_q^2 and components 3, 4 hold a = grad W / W.
// One thread per tile slot; slots past the mesh edge get zeros.
//
// The map at one quadrature point: physical point, reference-square
// Jacobian and W, grad W (W = 1 without wts).
template <int P, int NQ>
__device__ __forceinline__ void map_q(const double *cp, const double *wts, int nfx, int ex, int ey,
                                      const double *t1, const double *t2, int i, int j,
                                      double (&m)[9]) {
  constexpr int NB = P + 1, TAB = NQ * NB;
  double x = 0, y = 0, xxi = 0, xeta = 0, yxi = 0, yeta = 0, W = 0, Wxi = 0, Weta = 0;
  for (int b = 0; b < NB; ++b)
    for (int a = 0; a < NB; ++a) {
      const size_t A = (size_t)(ey + b) * nfx + ex + a;
      const double R = t1[i * NB + a] * t2[j * NB + b];
      const double dxi = t1[TAB + i * NB + a] * t2[j * NB + b];
      const double deta = t1[i * NB + a] * t2[TAB + j * NB + b];
      double px = cp[2 * A], py = cp[2 * A + 1];
      if (wts) {
        const double w = wts[A];
        px *= w;
        py *= w;
        W += w * R;
        Wxi += w * dxi;
        Weta += w * deta;
      }
      x += px * R;
      y += py * R;
      xxi += px * dxi;
      xeta += px * deta;
      yxi += py * dxi;
      yeta += py * deta;
    }
  if (wts) {
    // quotient rule for x = (sum w P B) / W
    x /= W;
    y /= W;
    xxi = (xxi - x * Wxi) / W;
    xeta = (xeta - x * Weta) / W;
    yxi = (yxi - y * Wxi) / W;
    yeta = (yeta - y * Weta) / W;
  } else {
    W = 1.0;
  }
  m[0] = x; m[1] = y; m[2] = xxi; m[3] = xeta; m[4] = yxi; m[5] = yeta;
  m[6] = W; m[7] = Wxi; m[8] = Weta;
}

template <typename T, int P, int NQ>
__global__ void k_geometry(T *G, const double *tabx, const double *taby,
                           const double *cp, const double *wts, const double *wq,
                           const double *kappa, int nex, int ney, int nfx,
                           int TX, int TY, int ntx) {
  constexpr int NQ2 = NQ * NQ, TAB = NQ * (P + 1);
  const int NT = TX * TY;
  const int tile = blockIdx.x;
  const bool split = gq_split(sizeof(T), TX, TY, NQ);
  auto put = [&](int c, int q, int t, double v) {
    if (wts)
      G[gq_at<T, NQ, 5>(tile, c, q, t, NT, split)] = (T)v;
    else
      G[gq_at<T, NQ, 3>(tile, c, q, t, NT, split)] = (T)v;
  };
  const int nc = wts ? 5 : 3;
  for (int t = threadIdx.x; t < NT; t += blockDim.x) {
    const int tx = tile % ntx, ty = tile / ntx;
    const int ex = tx * TX + t % TX, ey = ty * TY + t / TX;
    if (ex >= nex || ey >= ney) {
      for (int q = 0; q < NQ2; ++q)
        for (int c = 0; c < nc; ++c) put(c, q, t, 0.0);
      continue;
    }
    const double *t1 = tabx + (size_t)ex * 2 * TAB;
    const double *t2 = taby + (size_t)ey * 2 * TAB;
    for (int j = 0; j < NQ; ++j)
      for (int i = 0; i < NQ; ++i) {
        double m[9];
        map_q<P, NQ>(cp, wts, nfx, ex, ey, t1, t2, i, j, m);
        const double xxi = m[2], xeta = m[3], yxi = m[4], yeta = m[5];
        const double det = xxi * yeta - xeta * yxi;
        const int q = j * NQ + i;
        double s = wq[i] * wq[j] / det;
        if (kappa) s *= kappa[((size_t)ey * nex + ex) * NQ2 + q];
        if (wts) s /= m[6] * m[6];
        put(0, q, t, s * (xeta * xeta + yeta * yeta));
        put(1, q, t, -s * (xxi * xeta + yxi * yeta));
        put(2, q, t, s * (xxi * xxi + yxi * yxi));
        if (wts) {
          put(3, q, t, m[7] / m[6]);
          put(4, q, t, m[8] / m[6]);
        }
      }
  }
}

// --------------------------------------------------------------------------
// load vector (FP64): fJ[e][q] = f(x_q) w_q det(J_q), then a per-function
// gather over the adjacent elements (matrix_free_gpu_dbd/src/mult.cu:233-339
// is the per-function form of the reference's load kernel).
// --------------------------------------------------------------------------
template <int P, int NQ>
// fq: values of f at the quadrature points ([E][NQ*NQ], reference order),
// used instead of kind / value when not null
__global__ void k_load_qp(double *fJ, const double *tabx, const double *taby,
                          const double *cp, const double *wts, const double *wq, int nex,
                          int ney, int nfx, int kind, double value, const double *fq) {
  constexpr int NB = P + 1, NQ2 = NQ * NQ, TAB = NQ * NB;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nex * ney) return;
  const int ex = e % nex, ey = e / nex;
  const double *t1 = tabx + (size_t)ex * 2 * TAB;
  const double *t2 = taby + (size_t)ey * 2 * TAB;
  const double pi = 3.14159265358979323846264338327950288;
  for (int j = 0; j < NQ; ++j)
    for (int i = 0; i < NQ; ++i) {
      double m[9];
      map_q<P, NQ>(cp, wts, nfx, ex, ey, t1, t2, i, j, m);
      const double x = m[0], y = m[1];
      double f;
      if (fq) f = fq[(size_t)e * NQ2 + j * NQ + i];
      else if (kind == 0) f = x * (1.0 - x) * y * (1.0 - y);
      else if (kind == 1) f = value;
      else f = value * sin(pi * x) * sin(pi * y);
      // general rational: R_A = w_A B_A / W, the 1/W here, w_A in the gather
      double v = f * wq[i] * wq[j] * (m[2] * m[5] - m[3] * m[4]);
      if (wts) v /= m[6];
      fJ[(size_t)e * NQ2 + j * NQ + i] = v;
    }
}

// physical coordinates of every quadrature point: xq[e][q][2] (FP64)
template <int P, int NQ>
__global__ void k_qp_coords(double *xq, const double *tabx, const double *taby,
                            const double *cp, const double *wts, int nex, int ney, int nfx) {
  constexpr int NB = P + 1, NQ2 = NQ * NQ, TAB = NQ * NB;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nex * ney) return;
  const int ex = e % nex, ey = e / nex;
  const double *t1 = tabx + (size_t)ex * 2 * TAB;
  const double *t2 = taby + (size_t)ey * 2 * TAB;
  for (int j = 0; j < NQ; ++j)
    for (int i = 0; i < NQ; ++i) {
      double m[9];
      map_q<P, NQ>(cp, wts, nfx, ex, ey, t1, t2, i, j, m);
      const size_t o = ((size_t)e * NQ2 + j * NQ + i) * 2;
      xq[o] = m[0];
      xq[o + 1] = m[1];
    }
}

template <int P, int NQ>
__global__ void k_load_gather(double *F, const double *fJ, const double *tabx,
                              const double *taby, const double *wts, const uint8_t *dirmask,
                              int nex, int ney, int nfx, int nfy) {
  constexpr int NB = P + 1, NQ2 = NQ * NQ, TAB = NQ * NB;
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= nfx * nfy) return;
  const int gx = A % nfx, gy = A / nfx;
  if (is_dirichlet(dirmask, A, gx, gy, nfx, nfy)) {
    F[A] = 0.0;
    return;
  }
  double acc = 0.0;
  for (int b = P; b >= 0; --b) {
    const int ey = gy - b;
    if (ey < 0 || ey >= ney) continue;
    for (int a = P; a >= 0; --a) {
      const int ex = gx - a;
      if (ex < 0 || ex >= nex) continue;
      const double *t1 = tabx + (size_t)ex * 2 * TAB;
      const double *t2 = taby + (size_t)ey * 2 * TAB;
      const double *f = fJ + ((size_t)ey * nex + ex) * NQ2;
      double s = 0.0;
      for (int j = 0; j < NQ; ++j)
        for (int i = 0; i < NQ; ++i) s += t1[i * NB + a] * t2[j * NB + b] * f[j * NQ + i];
      acc += s;
    }
  }
  F[A] = wts ? acc * wts[A] : acc;
}

// --------------------------------------------------------------------------
// grad_kappa[e][q] = -grad x_q^T G0_q grad (P_int g)_q   (FP64)
// --------------------------------------------------------------------------
template <int P, int NQ>
__global__ void k_kappa_grad(double *out, const double *x, const double *g,
                             const double *G0, const double *tabx,
                             const double *taby, const double *wts, const uint8_t *dirmask,
                             int nex, int ney, int nfx, int nfy, int TX, int TY,
                             int ntx) {
  constexpr int NB = P + 1, NQ2 = NQ * NQ, TAB = NQ * NB;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nex * ney) return;
  const int ex = e % nex, ey = e / nex;
  const int NT = TX * TY;
  const int tile = (ey / TY) * ntx + ex / TX, t = (ey % TY) * TX + ex % TX;
  const double *t1 = tabx + (size_t)ex * 2 * TAB;
  const double *t2 = taby + (size_t)ey * 2 * TAB;
  auto gq = [&](int c, int q) {
    return wts ? G0[gq_index<double, NQ2, 5>(tile, c, q, t, NT)]
               : G0[gq_index<double, NQ2, 3>(tile, c, q, t, NT)];
  };
  double xe[NB][NB], ge[NB][NB];
  for (int b = 0; b < NB; ++b)
    for (int a = 0; a < NB; ++a) {
      const int gx = ex + a, gy = ey + b, A = gy * nfx + gx;
      xe[b][a] = x[A];
      ge[b][a] = is_dirichlet(dirmask, A, gx, gy, nfx, nfy) ? 0.0 : g[A];
      if (wts) {
        xe[b][a] *= wts[A];
        ge[b][a] *= wts[A];
      }
    }
  for (int j = 0; j < NQ; ++j)
    for (int i = 0; i < NQ; ++i) {
      double ux = 0, uy = 0, vx = 0, vy = 0, uv = 0, vv = 0;
      for (int b = 0; b < NB; ++b)
        for (int a = 0; a < NB; ++a) {
          const double dxi = t1[TAB + i * NB + a] * t2[j * NB + b];
          const double deta = t1[i * NB + a] * t2[TAB + j * NB + b];
          ux += dxi * xe[b][a];
          uy += deta * xe[b][a];
          vx += dxi * ge[b][a];
          vy += deta * ge[b][a];
          if (wts) {
            const double R = t1[i * NB + a] * t2[j * NB + b];
            uv += R * xe[b][a];
            vv += R * ge[b][a];
          }
        }
      const int q = j * NQ + i;
      if (wts) {
        const double ax = gq(3, q), ay = gq(4, q);
        ux -= uv * ax;
        uy -= uv * ay;
        vx -= vv * ax;
        vy -= vv * ay;
      }
      const double a11 = gq(0, q), a12 = gq(1, q), a22 = gq(2, q);
      out[(size_t)e * NQ2 + q] = -(ux * (a11 * vx + a12 * vy) + uy * (a12 * vx + a22 * vy));
    }
}

// --------------------------------------------------------------------------
// diagonal of A (FP64): diag[A] = -sum_{e,a} sum_q gradR_a^T G_q gradR_a,
// gathered per function in a fixed element order; Dirichlet rows get 1.
// --------------------------------------------------------------------------
template <int P, int NQ>
__global__ void k_diag(double *diag, const double *G, const double *tabx,
                       const double *taby, const double *wts, const uint8_t *dirmask, int nex,
                       int ney, int nfx, int nfy, int TX, int TY, int ntx) {
  constexpr int NB = P + 1, NQ2 = NQ * NQ, TAB = NQ * NB;
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= nfx * nfy) return;
  const int gx = A % nfx, gy = A / nfx;
  if (is_dirichlet(dirmask, A, gx, gy, nfx, nfy)) {
    diag[A] = 1.0;
    return;
  }
  const int NT = TX * TY;
  double acc = 0.0;
  for (int b = P; b >= 0; --b) {
    const int ey = gy - b;
    if (ey < 0 || ey >= ney) continue;
    for (int a = P; a >= 0; --a) {
      const int ex = gx - a;
      if (ex < 0 || ex >= nex) continue;
      const double *t1 = tabx + (size_t)ex * 2 * TAB;
      const double *t2 = taby + (size_t)ey * 2 * TAB;
      const int tile = (ey / TY) * ntx + ex / TX, t = (ey % TY) * TX + ex % TX;
      auto gq = [&](int c, int q) {
        return wts ? G[gq_index<double, NQ2, 5>(tile, c, q, t, NT)]
                   : G[gq_index<double, NQ2, 3>(tile, c, q, t, NT)];
      };
      for (int j = 0; j < NQ; ++j)
        for (int i = 0; i < NQ; ++i) {
          const int q = j * NQ + i;
          double dx = t1[TAB + i * NB + a] * t2[j * NB + b];
          double dy = t1[i * NB + a] * t2[TAB + j * NB + b];
          if (wts) {
            const double R = t1[i * NB + a] * t2[j * NB + b];
            dx -= R * gq(3, q);
            dy -= R * gq(4, q);
          }
          acc -= gq(0, q) * dx * dx + 2.0 * gq(1, q) * dx * dy + gq(2, q) * dy * dy;
        }
    }
  }
  diag[A] = wts ? acc * wts[A] * wts[A] : acc;
}

}  // namespace igak
