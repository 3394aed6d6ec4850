// mg.hip -- geometric multigrid on the Q1 FEM preconditioner matrix.
//
// The reference preconditions its matrix-free CG with the Q1 stiffness K on
// the control-point-sized node grid, applied by an inner CG + Jacobi to
// 1e-10 (matrix_free_gpu/matrixfreegpu.cu:18-24, 264-293): an approximation
// of K^-1 whose cost grows with the grid.  Here K^-1 is approximated by one
// V-cycle of geometric multigrid on the same matrix -- a fixed, symmetric,
// linear operator, so the outer iteration stays plain PCG -- with the outer
// iteration count of the exact K^-1 and a cost linear in the grid.
//
// The Q1 matrix lives on a structured nx x ny node grid, so every level keeps
// a 9-point stencil per node, structure-of-arrays [9][nx*ny] (offset
// k = (dy+1)*3 + (dx+1)), its couplings to Dirichlet (boundary) nodes dropped
// and a unit centre on boundary rows, whose values stay zero.
//   coarsening  per direction with more than 17 nodes: the even nodes plus the
//               last one; a fine node between two coarse nodes takes half of
//               each (linear interpolation), P = Py (x) Px, restriction P^T;
//   coarse A    Galerkin P^T A P, formed once on the host in FP64;
//   smoother    Chebyshev-Jacobi of degree 2 on [lmax/4, lmax], lmax the
//               Gershgorin bound of D^-1 A (Saad, Iterative Methods for
//               Sparse Linear Systems, Alg. 12.1), before and after the
//               coarse correction (the same polynomial: the cycle is
//               symmetric);
//   coarsest    at most 17 x 17 nodes: its interior system inverted on the host.
// The symmetric stencil is stored by its upper half (centre, E, N, NE, NW:
// 5 values per node instead of 9).  A level takes four kernels: the
// pre-smoothing; the residual, formed where the restriction needs it; the
// coarse correction fused with the first post-smoothing step (the
// interpolated correction formed on the fly); the second step.  Levels with
// many nodes run each over the whole GPU; all levels from NSMALL nodes down
// run in one single-workgroup kernel (their launch latency would exceed their
// work).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "op_internal.hpp"

namespace {

constexpr int BS = 256;          // threads per workgroup of the grid-wide kernels
constexpr int TAIL_BS = 1024;    // threads of the single-workgroup kernel
constexpr int NSMALL = 1200;     // levels with at most this many nodes: tail kernel
constexpr int MAXLEV = 32;
constexpr int NCOARSE = 17;      // coarsen a direction while it has more nodes
constexpr int MAXC = (NCOARSE - 2) * (NCOARSE - 2);  // coarsest interior unknowns

// Level state of one precision (device pointers; values by copy into kernels)
template <typename V>
struct Lev {
  int nx, ny, cnx, cny;  // this grid and its coarse grid (0 at the coarsest)
  int cx, cy;            // directions coarsened to the next level
  const V *st;           // [5][nx*ny] symmetric stencil: centre, E, N, NE, NW
  const V *dinv;         // [nx*ny] inverse centre coefficients (1 on the boundary)
  V *b, *x, *xb, *d, *r; // right-hand side, iterate, second iterate, step, residual
  V inv_theta, c1, c2;   // Chebyshev: 1/theta, and x2 = x1 + c1 d1 + c2 D^-1 r1
};

template <typename V>
struct Tail {
  int l0, nlev;              // first level handled, number of levels
  const V *minv;             // coarsest interior inverse [m][m]
  int m;                     // coarsest interior unknowns
};

constexpr int TX = 32, TY = 8;   // fine tile of the grid-wide kernels (one node per thread)
constexpr int CT = 16;           // coarse tile of the restriction (16 x 16 coarse nodes)
constexpr int FT = 2 * CT + 3;   // its fine footprint, at most 2 CT + 1 plus a node each side

// ---------------------------------------------------------------- per node
__device__ __forceinline__ bool inner(int ix, int iy, int nx, int ny) {
  return ix > 0 && iy > 0 && ix < nx - 1 && iy < ny - 1;
}

// sum_j A_ij v(dx, dy, j) over the 9-point neighbourhood of interior node i.
// A is symmetric and stored by its upper half: the couplings to W, S, SW, SE
// are the E, N, NE, NW entries of those neighbours.  Couplings to boundary
// nodes are zero in both directions.
template <typename V, typename F>
__device__ __forceinline__ V stencil_dot(const Lev<V> &L, int i, F v) {
  const size_t N = (size_t)L.nx * L.ny;
  const int nx = L.nx;
  const V *C = L.st, *E = L.st + N, *Nn = L.st + 2 * N, *NE = L.st + 3 * N, *NW = L.st + 4 * N;
  V s = C[i] * v(0, 0, i);
  s += E[i] * v(1, 0, i + 1) + E[i - 1] * v(-1, 0, i - 1);
  s += Nn[i] * v(0, 1, i + nx) + Nn[i - nx] * v(0, -1, i - nx);
  s += NE[i] * v(1, 1, i + nx + 1) + NE[i - nx - 1] * v(-1, -1, i - nx - 1);
  s += NW[i] * v(-1, 1, i + nx - 1) + NW[i - nx + 1] * v(1, -1, i - nx + 1);
  return s;
}

// pre-smoothing from x0 = 0: x1 = D^-1 b / theta, x2 = (1 + c1) x1 + c2 D^-1 (b - A x1)
template <typename V>
__device__ __forceinline__ void node_pre(const Lev<V> &L, const V *b, V *x, int ix, int iy) {
  const int i = iy * L.nx + ix;
  if (!inner(ix, iy, L.nx, L.ny)) {
    x[i] = V(0);
    return;
  }
  const V ax =
      stencil_dot(L, i, [&](int, int, int j) { return b[j] * L.dinv[j]; }) * L.inv_theta;
  const V di = L.dinv[i];
  const V x1 = b[i] * di * L.inv_theta;
  x[i] = (V(1) + L.c1) * x1 + L.c2 * (b[i] - ax) * di;
}

template <typename V>
__device__ __forceinline__ V residual(const Lev<V> &L, const V *b, const V *x, int ix, int iy) {
  if (!inner(ix, iy, L.nx, L.ny)) return V(0);
  const int i = iy * L.nx + ix;
  return b[i] - stencil_dot(L, i, [&](int, int, int j) { return x[j]; });
}

// 1D restriction: the fine node of coarse node c (n fine, nc coarse) -- its
// own node (even, or the last one); when coarsened the fine-only odd
// neighbours add with weight 1/2
__device__ __forceinline__ int fine_of(int c, int n, int nc, int coarsened) {
  if (!coarsened) return c;
  return (c == nc - 1 && ((n - 1) & 1)) ? n - 1 : 2 * c;
}
__device__ __forceinline__ bool fine_only(int f, int n) {
  return f > 0 && f < n - 1 && (f & 1);
}

// bc at coarse node (cx, cy) from fine residuals r(gx, gy)
template <typename V, typename R>
__device__ __forceinline__ V restrict_at(const Lev<V> &L, int cx, int cy, R r) {
  if (!inner(cx, cy, L.cnx, L.cny)) return V(0);
  const int fx = fine_of(cx, L.nx, L.cnx, L.cx), fy = fine_of(cy, L.ny, L.cny, L.cy);
  V s = V(0);
#pragma unroll
  for (int oy = -1; oy <= 1; ++oy) {
    const int gy = fy + oy;
    V wy;
    if (oy == 0) wy = V(1);
    else if (L.cy && fine_only(gy, L.ny)) wy = V(0.5);
    else continue;
#pragma unroll
    for (int ox = -1; ox <= 1; ++ox) {
      const int gx = fx + ox;
      V wx;
      if (ox == 0) wx = V(1);
      else if (L.cx && fine_only(gx, L.nx)) wx = V(0.5);
      else continue;
      s += wx * wy * r(gx, gy);
    }
  }
  return s;
}

// 1D interpolation: the coarse parents of fine node f and their weights
// (second weight 0 when there is one parent)
__device__ __forceinline__ void parents(int f, int n, int nc, int coarsened, int &c0, int &c1,
                                        float &w0, float &w1) {
  if (!coarsened) {
    c0 = c1 = f;
    w0 = 1.f;
    w1 = 0.f;
  } else if (!(f & 1)) {
    c0 = c1 = f >> 1;
    w0 = 1.f;
    w1 = 0.f;
  } else if (f == n - 1) {
    c0 = c1 = nc - 1;
    w0 = 1.f;
    w1 = 0.f;
  } else {
    c0 = (f - 1) >> 1;
    c1 = (f + 1) >> 1;
    w0 = w1 = 0.5f;
  }
}

// x + P xc at fine node (jx, jy); 0 on the boundary
template <typename V>
__device__ __forceinline__ V corrected(const Lev<V> &L, const V *x, const V *xc, int jx, int jy) {
  if (!inner(jx, jy, L.nx, L.ny)) return V(0);
  int ax0, ax1, ay0, ay1;
  float wx0, wx1, wy0, wy1;
  parents(jx, L.nx, L.cnx, L.cx, ax0, ax1, wx0, wx1);
  parents(jy, L.ny, L.cny, L.cy, ay0, ay1, wy0, wy1);
  const V *r0 = xc + (size_t)ay0 * L.cnx, *r1 = xc + (size_t)ay1 * L.cnx;
  return x[(size_t)jy * L.nx + jx] + V(wx0 * wy0) * r0[ax0] + V(wx1 * wy0) * r0[ax1] +
         V(wx0 * wy1) * r1[ax0] + V(wx1 * wy1) * r1[ax1];
}

// post-smoothing step 1 from x3: d = D^-1 (b - A x3) / theta, xb = x3 + d,
// x3 given by a functor of (dx, dy, j)
template <typename V, typename F>
__device__ __forceinline__ void node_step1(const Lev<V> &L, const V *b, F x3, V *xb, V *d,
                                           int ix, int iy) {
  const int i = iy * L.nx + ix;
  if (!inner(ix, iy, L.nx, L.ny)) {
    xb[i] = V(0);
    d[i] = V(0);
    return;
  }
  const V r = b[i] - stencil_dot(L, i, x3);
  const V di = r * L.dinv[i] * L.inv_theta;
  d[i] = di;
  xb[i] = x3(0, 0, i) + di;
}

// step 2: out = xb + c1 d + c2 D^-1 (b - A xb)
template <typename V>
__device__ __forceinline__ void node_step2(const Lev<V> &L, const V *b, const V *xb, const V *d,
                                           V *out, int ix, int iy) {
  const int i = iy * L.nx + ix;
  if (!inner(ix, iy, L.nx, L.ny)) {
    out[i] = V(0);
    return;
  }
  const V r = b[i] - stencil_dot(L, i, [&](int, int, int j) { return xb[j]; });
  out[i] = xb[i] + L.c1 * d[i] + L.c2 * r * L.dinv[i];
}

__device__ __forceinline__ bool skipped(const int *s1, const int *s2) {
  return (s1 && *s1) || (s2 && *s2);
}

// ---------------------------------------------------------------- grid-wide kernels
// fine kernels: a 32 x 8 tile of nodes per workgroup, grid (ceil(nx/32), ceil(ny/8))
#define MG_TILE_NODE                                            \
  if (skipped(s1, s2)) return;                                  \
  const int ix = blockIdx.x * TX + (threadIdx.x & (TX - 1));    \
  const int iy = blockIdx.y * TY + (threadIdx.x / TX);

template <typename V>
__global__ __launch_bounds__(BS) void k_mg_pre(Lev<V> L, const V *b, V *x, const int *s1,
                                               const int *s2) {
  MG_TILE_NODE
  if (ix >= L.nx || iy >= L.ny) return;
  node_pre(L, b, x, ix, iy);
}

// residual of the fine level and its restriction: a 16 x 16 tile of coarse
// nodes per workgroup; the fine residuals of its footprint go to LDS first
template <typename V>
__global__ __launch_bounds__(BS) void k_mg_resid_restrict(Lev<V> L, const V *b, const V *x,
                                                          V *bc, const int *s1, const int *s2) {
  if (skipped(s1, s2)) return;
  __shared__ V rt[FT * FT];
  const int cx0 = blockIdx.x * CT, cy0 = blockIdx.y * CT;
  const int cx1 = min(cx0 + CT, L.cnx) - 1, cy1 = min(cy0 + CT, L.cny) - 1;
  const int fx0 = max(fine_of(cx0, L.nx, L.cnx, L.cx) - 1, 0);
  const int fy0 = max(fine_of(cy0, L.ny, L.cny, L.cy) - 1, 0);
  const int fx1 = min(fine_of(cx1, L.nx, L.cnx, L.cx) + 1, L.nx - 1);
  const int fy1 = min(fine_of(cy1, L.ny, L.cny, L.cy) + 1, L.ny - 1);
  const int W = fx1 - fx0 + 1, H = fy1 - fy0 + 1;  // at most FT each
  for (int k = threadIdx.x; k < W * H; k += BS) {
    const int ly = k / W, lx = k - ly * W;
    rt[ly * FT + lx] = residual(L, b, x, fx0 + lx, fy0 + ly);
  }
  __syncthreads();
  const int cx = cx0 + (threadIdx.x & (CT - 1)), cy = cy0 + threadIdx.x / CT;
  if (cx > cx1 || cy > cy1) return;
  bc[(size_t)cy * L.cnx + cx] =
      restrict_at(L, cx, cy, [&](int gx, int gy) { return rt[(gy - fy0) * FT + (gx - fx0)]; });
}

// coarse correction and post-smoothing step 1: the corrected iterate of the
// tile and a one-node halo goes to LDS first
template <typename V>
__global__ __launch_bounds__(BS) void k_mg_prolong_step1(Lev<V> L, const V *b, const V *x,
                                                         const V *xc, V *xb, V *d, const int *s1,
                                                         const int *s2) {
  if (skipped(s1, s2)) return;
  constexpr int HX = TX + 2, HY = TY + 2;
  __shared__ V x3[HX * HY];
  const int fx0 = blockIdx.x * TX - 1, fy0 = blockIdx.y * TY - 1;
  for (int k = threadIdx.x; k < HX * HY; k += BS) {
    const int ly = k / HX, lx = k - ly * HX;
    const int jx = fx0 + lx, jy = fy0 + ly;
    x3[k] = (jx < 0 || jy < 0 || jx >= L.nx || jy >= L.ny) ? V(0) : corrected(L, x, xc, jx, jy);
  }
  __syncthreads();
  const int lx = (threadIdx.x & (TX - 1)) + 1, ly = threadIdx.x / TX + 1;
  const int ix = fx0 + lx, iy = fy0 + ly;
  if (ix >= L.nx || iy >= L.ny) return;
  node_step1(L, b, [&](int dx, int dy, int) { return x3[(ly + dy) * HX + lx + dx]; }, xb, d, ix,
             iy);
}

template <typename V>
__global__ __launch_bounds__(BS) void k_mg_step2(Lev<V> L, const V *b, const V *xb, const V *d,
                                                 V *out, const int *s1, const int *s2) {
  MG_TILE_NODE
  if (ix >= L.nx || iy >= L.ny) return;
  node_step2(L, b, xb, d, out, ix, iy);
}

// ---------------------------------------------------------------- the tail
// Levels l0 .. l0+nlev-1 (the last one the coarsest) in one workgroup: the
// same node steps unfused (residual, restriction, correction, smoothing), a
// workgroup barrier between steps.  b_top / x_top: the right-hand side and
// result of level l0 (caller's vectors at level 0).
template <typename V>
__global__ __launch_bounds__(TAIL_BS) void k_mg_tail(const Lev<V> *__restrict__ levs, Tail<V> T,
                                                     const V *b_top, V *x_top, const int *s1,
                                                     const int *s2) {
  if (skipped(s1, s2)) return;
  const int t = threadIdx.x;
  auto bof = [&](int l) -> const V * { return l == T.l0 ? b_top : levs[l].b; };
  const int lc = T.l0 + T.nlev - 1;
  for (int l = T.l0; l < lc; ++l) {
    const Lev<V> L = levs[l];
    const int N = L.nx * L.ny;
    const V *b = bof(l);
    for (int i = t; i < N; i += TAIL_BS) node_pre(L, b, L.x, i % L.nx, i / L.nx);
    __syncthreads();
    for (int i = t; i < N; i += TAIL_BS) L.r[i] = residual(L, b, L.x, i % L.nx, i / L.nx);
    __syncthreads();
    V *bc = levs[l + 1].b;
    for (int i = t; i < L.cnx * L.cny; i += TAIL_BS)
      bc[i] = restrict_at(L, i % L.cnx, i / L.cnx,
                          [&](int gx, int gy) { return L.r[(size_t)gy * L.nx + gx]; });
    __syncthreads();
  }
  {
    // coarsest: x = K_c^-1 b on the interior unknowns (row-major numbering)
    const Lev<V> L = levs[lc];
    const V *b = bof(lc);
    V *x = lc == T.l0 ? x_top : L.x;
    const int N = L.nx * L.ny, mx = L.nx - 2;
    for (int i = t; i < N; i += TAIL_BS) {
      const int ix = i % L.nx, iy = i / L.nx;
      if (!inner(ix, iy, L.nx, L.ny)) {
        x[i] = V(0);
        continue;
      }
      const int u = (iy - 1) * mx + (ix - 1);
      V s = V(0);
      for (int v = 0; v < T.m; ++v) {
        const int jx = v % mx + 1, jy = v / mx + 1;
        s += T.minv[u * T.m + v] * b[jy * L.nx + jx];
      }
      x[i] = s;
    }
    __syncthreads();
  }
  for (int l = lc - 1; l >= T.l0; --l) {
    const Lev<V> L = levs[l];
    const int N = L.nx * L.ny;
    const V *b = bof(l);
    const V *xc = levs[l + 1].x;
    // x3 = x + P xc into r (free here)
    for (int i = t; i < N; i += TAIL_BS) L.r[i] = corrected(L, L.x, xc, i % L.nx, i / L.nx);
    __syncthreads();
    for (int i = t; i < N; i += TAIL_BS)
      node_step1(L, b, [&](int, int, int j) { return L.r[j]; }, L.xb, L.d, i % L.nx, i / L.nx);
    __syncthreads();
    V *out = l == T.l0 ? x_top : L.x;
    for (int i = t; i < N; i += TAIL_BS) node_step2(L, b, L.xb, L.d, out, i % L.nx, i / L.nx);
    __syncthreads();
  }
}

// ---------------------------------------------------------------- host set-up
struct HostLevel {
  int nx = 0, ny = 0, cx = 0, cy = 0, cnx = 0, cny = 0;
  std::vector<double> st;  // [9][N]
  double inv_theta = 0, c1 = 0, c2 = 0;
};

int coarse_count(int n, int coarsened) {
  if (!coarsened) return n;
  return (n - 1) / 2 + 1 + ((n - 1) & 1);
}

// host twin of parents()
int hparents(int f, int n, int nc, int coarsened, int (&c)[2], double (&w)[2]) {
  if (!coarsened) {
    c[0] = f;
    w[0] = 1.0;
    return 1;
  }
  if (!(f & 1)) {
    c[0] = f >> 1;
    w[0] = 1.0;
    return 1;
  }
  if (f == n - 1) {
    c[0] = nc - 1;
    w[0] = 1.0;
    return 1;
  }
  c[0] = (f - 1) >> 1;
  c[1] = (f + 1) >> 1;
  w[0] = w[1] = 0.5;
  return 2;
}

bool is_int(int ix, int iy, int nx, int ny) {
  return ix > 0 && iy > 0 && ix < nx - 1 && iy < ny - 1;
}

// Galerkin P^T A P of level F into level C (structured)
void galerkin(const HostLevel &F, HostLevel &C) {
  const int nx = F.nx, ny = F.ny, cnx = F.cnx, cny = F.cny;
  const size_t N = (size_t)nx * ny, NC = (size_t)cnx * cny;
  C.nx = cnx;
  C.ny = cny;
  C.st.assign(9 * NC, 0.0);
  for (int fy = 1; fy < ny - 1; ++fy)
    for (int fx = 1; fx < nx - 1; ++fx) {
      const size_t f = (size_t)fy * nx + fx;
      int pfx[2], pfy[2];
      double wfx[2], wfy[2];
      const int nfx_ = hparents(fx, nx, cnx, F.cx, pfx, wfx);
      const int nfy_ = hparents(fy, ny, cny, F.cy, pfy, wfy);
      for (int k = 0; k < 9; ++k) {
        const double a = F.st[k * N + f];
        if (a == 0.0) continue;
        const int gx = fx + k % 3 - 1, gy = fy + k / 3 - 1;
        int pgx[2], pgy[2];
        double wgx[2], wgy[2];
        const int ngx = hparents(gx, nx, cnx, F.cx, pgx, wgx);
        const int ngy = hparents(gy, ny, cny, F.cy, pgy, wgy);
        for (int by = 0; by < nfy_; ++by)
          for (int bx = 0; bx < nfx_; ++bx) {
            const int Ix = pfx[bx], Iy = pfy[by];
            if (!is_int(Ix, Iy, cnx, cny)) continue;
            const double wI = wfx[bx] * wfy[by] * a;
            for (int dy = 0; dy < ngy; ++dy)
              for (int dx = 0; dx < ngx; ++dx) {
                const int Jx = pgx[dx], Jy = pgy[dy];
                if (!is_int(Jx, Jy, cnx, cny)) continue;
                const int ox = Jx - Ix, oy = Jy - Iy;  // within [-1, 1] by construction
                C.st[(size_t)((oy + 1) * 3 + ox + 1) * NC + (size_t)Iy * cnx + Ix] +=
                    wI * wgx[dx] * wgy[dy];
              }
          }
      }
    }
  for (int y = 0; y < cny; ++y)
    for (int x = 0; x < cnx; ++x)
      if (!is_int(x, y, cnx, cny)) C.st[4 * NC + (size_t)y * cnx + x] = 1.0;
}

// Make the 9-point stencil exactly symmetric, A_ji := A_ij for j after i
// (the upper half is what the device stores); the two agree to rounding.
void symmetrize(HostLevel &L) {
  const size_t N = (size_t)L.nx * L.ny;
  for (int y = 1; y < L.ny - 1; ++y)
    for (int x = 1; x < L.nx - 1; ++x) {
      const size_t i = (size_t)y * L.nx + x;
      // (dx, dy) in E, N, NE, NW and the opposite entry of that neighbour
      const int off[4][2] = {{1, 0}, {0, 1}, {1, 1}, {-1, 1}};
      for (auto &o : off) {
        const int jx = x + o[0], jy = y + o[1];
        const int k = (o[1] + 1) * 3 + (o[0] + 1), kr = (1 - o[1]) * 3 + (1 - o[0]);
        if (!is_int(jx, jy, L.nx, L.ny)) {
          L.st[k * N + i] = 0.0;
          continue;
        }
        L.st[kr * N + (size_t)jy * L.nx + jx] = L.st[k * N + i];
      }
    }
}

// Chebyshev-Jacobi coefficients of degree 2 on [lmax/4, lmax]
void chebyshev(HostLevel &L) {
  const size_t N = (size_t)L.nx * L.ny;
  double lmax = 0.0;
  for (int y = 1; y < L.ny - 1; ++y)
    for (int x = 1; x < L.nx - 1; ++x) {
      const size_t i = (size_t)y * L.nx + x;
      double s = 0.0;
      for (int k = 0; k < 9; ++k) s += std::fabs(L.st[k * N + i]);
      lmax = std::max(lmax, s / std::fabs(L.st[4 * N + i]));
    }
  const double lmin = lmax / 4.0;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta, rho0 = 1.0 / sigma, rho1 = 1.0 / (2.0 * sigma - rho0);
  L.inv_theta = 1.0 / theta;
  L.c1 = rho1 * rho0;
  L.c2 = 2.0 * rho1 / delta;
}

// dense inverse of the coarsest interior system (Gauss-Jordan, partial pivoting)
bool coarsest_inverse(const HostLevel &L, std::vector<double> &inv, int &m) {
  const int mx = L.nx - 2, my = L.ny - 2;
  m = mx * my;
  if (m < 1 || m > MAXC) return false;
  const size_t N = (size_t)L.nx * L.ny;
  std::vector<double> A((size_t)m * 2 * m, 0.0);
  for (int u = 0; u < m; ++u) {
    const int ix = u % mx + 1, iy = u / mx + 1;
    const size_t i = (size_t)iy * L.nx + ix;
    for (int k = 0; k < 9; ++k) {
      const int jx = ix + k % 3 - 1, jy = iy + k / 3 - 1;
      if (!is_int(jx, jy, L.nx, L.ny)) continue;
      const int v = (jy - 1) * mx + (jx - 1);
      A[(size_t)u * 2 * m + v] += L.st[k * N + i];
    }
    A[(size_t)u * 2 * m + m + u] = 1.0;
  }
  for (int c = 0; c < m; ++c) {
    int piv = c;
    for (int r = c + 1; r < m; ++r)
      if (std::fabs(A[(size_t)r * 2 * m + c]) > std::fabs(A[(size_t)piv * 2 * m + c])) piv = r;
    if (A[(size_t)piv * 2 * m + c] == 0.0) return false;
    for (int k = 0; k < 2 * m; ++k) std::swap(A[(size_t)c * 2 * m + k], A[(size_t)piv * 2 * m + k]);
    const double d = A[(size_t)c * 2 * m + c];
    for (int k = 0; k < 2 * m; ++k) A[(size_t)c * 2 * m + k] /= d;
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      const double f = A[(size_t)r * 2 * m + c];
      if (f == 0.0) continue;
      for (int k = 0; k < 2 * m; ++k) A[(size_t)r * 2 * m + k] -= f * A[(size_t)c * 2 * m + k];
    }
  }
  inv.resize((size_t)m * m);
  for (int u = 0; u < m; ++u)
    for (int v = 0; v < m; ++v) inv[(size_t)u * m + v] = A[(size_t)u * 2 * m + m + v];
  return true;
}

}  // namespace

namespace igai {

template <typename V>
struct MgSet {
  std::vector<Lev<V>> lev;  // host copies of the device level records
  Lev<V> *dlev = nullptr;   // device copies (tail kernel)
  V *minv = nullptr;
};

struct MgHier {
  int nlev = 0, tail0 = 0, m = 0;
  MgSet<double> s64;
  MgSet<float> s32;
  bool has64 = false, has32 = false;
  std::vector<void *> allocs;
  int64_t bytes = 0;
};

void mg_destroy(MgHier *h) {
  if (!h) return;
  for (void *p : h->allocs) (void)hipFree(p);
  delete h;
}

namespace {
template <typename V>
int upload(MgHier *h, MgSet<V> &S, const std::vector<HostLevel> &H,
           const std::vector<double> &minv) {
  auto al = [&](void **p, size_t bytes) -> int {
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return IGA_ERR_OUT_OF_MEMORY;
    }
    h->allocs.push_back(*p);
    h->bytes += (int64_t)bytes;
    return IGA_OK;
  };
  const int L = (int)H.size();
  S.lev.resize(L);
  for (int l = 0; l < L; ++l) {
    const HostLevel &hl = H[l];
    const size_t N = (size_t)hl.nx * hl.ny;
    Lev<V> &d = S.lev[l];
    d.nx = hl.nx;
    d.ny = hl.ny;
    d.cnx = hl.cnx;
    d.cny = hl.cny;
    d.cx = hl.cx;
    d.cy = hl.cy;
    d.inv_theta = (V)hl.inv_theta;
    d.c1 = (V)hl.c1;
    d.c2 = (V)hl.c2;
    V *st = nullptr;
    int rc = al((void **)&st, 5 * N * sizeof(V));
    if (rc) return rc;
    // centre, E, N, NE, NW: offsets (0,0), (1,0), (0,1), (1,1), (-1,1)
    const int keep[5] = {4, 5, 7, 8, 6};
    std::vector<V> tmp(5 * N);
    for (int c = 0; c < 5; ++c)
      for (size_t i = 0; i < N; ++i) tmp[c * N + i] = (V)hl.st[keep[c] * N + i];
    IGA_HIP_TRY(hipMemcpy(st, tmp.data(), 5 * N * sizeof(V), hipMemcpyHostToDevice));
    d.st = st;
    V *dv = nullptr;
    rc = al((void **)&dv, N * sizeof(V));
    if (rc) return rc;
    std::vector<V> dh(N);
    for (size_t i = 0; i < N; ++i) dh[i] = (V)(1.0 / hl.st[4 * N + i]);  // centres nonzero
    IGA_HIP_TRY(hipMemcpy(dv, dh.data(), N * sizeof(V), hipMemcpyHostToDevice));
    d.dinv = dv;
    V **vec[5] = {&d.b, &d.x, &d.xb, &d.d, &d.r};
    for (V **v : vec) {
      rc = al((void **)v, N * sizeof(V));
      if (rc) return rc;
      IGA_HIP_TRY(hipMemset(*v, 0, N * sizeof(V)));
    }
  }
  int rc = al((void **)&S.dlev, L * sizeof(Lev<V>));
  if (rc) return rc;
  IGA_HIP_TRY(hipMemcpy(S.dlev, S.lev.data(), L * sizeof(Lev<V>), hipMemcpyHostToDevice));
  rc = al((void **)&S.minv, std::max<size_t>(1, minv.size()) * sizeof(V));
  if (rc) return rc;
  std::vector<V> mv(minv.begin(), minv.end());
  if (!mv.empty())
    IGA_HIP_TRY(hipMemcpy(S.minv, mv.data(), mv.size() * sizeof(V), hipMemcpyHostToDevice));
  return IGA_OK;
}
}  // namespace

int mg_create(const int *rowptr, const int *col, const double *val, int nx, int ny, bool want64,
              bool want32, MgHier **out) {
  *out = nullptr;
  if (nx < 3 || ny < 3) return IGA_ERR_TOO_FEW_ELEMENTS;
  std::vector<HostLevel> H(1);
  {
    HostLevel &L = H[0];
    L.nx = nx;
    L.ny = ny;
    const size_t N = (size_t)nx * ny;
    L.st.assign(9 * N, 0.0);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const size_t i = (size_t)iy * nx + ix;
        if (!is_int(ix, iy, nx, ny)) {
          L.st[4 * N + i] = 1.0;
          continue;
        }
        for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
          const int j = col[e], jx = j % nx, jy = j / nx;
          const int dx = jx - ix, dy = jy - iy;
          if (dx < -1 || dx > 1 || dy < -1 || dy > 1) return IGA_ERR_BAD_MESH;
          if (!is_int(jx, jy, nx, ny)) continue;  // Dirichlet node: value zero
          L.st[(size_t)((dy + 1) * 3 + dx + 1) * N + i] += val[e];
        }
        if (!(L.st[4 * N + i] != 0.0)) return IGA_ERR_BAD_MESH;
      }
    symmetrize(L);
  }
  for (;;) {
    HostLevel &F = H.back();
    if ((F.nx <= NCOARSE && F.ny <= NCOARSE) || (int)H.size() >= MAXLEV) break;
    F.cx = F.nx > NCOARSE;
    F.cy = F.ny > NCOARSE;
    F.cnx = coarse_count(F.nx, F.cx);
    F.cny = coarse_count(F.ny, F.cy);
    HostLevel C;
    galerkin(F, C);
    symmetrize(C);
    H.push_back(std::move(C));
  }
  for (HostLevel &L : H) chebyshev(L);
  std::vector<double> minv;
  int m = 0;
  if (!coarsest_inverse(H.back(), minv, m)) return IGA_ERR_BAD_MESH;
  MgHier *h = new (std::nothrow) MgHier();
  if (!h) return IGA_ERR_OUT_OF_MEMORY;
  h->nlev = (int)H.size();
  h->m = m;
  h->tail0 = h->nlev - 1;
  while (h->tail0 > 0 && (int64_t)H[h->tail0 - 1].nx * H[h->tail0 - 1].ny <= NSMALL) --h->tail0;
  int rc = IGA_OK;
  if (want64) {
    rc = upload<double>(h, h->s64, H, minv);
    h->has64 = rc == IGA_OK;
  }
  if (!rc && want32) {
    rc = upload<float>(h, h->s32, H, minv);
    h->has32 = rc == IGA_OK;
  }
  if (rc) {
    mg_destroy(h);
    return rc;
  }
  *out = h;
  return IGA_OK;
}

int64_t mg_bytes(const MgHier *h) { return h ? h->bytes : 0; }
int mg_levels(const MgHier *h) { return h ? h->nlev : 0; }

template <typename V>
int mg_apply(MgHier *h, const V *r, V *z, hipStream_t st, const int *s1, const int *s2) {
  MgSet<V> *Sp;
  if constexpr (std::is_same_v<V, double>)
    Sp = &h->s64;
  else
    Sp = &h->s32;
  MgSet<V> &S = *Sp;
  if (S.lev.empty()) return IGA_ERR_UNSUPPORTED;
  const int t0 = h->tail0;
  auto fine = [](const Lev<V> &L) { return dim3((L.nx + TX - 1) / TX, (L.ny + TY - 1) / TY); };
  // down the grid-wide levels
  for (int l = 0; l < t0; ++l) {
    const Lev<V> &L = S.lev[l];
    const V *b = l == 0 ? r : L.b;
    hipLaunchKernelGGL((k_mg_pre<V>), fine(L), dim3(BS), 0, st, L, b, L.x, s1, s2);
    hipLaunchKernelGGL((k_mg_resid_restrict<V>),
                       dim3((L.cnx + CT - 1) / CT, (L.cny + CT - 1) / CT), dim3(BS), 0, st, L, b,
                       L.x, S.lev[l + 1].b, s1, s2);
  }
  Tail<V> T{t0, h->nlev - t0, S.minv, h->m};
  const V *bt = t0 == 0 ? r : S.lev[t0].b;
  V *xt = t0 == 0 ? z : S.lev[t0].x;
  hipLaunchKernelGGL((k_mg_tail<V>), dim3(1), dim3(TAIL_BS), 0, st, S.dlev, T, bt, xt, s1, s2);
  // up the grid-wide levels
  for (int l = t0 - 1; l >= 0; --l) {
    const Lev<V> &L = S.lev[l];
    const V *b = l == 0 ? r : L.b;
    hipLaunchKernelGGL((k_mg_prolong_step1<V>), fine(L), dim3(BS), 0, st, L, b, L.x,
                       S.lev[l + 1].x, L.xb, L.d, s1, s2);
    hipLaunchKernelGGL((k_mg_step2<V>), fine(L), dim3(BS), 0, st, L, b, L.xb, L.d,
                       l == 0 ? z : L.x, s1, s2);
  }
  return hipGetLastError() == hipSuccess ? IGA_OK : IGA_ERR_HIP;
}

template int mg_apply<double>(MgHier *, const double *, double *, hipStream_t, const int *,
                              const int *);
template int mg_apply<float>(MgHier *, const float *, float *, hipStream_t, const int *,
                             const int *);

}  // namespace igai
