/*
 * iga_oracle.c -- CPU FP64 restatement of the reference's matrix-free IGA
 * Poisson path (SeaVegetable/Parallel-Computing, trees matrix_free/,
 * matrix_free_gpu/, matrix_free_gpu_dbd/).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * the timed CPU baseline.  The product never links it.
 *
 * Each function names the reference code it restates.  Pinning (what checks
 * each function against something other than itself) is listed in the test
 * files; functions without such a check are marked "parity unpinned".
 */
#include "iga_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* random numbers                                                            */
/* ------------------------------------------------------------------------ */

/* splitmix64 (Steele, Lea, Flood 2014): the one generator every side of the
 * build uses for synthetic inputs. */
uint64_t orc_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double u01_open(uint64_t *st) {
  /* 53 random bits, shifted by half an ulp: strictly inside (0,1). */
  return ((double)(orc_splitmix64(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

void orc_rng_uniform(uint64_t seed, int64_t n, double lo, double hi,
                     double *out) {
  uint64_t st = seed;
  for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * u01_open(&st);
}

void orc_rng_normal(uint64_t seed, int64_t n, double *out) {
  /* Box-Muller on consecutive uniform pairs. */
  uint64_t st = seed;
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < n; i += 2) {
    double u1 = u01_open(&st), u2 = u01_open(&st);
    double r = sqrt(-2.0 * log(u1));
    out[i] = r * cos(two_pi * u2);
    if (i + 1 < n) out[i + 1] = r * sin(two_pi * u2);
  }
}

/* ------------------------------------------------------------------------ */
/* 1D tools                                                                  */
/* ------------------------------------------------------------------------ */

/* Open uniform knot vector on [0,L]: matrix_free/preprocess.cpp:50-56
 * (p+1 zeros, i*h for i=1..nelem-1, p+1 copies of L).  Returns the length. */
int orc_open_uniform_knots(int p, int nelem, double L, double *S) {
  const double h = L / nelem;
  int k = 0;
  for (int i = 0; i < p + 1; ++i) S[k++] = 0.0;
  for (int i = 1; i < nelem; ++i) S[k++] = i * h;
  for (int i = 0; i < p + 1; ++i) S[k++] = L;
  return k;
}

static double binom(int n, int k) {
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

/* Bernstein polynomials of degree p on [0,1] and their derivatives:
 * the closed forms that matrix_free/src/BernsteinBasis.cpp:3-143 expands
 * case by case.  B_k = C(p,k) xi^k (1-xi)^(p-k),
 * dB_k = p (B^{p-1}_{k-1} - B^{p-1}_k). */
void orc_bernstein(int p, double xi, double *B, double *dB) {
  for (int k = 0; k <= p; ++k)
    B[k] = binom(p, k) * pow(xi, k) * pow(1.0 - xi, p - k);
  for (int k = 0; k <= p; ++k) {
    double lo = (k >= 1) ? binom(p - 1, k - 1) * pow(xi, k - 1) *
                               pow(1.0 - xi, p - k)
                         : 0.0;
    double hi = (k <= p - 1) ? binom(p - 1, k) * pow(xi, k) *
                                   pow(1.0 - xi, p - 1 - k)
                             : 0.0;
    dB[k] = p * (lo - hi);
  }
}

/* Gauss-Legendre rule on [0,1] (weights sum to 1), computed by Newton
 * iteration on P_n.  The reference tabulates the same rule to 16 digits and
 * maps it with x -> (x+1)/2, w -> w/2 (matrix_free/src/QuadraturePoint.cpp:
 * 3-58).  n in [1,10] like the reference. */
int orc_gauss_legendre01(int n, double *x, double *w) {
  if (n < 1 || n > 10) return -1;
  const double pi = 3.14159265358979323846264338327950288;
  for (int i = 0; i < n; ++i) {
    double t = cos(pi * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        double pk = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
      }
      if (n == 1) { p1 = t; p0 = 1.0; }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      double dt = p1 / dp;
      t -= dt;
      if (fabs(dt) < 1e-17) break;
    }
    {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        double pk = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
      }
      if (n == 1) { p1 = t; p0 = 1.0; }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
    }
    double wi = 2.0 / ((1.0 - t * t) * dp * dp);
    /* cos guesses run from +1 to -1; store ascending like the reference. */
    x[n - 1 - i] = 0.5 * t + 0.5;
    w[n - 1 - i] = 0.5 * wi;
  }
  return 0;
}

/* Cox-de Boor B-spline basis at u (The NURBS Book A2.1 FindSpan + A2.2
 * BasisFuns, as in matrix_free/src/BSplineBasis.cpp:3-41).  N[0..p] are the
 * nonzero functions span-p..span.  Returns 0. */
int orc_bspline_basis(int p, const double *S, int nknots, double u, double *N,
                      int *span_out) {
  const int n = nknots - p - 2; /* last function index */
  int span;
  if (u >= S[n + 1]) {
    span = n;
  } else {
    int lo = p, hi = n + 1;
    span = (lo + hi) / 2;
    while (u < S[span] || u >= S[span + 1]) {
      if (u < S[span]) hi = span; else lo = span;
      span = (lo + hi) / 2;
    }
  }
  double left[16], right[16];
  N[0] = 1.0;
  for (int j = 1; j <= p; ++j) {
    left[j] = u - S[span + 1 - j];
    right[j] = S[span + j] - u;
    double saved = 0.0;
    for (int r = 0; r < j; ++r) {
      double tmp = N[r] / (right[r + 1] + left[j - r]);
      N[r] = saved + right[r + 1] * tmp;
      saved = left[j - r] * tmp;
    }
    N[j] = saved;
  }
  if (span_out) *span_out = span;
  return 0;
}

/* Bezier extraction of a univariate B-spline, Borden et al. 2011 IJNME 87,
 * Algorithm 1, as implemented by
 * matrix_free/src/NURBSExtractionGenerator.cpp:4-82 (index map idxC at
 * matrix_free/include/NURBSExtractionGenerator.hpp:14-17).  Output
 * C[e][row][col], row = B-spline, col = Bernstein.  The operator counter and
 * the knot cursor advance after every knot group; the reference advances them
 * only when mult < p, which never terminates for p = 1 (every interior knot
 * then has mult == p); for p >= 2 with simple interior knots the two agree.
 * Returns the number of elements. */
int orc_extraction_1d(int p, const double *S, int nknots, double *C) {
  const int m = nknots;
  const int nfunc = m - p - 1;
  const int nelem = nfunc - p;
  const int a = p + 1;
  memset(C, 0, sizeof(double) * (size_t)a * a * nelem);
#define UK(i) S[(i) - 1]                                /* 1-based knots */
#define CE(r, c, e) C[((e) - 1) * a * a + ((r) - 1) * a + ((c) - 1)]
  for (int i = 1; i <= a; ++i) CE(i, i, 1) = 1.0;
  int ka = p + 1, kb = ka + 1, nb = 1;
  double alphas[32];
  while (kb < m) {
    if (nb + 1 <= nelem)
      for (int i = 1; i <= a; ++i) CE(i, i, nb + 1) = 1.0;
    const int i0 = kb;
    while (kb < m && UK(kb + 1) == UK(kb)) kb++;
    const int mult = kb - i0 + 1;
    if (mult < p) {
      const double numer = UK(kb) - UK(ka);
      for (int j = p; j > mult; --j)
        alphas[j - mult] = numer / (UK(ka + j) - UK(ka));
      const int r = p - mult;
      for (int j = 1; j <= r; ++j) {
        const int save = r - j + 1;
        const int s = mult + j;
        for (int k = p + 1; k > s; --k) {
          const double alpha = alphas[k - s];
          for (int row = 1; row <= a; ++row)
            CE(row, k, nb) = alpha * CE(row, k, nb) +
                             (1.0 - alpha) * CE(row, k - 1, nb);
        }
        if (kb < m && nb + 1 <= nelem)
          for (int off = 0; off <= j; ++off)
            CE(save + off, save, nb + 1) = CE(p - j + 1 + off, p + 1, nb);
      }
    }
    nb++;
    if (nb > nelem) break;
    if (kb < m) { ka = kb; kb++; }
  }
#undef UK
#undef CE
  return nelem;
}

/* Greville abscissae g_i = (S[i+1]+...+S[i+p])/p
 * (matrix_free/src/AbscissaeGenerator.cpp:27-41). */
void orc_greville(int p, const double *S, int nknots, double *g) {
  const int n = nknots - p - 1;
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 1; j <= p; ++j) s += S[i + j];
    g[i] = s / p;
  }
}

/* ------------------------------------------------------------------------ */
/* topology                                                                  */
/* ------------------------------------------------------------------------ */

/* IEN for a tensor-product patch: the loop of
 * matrix_free/src/IENGenerator.cpp:68-103 (functions counted x-fastest,
 * element local order walks jloc = q..0, iloc = p..0 backwards from A). */
void orc_ien_2d(int p, int q, int nex, int ney, int *IEN) {
  const int m = nex + p, n = ney + q;
  int A = 0;
  int64_t k = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      A++;
      if (i >= p && j >= q)
        for (int jloc = q; jloc >= 0; --jloc)
          for (int iloc = p; iloc >= 0; --iloc)
            IEN[k++] = A - iloc - jloc * m - 1;
    }
}

/* ID / Dir of the serial partition: boundary ring -> -1 and listed in Dir
 * (ascending), interior function i -> i.  matrix_free/src/IDGenerator.cpp:
 * 43-67 marks the ring; matrix_free_gpu/src/Partition.cpp:258-269 renumbers
 * the survivors to themselves and collects Dir.  Returns ndir. */
int orc_id_2d(int nfx, int nfy, int *ID, int *Dir) {
  int nd = 0;
  for (int j = 0; j < nfy; ++j)
    for (int i = 0; i < nfx; ++i) {
      const int A = j * nfx + i;
      const int bnd = (i == 0 || j == 0 || i == nfx - 1 || j == nfy - 1);
      ID[A] = bnd ? -1 : A;
      if (bnd) Dir[nd++] = A;
    }
  return nd;
}

/* ------------------------------------------------------------------------ */
/* geometries                                                                */
/* ------------------------------------------------------------------------ */

/* Identity parametrisation of [0,Lx]x[0,Ly]: control points at the Greville
 * abscissae, x-fastest (matrix_free/src/ControlPointGenerator.cpp:3-71 solves
 * for the same points with an Eigen LU of the first element's derivative
 * conditions; the Greville closed form is the exact solution).  W = 1. */
void orc_geometry_unit_square(int p, int q, int nex, int ney, double Lx,
                              double Ly, double *CP, double *W) {
  const int nfx = nex + p, nfy = ney + q;
  double *S = (double *)malloc(sizeof(double) * (nex + 2 * p + 1));
  double *T = (double *)malloc(sizeof(double) * (ney + 2 * q + 1));
  double *gx = (double *)malloc(sizeof(double) * nfx);
  double *gy = (double *)malloc(sizeof(double) * nfy);
  int ns = orc_open_uniform_knots(p, nex, Lx, S);
  int nt = orc_open_uniform_knots(q, ney, Ly, T);
  orc_greville(p, S, ns, gx);
  orc_greville(q, T, nt, gy);
  for (int j = 0; j < nfy; ++j)
    for (int i = 0; i < nfx; ++i) {
      CP[2 * (j * nfx + i)] = gx[i];
      CP[2 * (j * nfx + i) + 1] = gy[j];
      if (W) W[j * nfx + i] = 1.0;
    }
  free(S); free(T); free(gx); free(gy);
}

/* Solve the B-spline collocation system sum_j N_j(tau_k) c_j = f_k at the
 * Greville points for ncomp right-hand sides (banded Gaussian elimination
 * without pivoting; B-spline collocation matrices are totally positive). */
static int collocate(int p, const double *S, int nknots, int ncomp,
                     double *rhs /* [n][ncomp] in, coefficients out */) {
  const int n = nknots - p - 1;
  const int bw = p;
  const int wd = 2 * bw + 1;
  double *A = (double *)calloc((size_t)n * wd, sizeof(double));
  double *tau = (double *)malloc(sizeof(double) * n);
  double Nv[16];
  orc_greville(p, S, nknots, tau);
  for (int k = 0; k < n; ++k) {
    int span;
    orc_bspline_basis(p, S, nknots, tau[k], Nv, &span);
    for (int r = 0; r <= p; ++r) {
      int j = span - p + r;
      int off = j - k + bw;
      if (off < 0 || off >= wd) { free(A); free(tau); return -1; }
      A[(size_t)k * wd + off] = Nv[r];
    }
  }
  for (int c = 0; c < n; ++c) {
    double piv = A[(size_t)c * wd + bw];
    if (piv == 0.0) { free(A); free(tau); return -2; }
    for (int r = c + 1; r <= c + bw && r < n; ++r) {
      double f = A[(size_t)r * wd + (c - r + bw)] / piv;
      if (f == 0.0) continue;
      for (int cc = c; cc <= c + bw && cc < n; ++cc)
        A[(size_t)r * wd + (cc - r + bw)] -= f * A[(size_t)c * wd + (cc - c + bw)];
      for (int t = 0; t < ncomp; ++t) rhs[r * ncomp + t] -= f * rhs[c * ncomp + t];
    }
  }
  for (int c = n - 1; c >= 0; --c) {
    for (int t = 0; t < ncomp; ++t) {
      double s = rhs[c * ncomp + t];
      for (int cc = c + 1; cc <= c + bw && cc < n; ++cc)
        s -= A[(size_t)c * wd + (cc - c + bw)] * rhs[cc * ncomp + t];
      rhs[c * ncomp + t] = s / A[(size_t)c * wd + bw];
    }
  }
  free(A); free(tau);
  return 0;
}

/* Quarter annulus r in [r0,r1], theta in [0,pi/2] (extension: the reference
 * has no weighted geometry).  xi runs radially (degree p, linear map), eta
 * runs along the arc: the exact quarter circle is the rational quadratic
 * Bezier with weights 1, sqrt(2)/2, 1; its homogeneous coordinates are
 * polynomials of degree 2, so they lie in the degree-p spline space of the
 * refined knot vector and are recovered exactly by collocation at the
 * Greville points.  CP = r_i * c_j, W = w_j. */
int orc_geometry_quarter_annulus(int p, int nex, int ney, double r0, double r1,
                                 double *CP, double *W) {
  if (p < 2) return -1;
  const int nfx = nex + p, nfy = ney + p;
  double *S = (double *)malloc(sizeof(double) * (nex + 2 * p + 1));
  double *T = (double *)malloc(sizeof(double) * (ney + 2 * p + 1));
  double *gx = (double *)malloc(sizeof(double) * nfx);
  double *hom = (double *)malloc(sizeof(double) * 3 * nfy);
  double *tau = (double *)malloc(sizeof(double) * nfy);
  int ns = orc_open_uniform_knots(p, nex, 1.0, S);
  int nt = orc_open_uniform_knots(p, ney, 1.0, T);
  orc_greville(p, S, ns, gx);
  orc_greville(p, T, nt, tau);
  const double s2 = sqrt(0.5);
  for (int k = 0; k < nfy; ++k) {
    double t = tau[k];
    double b0 = (1 - t) * (1 - t), b1 = 2 * t * (1 - t), b2 = t * t;
    hom[3 * k + 0] = b0 * 1.0 + b1 * s2 + b2 * 0.0;
    hom[3 * k + 1] = b0 * 0.0 + b1 * s2 + b2 * 1.0;
    hom[3 * k + 2] = b0 * 1.0 + b1 * s2 + b2 * 1.0;
  }
  int rc = collocate(p, T, nt, 3, hom);
  if (rc == 0) {
    for (int j = 0; j < nfy; ++j) {
      const double w = hom[3 * j + 2];
      const double cx = hom[3 * j] / w, cy = hom[3 * j + 1] / w;
      for (int i = 0; i < nfx; ++i) {
        const double r = r0 + (r1 - r0) * gx[i];
        CP[2 * (j * nfx + i)] = r * cx;
        CP[2 * (j * nfx + i) + 1] = r * cy;
        W[j * nfx + i] = w;
      }
    }
  }
  free(S); free(T); free(gx); free(hom); free(tau);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* element numerics                                                          */
/* ------------------------------------------------------------------------ */

/* Per-quadrature-point rational basis, physical gradients and JxW of one
 * element: matrix_free/src/ElementMF.cpp:106-207 driven by :250-296
 * (N = C.B, dN = C.dB/h, tensor product, w = sum N, R = N/w,
 * dR = (dN - dw R)/w, Jacobian from eCP, inverse, JxW = w1 w2 det hx hy with
 * the signed det).  eW (per-function NURBS weights) generalises N_a to
 * W_a N_a; eW == NULL is the reference (all weights one).  Quadrature points
 * y-outer, x-inner; arrays R/dRdx/dRdy are [nqp][nloc]; xq is [nqp][2]. */
void orc_element_qp(int p, int q, int nq1, int nq2, const double *C1e,
                    const double *C2e, double hx, double hy, const double *eCP,
                    const double *eW, double *R, double *dRdx, double *dRdy,
                    double *JxW, double *xq) {
  const int a1 = p + 1, a2 = q + 1, nloc = a1 * a2;
  double g1[10], w1[10], g2[10], w2[10];
  double B[16], dB[16];
  double N1[10][16], dN1[10][16], N2[10][16], dN2[10][16];
  orc_gauss_legendre01(nq1, g1, w1);
  orc_gauss_legendre01(nq2, g2, w2);
  for (int ii = 0; ii < nq1; ++ii) {
    orc_bernstein(p, g1[ii], B, dB);
    for (int r = 0; r < a1; ++r) {
      double s = 0.0, ds = 0.0;
      for (int k = 0; k < a1; ++k) {
        s += C1e[r * a1 + k] * B[k];
        ds += C1e[r * a1 + k] * dB[k];
      }
      N1[ii][r] = s;
      dN1[ii][r] = ds / hx;
    }
  }
  for (int jj = 0; jj < nq2; ++jj) {
    orc_bernstein(q, g2[jj], B, dB);
    for (int r = 0; r < a2; ++r) {
      double s = 0.0, ds = 0.0;
      for (int k = 0; k < a2; ++k) {
        s += C2e[r * a2 + k] * B[k];
        ds += C2e[r * a2 + k] * dB[k];
      }
      N2[jj][r] = s;
      dN2[jj][r] = ds / hy;
    }
  }
  double N[128], dNxi[128], dNeta[128], dRxi[128], dReta[128];
  for (int jj = 0; jj < nq2; ++jj)
    for (int ii = 0; ii < nq1; ++ii) {
      const int qp = jj * nq1 + ii;
      double w = 0.0, dwxi = 0.0, dweta = 0.0;
      for (int j = 0; j < a2; ++j)
        for (int i = 0; i < a1; ++i) {
          const int l = j * a1 + i;
          const double wt = eW ? eW[l] : 1.0;
          N[l] = wt * N1[ii][i] * N2[jj][j];
          dNxi[l] = wt * dN1[ii][i] * N2[jj][j];
          dNeta[l] = wt * N1[ii][i] * dN2[jj][j];
          w += N[l];
          dwxi += dNxi[l];
          dweta += dNeta[l];
        }
      double *Rq = R + qp * nloc;
      for (int l = 0; l < nloc; ++l) {
        Rq[l] = N[l] / w;
        dRxi[l] = (dNxi[l] - dwxi * Rq[l]) / w;
        dReta[l] = (dNeta[l] - dweta * Rq[l]) / w;
      }
      double x = 0, y = 0, xxi = 0, xeta = 0, yxi = 0, yeta = 0;
      for (int l = 0; l < nloc; ++l) {
        x += eCP[2 * l] * Rq[l];
        y += eCP[2 * l + 1] * Rq[l];
        xxi += eCP[2 * l] * dRxi[l];
        xeta += eCP[2 * l] * dReta[l];
        yxi += eCP[2 * l + 1] * dRxi[l];
        yeta += eCP[2 * l + 1] * dReta[l];
      }
      const double det = xxi * yeta - xeta * yxi;
      const double xix = yeta / det, xiy = -xeta / det;
      const double etax = -yxi / det, etay = xxi / det;
      for (int l = 0; l < nloc; ++l) {
        dRdx[qp * nloc + l] = xix * dRxi[l] + etax * dReta[l];
        dRdy[qp * nloc + l] = xiy * dRxi[l] + etay * dReta[l];
      }
      JxW[qp] = w1[ii] * w2[jj] * (det * hx * hy);
      if (xq) { xq[2 * qp] = x; xq[2 * qp + 1] = y; }
    }
}

/* gather the element's control points / weights / extraction operators */
static void element_inputs(const orc_mesh *m, int ex, int ey, int *eIEN,
                           double *eCP, double *eW, const double **C1e,
                           const double **C2e) {
  const int nloc = (m->p + 1) * (m->q + 1);
  const int e = ey * m->nex + ex;
  for (int l = 0; l < nloc; ++l) {
    const int A = m->IEN[(int64_t)e * nloc + l];
    eIEN[l] = A;
    eCP[2 * l] = m->CP[2 * (int64_t)A];
    eCP[2 * l + 1] = m->CP[2 * (int64_t)A + 1];
    eW[l] = m->W ? m->W[A] : 1.0;
  }
  *C1e = m->C1 + (int64_t)ex * (m->p + 1) * (m->p + 1);
  *C2e = m->C2 + (int64_t)ey * (m->q + 1) * (m->q + 1);
}

/* One element's action y_e = K_e x_e:
 * matrix_free/src/LocalAssemblyMF.cpp:31-67 (t = grad R . x_e per qp, scaled
 * by -JxW (times kappa_q when given), then y_a += dR_a/dx t_x + dR_a/dy t_y). */
static void element_action(const orc_mesh *m, int e, const double *R,
                           const double *dRdx, const double *dRdy,
                           const double *JxW, const double *xe, double *ye) {
  const int nloc = (m->p + 1) * (m->q + 1);
  const int nqp = m->nq1 * m->nq2;
  double tx[100], ty[100];
  (void)R;
  for (int qp = 0; qp < nqp; ++qp) {
    tx[qp] = 0.0;
    ty[qp] = 0.0;
    for (int b = 0; b < nloc; ++b) {
      tx[qp] += dRdx[qp * nloc + b] * xe[b];
      ty[qp] += dRdy[qp * nloc + b] * xe[b];
    }
  }
  for (int qp = 0; qp < nqp; ++qp) {
    double s = -JxW[qp];
    if (m->kappa) s *= m->kappa[(int64_t)e * nqp + qp];
    tx[qp] *= s;
    ty[qp] *= s;
  }
  for (int l = 0; l < nloc; ++l) ye[l] = 0.0;
  for (int qp = 0; qp < nqp; ++qp)
    for (int a = 0; a < nloc; ++a) {
      ye[a] += dRdx[qp * nloc + a] * tx[qp];
      ye[a] += dRdy[qp * nloc + a] * ty[qp];
    }
}

/* y = P_int sum_e P_e^T K_e x[IEN_e]: matrix_free/src/GlobalAssemblyMF.cpp:
 * 141-210.  x is gathered through IEN for every function (Dirichlet ones
 * included, :188), y is scattered through ID with negative IDs dropped
 * (:192 with VEC_IGNORE_NEGATIVE_INDICES), then y[Dir] = 0 (:201-205).
 * y is overwritten (the caller's VecSet(y,0) of MyMatMult is folded in). */
void orc_apply_mf(const orc_mesh *m, const double *x, double *y) {
  const int nloc = (m->p + 1) * (m->q + 1);
  const int nqp = m->nq1 * m->nq2;
  int eIEN[128];
  double eCP[256], eW[128], xe[128], ye[128];
  double *R = (double *)malloc(sizeof(double) * nqp * nloc);
  double *dx = (double *)malloc(sizeof(double) * nqp * nloc);
  double *dy = (double *)malloc(sizeof(double) * nqp * nloc);
  double JxW[100];
  for (int i = 0; i < m->nfunc; ++i) y[i] = 0.0;
  for (int ey = 0; ey < m->ney; ++ey)
    for (int ex = 0; ex < m->nex; ++ex) {
      const double *C1e, *C2e;
      element_inputs(m, ex, ey, eIEN, eCP, eW, &C1e, &C2e);
      orc_element_qp(m->p, m->q, m->nq1, m->nq2, C1e, C2e, m->h1[ex],
                     m->h2[ey], eCP, m->W ? eW : NULL, R, dx, dy, JxW, NULL);
      for (int l = 0; l < nloc; ++l) xe[l] = x[eIEN[l]];
      element_action(m, ey * m->nex + ex, R, dx, dy, JxW, xe, ye);
      for (int l = 0; l < nloc; ++l) {
        const int I = m->ID[eIEN[l]];
        if (I >= 0) y[I] += ye[l];
      }
    }
  for (int d = 0; d < m->ndir; ++d) y[m->Dir[d]] = 0.0;
  free(R); free(dx); free(dy);
}

/* The same apply, parallelised with OpenMP over a (p+1)x(q+1) element
 * colouring (elements of one colour share no function), for the CPU
 * baseline.  Sums land in a different order than orc_apply_mf. */
int orc_apply_mf_colored(const orc_mesh *m, const double *x, double *y,
                         int nthreads) {
  const int nloc = (m->p + 1) * (m->q + 1);
  const int nqp = m->nq1 * m->nq2;
  const int cx = m->p + 1, cy = m->q + 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (int i = 0; i < m->nfunc; ++i) y[i] = 0.0;
  for (int col = 0; col < cx * cy; ++col) {
    const int ox = col % cx, oy = col / cx;
    const int nbx = (m->nex - ox + cx - 1) / cx;
    const int nby = (m->ney - oy + cy - 1) / cy;
    const int64_t nb = (int64_t)(nbx > 0 ? nbx : 0) * (nby > 0 ? nby : 0);
#pragma omp parallel
    {
      int eIEN[128];
      double eCP[256], eW[128], xe[128], ye[128], JxW[100];
      double *R = (double *)malloc(sizeof(double) * nqp * nloc);
      double *dx = (double *)malloc(sizeof(double) * nqp * nloc);
      double *dy = (double *)malloc(sizeof(double) * nqp * nloc);
#pragma omp for schedule(static)
      for (int64_t b = 0; b < nb; ++b) {
        const int ex = ox + (int)(b % nbx) * cx;
        const int ey = oy + (int)(b / nbx) * cy;
        const double *C1e, *C2e;
        element_inputs(m, ex, ey, eIEN, eCP, eW, &C1e, &C2e);
        orc_element_qp(m->p, m->q, m->nq1, m->nq2, C1e, C2e, m->h1[ex],
                       m->h2[ey], eCP, m->W ? eW : NULL, R, dx, dy, JxW, NULL);
        for (int l = 0; l < nloc; ++l) xe[l] = x[eIEN[l]];
        element_action(m, ey * m->nex + ex, R, dx, dy, JxW, xe, ye);
        for (int l = 0; l < nloc; ++l) {
          const int I = m->ID[eIEN[l]];
          if (I >= 0) y[I] += ye[l];
        }
      }
      free(R); free(dx); free(dy);
    }
  }
  for (int d = 0; d < m->ndir; ++d) y[m->Dir[d]] = 0.0;
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* The reference's sum-factorised element matrix (mult_sf semantics):
 * matrix_free/src/LocalAssemblyMFSF.cpp:43-102 with the per-qp data of
 * matrix_free/src/ElementMFSF.cpp:101-179,237-297 (1D tables built at the
 * x-direction points for both directions; W, J, dW stored jj-outer but read
 * as [ii*nq2+jj]; the temp3/temp4 and |grad W|^2 terms as written).  Those
 * quirks vanish on identity geometry with unit weights, where this equals
 * orc_apply_mf; off it the reference formula is not the Laplacian, and this
 * restatement is only called on identity geometry.  Returns -1 if p != q or
 * nq1 != nq2 (the reference's indexing assumes both). */
int orc_apply_sf_ref(const orc_mesh *m, const double *x, double *y) {
  const int p = m->p, q = m->q, a1 = p + 1, a2 = q + 1, nloc = a1 * a2;
  const int nq1 = m->nq1, nq2 = m->nq2;
  if (p != q || nq1 != nq2) return -1;
  double g[10], qw[10], Bb[16], dBb[16];
  orc_gauss_legendre01(nq1, g, qw);
  double *Kloc = (double *)malloc(sizeof(double) * nloc * nloc);
  int eIEN[128];
  double eCP[256], eW[128], xe[128];
  double B1[160], dB1[160], B2[160], dB2[160];
  double Wq[100], Jq[100], dWx[100], dWy[100];
  for (int i = 0; i < m->nfunc; ++i) y[i] = 0.0;
  for (int ey = 0; ey < m->ney; ++ey)
    for (int ex = 0; ex < m->nex; ++ex) {
      const double *C1e, *C2e;
      element_inputs(m, ex, ey, eIEN, eCP, eW, &C1e, &C2e);
      const double hx = m->h1[ex], hy = m->h2[ey];
      for (int ii = 0; ii < nq1; ++ii) {
        orc_bernstein(p, g[ii], Bb, dBb);
        for (int r = 0; r < a1; ++r) {
          double s1 = 0, d1 = 0, s2 = 0, d2 = 0;
          for (int k = 0; k < a1; ++k) {
            s1 += C1e[r * a1 + k] * Bb[k];
            d1 += C1e[r * a1 + k] * dBb[k];
            s2 += C2e[r * a2 + k] * Bb[k];
            d2 += C2e[r * a2 + k] * dBb[k];
          }
          B1[ii * a1 + r] = s1;
          dB1[ii * a1 + r] = d1 / hx;
          B2[ii * a2 + r] = s2;
          dB2[ii * a2 + r] = d2 / hy;
        }
      }
      for (int jj = 0; jj < nq2; ++jj)
        for (int ii = 0; ii < nq1; ++ii) {
          double w = 0, wxi = 0, weta = 0;
          double N[128], Nxi[128], Neta[128];
          for (int j = 0; j < a2; ++j)
            for (int i = 0; i < a1; ++i) {
              const int l = j * a1 + i;
              N[l] = B1[ii * a1 + i] * B2[jj * a2 + j];
              Nxi[l] = dB1[ii * a1 + i] * B2[jj * a2 + j];
              Neta[l] = B1[ii * a1 + i] * dB2[jj * a2 + j];
              w += N[l]; wxi += Nxi[l]; weta += Neta[l];
            }
          double xxi = 0, xeta = 0, yxi = 0, yeta = 0;
          for (int l = 0; l < nloc; ++l) {
            const double R = N[l] / w;
            const double Rxi = (Nxi[l] - wxi * R) / w;
            const double Reta = (Neta[l] - weta * R) / w;
            xxi += eCP[2 * l] * Rxi;
            xeta += eCP[2 * l] * Reta;
            yxi += eCP[2 * l + 1] * Rxi;
            yeta += eCP[2 * l + 1] * Reta;
          }
          const double det = xxi * yeta - xeta * yxi;
          const double xix = yeta / det, xiy = -xeta / det;
          const double etax = -yxi / det, etay = xxi / det;
          const int s = jj * nq1 + ii; /* storage order of the reference */
          Wq[s] = w;
          Jq[s] = det * hx * hy;
          dWx[s] = wxi * xix + weta * etax;
          dWy[s] = wxi * xiy + weta * etay;
        }
      for (int l = 0; l < nloc * nloc; ++l) Kloc[l] = 0.0;
      for (int i = 0; i < nloc; ++i) {
        const int i1 = i % a1, i2 = i / a1;
        for (int j = 0; j < nloc; ++j) {
          const int j1 = j % a1, j2 = j / a1;
          for (int ii = 0; ii < nq1; ++ii) {
            double t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0, t6 = 0, t7 = 0;
            for (int jj = 0; jj < nq2; ++jj) {
              const int r = ii * nq2 + jj; /* the reference's read index */
              const double W2 = Wq[r] * Wq[r], W3 = W2 * Wq[r];
              const double b2i = B2[nq2 * jj + i2], b2j = B2[nq2 * jj + j2];
              const double d2i = dB2[nq2 * jj + i2], d2j = dB2[nq2 * jj + j2];
              t1 += qw[jj] * b2i * b2j * Jq[r] / W2;
              t2 += qw[jj] * d2i * d2j * Jq[r] / W2;
              t3 += qw[jj] * b2i * b2j * Jq[r] / W3 * dWx[r];
              t4 += qw[jj] * b2i * b2j * Jq[r] / W3 * dWx[r];
              t5 += qw[jj] * b2i * d2j * Jq[r] / W3 * dWy[r];
              t6 += qw[jj] * d2i * b2j * Jq[r] / W3 * dWy[r];
              t7 += qw[jj] * b2i * b2j *
                    (dWx[r] * dWx[r] + dWy[r] * dWy[r]) * Jq[r] / W2;
            }
            const double b1i = B1[nq1 * ii + i1], b1j = B1[nq1 * ii + j1];
            const double d1i = dB1[nq1 * ii + i1], d1j = dB1[nq1 * ii + j1];
            Kloc[i * nloc + j] -= qw[ii] * (d1i * d1j * t1 + b1i * b1j * t2 -
                                            b1i * d1j * t3 - d1i * b1j * t4 -
                                            b1i * b1j * t5 - b1i * b1j * t6 -
                                            b1i * b1j * t7);
          }
        }
      }
      for (int l = 0; l < nloc; ++l) xe[l] = x[eIEN[l]];
      for (int i = 0; i < nloc; ++i) {
        double s = 0.0;
        for (int j = 0; j < nloc; ++j) s += Kloc[i * nloc + j] * xe[j];
        const int I = m->ID[eIEN[i]];
        if (I >= 0) y[I] += s;
      }
    }
  for (int d = 0; d < m->ndir; ++d) y[m->Dir[d]] = 0.0;
  free(Kloc);
  return 0;
}

/* Load vector F_a = sum_q R_a f(x_q) JxW_q, scattered through ID, F[Dir] = 0:
 * matrix_free/src/LocalAssemblyMF.cpp:3-29 and GlobalAssemblyMF.cpp:26-80.
 * kind 0: the reference's f = x(1-x)y(1-y) (LocalAssemblyMF.hpp:52-55);
 * kind 1: f = value; kind 2: f = value*sin(pi x)sin(pi y). */
static void load_impl(const orc_mesh *m, int kind, double value, const double *fq,
                      double *F);

void orc_load(const orc_mesh *m, int kind, double value, double *F) {
  load_impl(m, kind, value, NULL, F);
}

/* The same with f given at the quadrature points, fq[e][qp] (element and
 * point order of orc_element_qp): AssemLocalLoad's sum with f(x_q) replaced
 * by the given value (extension: the reference hard-codes f). */
void orc_load_qp(const orc_mesh *m, const double *fq, double *F) {
  load_impl(m, 0, 0.0, fq, F);
}

static void load_impl(const orc_mesh *m, int kind, double value, const double *fq,
                      double *F) {
  const int nloc = (m->p + 1) * (m->q + 1);
  const int nqp = m->nq1 * m->nq2;
  const double pi = 3.14159265358979323846264338327950288;
  int eIEN[128];
  double eCP[256], eW[128], JxW[100], xq[200];
  double *R = (double *)malloc(sizeof(double) * nqp * nloc);
  double *dx = (double *)malloc(sizeof(double) * nqp * nloc);
  double *dy = (double *)malloc(sizeof(double) * nqp * nloc);
  for (int i = 0; i < m->nfunc; ++i) F[i] = 0.0;
  for (int ey = 0; ey < m->ney; ++ey)
    for (int ex = 0; ex < m->nex; ++ex) {
      const double *C1e, *C2e;
      element_inputs(m, ex, ey, eIEN, eCP, eW, &C1e, &C2e);
      orc_element_qp(m->p, m->q, m->nq1, m->nq2, C1e, C2e, m->h1[ex],
                     m->h2[ey], eCP, m->W ? eW : NULL, R, dx, dy, JxW, xq);
      double Fe[128];
      for (int l = 0; l < nloc; ++l) Fe[l] = 0.0;
      for (int qp = 0; qp < nqp; ++qp) {
        const double X = xq[2 * qp], Y = xq[2 * qp + 1];
        double f;
        if (fq) f = fq[((int64_t)ey * m->nex + ex) * nqp + qp];
        else if (kind == 0) f = X * (1 - X) * Y * (1 - Y);
        else if (kind == 1) f = value;
        else f = value * sin(pi * X) * sin(pi * Y);
        for (int l = 0; l < nloc; ++l) Fe[l] += R[qp * nloc + l] * f * JxW[qp];
      }
      for (int l = 0; l < nloc; ++l) {
        const int I = m->ID[eIEN[l]];
        if (I >= 0) F[I] += Fe[l];
      }
    }
  for (int d = 0; d < m->ndir; ++d) F[m->Dir[d]] = 0.0;
  free(R); free(dx); free(dy);
}

/* Gradient of g^T A(kappa) u with respect to the per-qp conductivity:
 * grad_kappa[e][q] = -JxW_q grad(u)_q . grad(P_int g)_q (extension: the
 * reference has no kappa; the per-qp chain is the one of orc_apply_mf). */
void orc_kappa_grad(const orc_mesh *m, const double *u, const double *g,
                    double *grad_kappa) {
  const int nloc = (m->p + 1) * (m->q + 1);
  const int nqp = m->nq1 * m->nq2;
  int eIEN[128];
  double eCP[256], eW[128], JxW[100];
  double *R = (double *)malloc(sizeof(double) * nqp * nloc);
  double *dx = (double *)malloc(sizeof(double) * nqp * nloc);
  double *dy = (double *)malloc(sizeof(double) * nqp * nloc);
  for (int ey = 0; ey < m->ney; ++ey)
    for (int ex = 0; ex < m->nex; ++ex) {
      const int e = ey * m->nex + ex;
      const double *C1e, *C2e;
      element_inputs(m, ex, ey, eIEN, eCP, eW, &C1e, &C2e);
      orc_element_qp(m->p, m->q, m->nq1, m->nq2, C1e, C2e, m->h1[ex],
                     m->h2[ey], eCP, m->W ? eW : NULL, R, dx, dy, JxW, NULL);
      for (int qp = 0; qp < nqp; ++qp) {
        double ux = 0, uy = 0, gx = 0, gy = 0;
        for (int l = 0; l < nloc; ++l) {
          const int A = eIEN[l];
          const double gv = (m->ID[A] >= 0) ? g[m->ID[A]] : 0.0;
          ux += dx[qp * nloc + l] * u[A];
          uy += dy[qp * nloc + l] * u[A];
          gx += dx[qp * nloc + l] * gv;
          gy += dy[qp * nloc + l] * gv;
        }
        grad_kappa[(int64_t)e * nqp + qp] = -JxW[qp] * (ux * gx + uy * gy);
      }
    }
  free(R); free(dx); free(dy);
}

/* Element stiffness matrices K_e[a][b] = -sum_q kappa_q JxW_q
 * grad R_a . grad R_b, [E][nloc][nloc]; an assembled-matrix cross-check of
 * the matrix-free apply (the reference's assembled IGA path is
 * matrix_free/src/Element.cpp + LocalAssembly.cpp). */
int64_t orc_element_matrices(const orc_mesh *m, int64_t cap, double *Ke) {
  const int nloc = (m->p + 1) * (m->q + 1);
  const int nqp = m->nq1 * m->nq2;
  const int64_t E = (int64_t)m->nex * m->ney;
  if (cap < E * nloc * nloc) return -(E * nloc * nloc);
  int eIEN[128];
  double eCP[256], eW[128], JxW[100];
  double *R = (double *)malloc(sizeof(double) * nqp * nloc);
  double *dx = (double *)malloc(sizeof(double) * nqp * nloc);
  double *dy = (double *)malloc(sizeof(double) * nqp * nloc);
  for (int ey = 0; ey < m->ney; ++ey)
    for (int ex = 0; ex < m->nex; ++ex) {
      const int64_t e = (int64_t)ey * m->nex + ex;
      const double *C1e, *C2e;
      element_inputs(m, ex, ey, eIEN, eCP, eW, &C1e, &C2e);
      orc_element_qp(m->p, m->q, m->nq1, m->nq2, C1e, C2e, m->h1[ex],
                     m->h2[ey], eCP, m->W ? eW : NULL, R, dx, dy, JxW, NULL);
      double *K = Ke + e * nloc * nloc;
      for (int l = 0; l < nloc * nloc; ++l) K[l] = 0.0;
      for (int qp = 0; qp < nqp; ++qp) {
        double s = -JxW[qp];
        if (m->kappa) s *= m->kappa[e * nqp + qp];
        for (int a = 0; a < nloc; ++a)
          for (int b = 0; b < nloc; ++b)
            K[a * nloc + b] += s * (dx[qp * nloc + a] * dx[qp * nloc + b] +
                                    dy[qp * nloc + a] * dy[qp * nloc + b]);
      }
    }
  free(R); free(dx); free(dy);
  return E * nloc * nloc;
}

/* ------------------------------------------------------------------------ */
/* FEM preconditioner                                                        */
/* ------------------------------------------------------------------------ */

/* Q1 preconditioner nodes: matrix_free_gpu/matrixfreegpu.cu:104-122 (knot
 * vector with p-1 repeated ends and interior knots i*L/(nelem+2), Greville
 * abscissae of degree p-2, x-fastest).  For p = 2 the degree-0 Greville
 * average divides by zero in the reference (all nodes NaN); this restatement
 * then uses the uniform grid of nelem+p nodes on [0,L], which is what the
 * degree-1 case (p = 3) produces.  Returns the node count or -1 (p < 2). */
int orc_fem_nodes(int p, int q, int nex, int ney, double Lx, double Ly,
                  double *CPf) {
  if (p < 2 || q < 2) return -1;
  const int nfx = p + nex, nfy = q + ney;
  double *gx = (double *)malloc(sizeof(double) * nfx);
  double *gy = (double *)malloc(sizeof(double) * nfy);
  for (int dir = 0; dir < 2; ++dir) {
    const int pd = dir ? q : p, ne = dir ? ney : nex, nf = dir ? nfy : nfx;
    const double L = dir ? Ly : Lx;
    double *g = dir ? gy : gx;
    if (pd == 2) {
      for (int i = 0; i < nf; ++i) g[i] = (i == nf - 1) ? L : i * (L / (nf - 1));
      continue;
    }
    const double h = L / (ne + 2);
    double *S = (double *)malloc(sizeof(double) * (ne + 2 * pd + 4));
    int k = 0;
    for (int i = 0; i < pd - 1; ++i) S[k++] = 0.0;
    for (int i = 1; i < ne + 2; ++i) S[k++] = i * h;
    for (int i = 0; i < pd - 1; ++i) S[k++] = L;
    orc_greville(pd - 2, S, k, g);
    free(S);
  }
  for (int j = 0; j < nfy; ++j)
    for (int i = 0; i < nfx; ++i) {
      CPf[2 * (j * nfx + i)] = gx[i];
      CPf[2 * (j * nfx + i) + 1] = gy[j];
    }
  free(gx); free(gy);
  return nfx * nfy;
}

static int cmp_pair(const void *a, const void *b) {
  const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

/* Sparsity pattern of the Q1 preconditioner: every (ID_a, ID_b) of an element
 * with both IDs >= 0, plus (d,d) for each Dirichlet node, sorted
 * lexicographically and unique.  That is what the dry run
 * (matrix_free_gpu/src/GlobalAssemblyDR.cpp:33-94: MatSetValues ignoring
 * negative indices, DirichletBCK, MatGetRowIJ) writes and
 * matrix_free_gpu/matrixfreegpu.cu:149-188 re-sorts.  Q1 IEN and ID follow
 * IENGenerator::GenerateIEN2D(nElemX,nElemY) (matrix_free/src/
 * IENGenerator.cpp:47-66) and the interior renumbering of matrixfreegpu.cu:
 * 127-137.  Brute force: collect, sort, unique.  Returns nnz (negative: the
 * required capacity when cap is too small). */
int64_t orc_fem_pattern(int nfx, int nfy, int64_t cap, int *rows, int *cols) {
  const int nex = nfx - 1, ney = nfy - 1;
  int64_t n = 0;
  int64_t maxn = (int64_t)nex * ney * 16 + (int64_t)nfx * nfy;
  int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * maxn);
  for (int ey = 0; ey < ney; ++ey)
    for (int ex = 0; ex < nex; ++ex) {
      int nodes[4] = {ey * nfx + ex, ey * nfx + ex + 1, (ey + 1) * nfx + ex,
                      (ey + 1) * nfx + ex + 1};
      int ids[4];
      for (int a = 0; a < 4; ++a) {
        const int i = nodes[a] % nfx, j = nodes[a] / nfx;
        ids[a] = (i == 0 || j == 0 || i == nfx - 1 || j == nfy - 1) ? -1 : nodes[a];
      }
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
          if (ids[a] >= 0 && ids[b] >= 0)
            keys[n++] = (int64_t)ids[a] * nfx * nfy + ids[b];
    }
  for (int j = 0; j < nfy; ++j)
    for (int i = 0; i < nfx; ++i)
      if (i == 0 || j == 0 || i == nfx - 1 || j == nfy - 1) {
        const int64_t d = (int64_t)j * nfx + i;
        keys[n++] = d * nfx * nfy + d;
      }
  qsort(keys, n, sizeof(int64_t), cmp_pair);
  int64_t u = 0;
  for (int64_t k = 0; k < n; ++k)
    if (k == 0 || keys[k] != keys[k - 1]) keys[u++] = keys[k];
  if (u > cap) { free(keys); return -u; }
  const int64_t N = (int64_t)nfx * nfy;
  for (int64_t k = 0; k < u; ++k) {
    rows[k] = (int)(keys[k] / N);
    cols[k] = (int)(keys[k] % N);
  }
  free(keys);
  return u;
}

/* elem2coo[e][a][b] = COO position of (ID_a, ID_b), or -1 when either ID is
 * negative: matrix_free_gpu/src/Elem2COOGenerator.cpp:3-55 (row offsets
 * from counts, linear search of the column inside the row). */
void orc_fem_elem2coo(int nfx, int nfy, const int *rows, const int *cols,
                      int64_t nnz, int *elem2coo) {
  const int nex = nfx - 1, ney = nfy - 1;
  const int64_t N = (int64_t)nfx * nfy;
  int64_t *off = (int64_t *)calloc(N + 1, sizeof(int64_t));
  for (int64_t k = 0; k < nnz; ++k) off[rows[k] + 1]++;
  for (int64_t i = 0; i < N; ++i) off[i + 1] += off[i];
  for (int ey = 0; ey < ney; ++ey)
    for (int ex = 0; ex < nex; ++ex) {
      const int64_t e = (int64_t)ey * nex + ex;
      int nodes[4] = {ey * nfx + ex, ey * nfx + ex + 1, (ey + 1) * nfx + ex,
                      (ey + 1) * nfx + ex + 1};
      int ids[4];
      for (int a = 0; a < 4; ++a) {
        const int i = nodes[a] % nfx, j = nodes[a] / nfx;
        ids[a] = (i == 0 || j == 0 || i == nfx - 1 || j == nfy - 1) ? -1 : nodes[a];
      }
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
          int pos = -1;
          if (ids[a] >= 0 && ids[b] >= 0) {
            for (int64_t k = off[ids[a]]; k < off[ids[a] + 1]; ++k)
              if (cols[k] == ids[b]) { pos = (int)k; break; }
          }
          elem2coo[e * 16 + a * 4 + b] = pos;
        }
    }
  free(off);
}

/* dir2coo[i] = position of the first COO entry in row Dir_i, nnz when the
 * row is empty: matrix_free_gpu/src/Elem2COOGenerator.cpp:57-69, which finds
 * it with std::find over rows.  rows is sorted, so the first entry of row d
 * is the lower bound of d (a binary search: the linear scan is O(ndir nnz),
 * minutes at 2048^2). */
int orc_fem_dir2coo(int nfx, int nfy, const int *rows, int64_t nnz,
                    int *dir2coo) {
  int nd = 0;
  for (int j = 0; j < nfy; ++j)
    for (int i = 0; i < nfx; ++i)
      if (i == 0 || j == 0 || i == nfx - 1 || j == nfy - 1) {
        const int d = j * nfx + i;
        int64_t lo = 0, hi = nnz;  /* first k with rows[k] >= d */
        while (lo < hi) {
          const int64_t mid = lo + (hi - lo) / 2;
          if (rows[mid] < d) lo = mid + 1; else hi = mid;
        }
        dir2coo[nd++] = (int)((lo < nnz && rows[lo] == d) ? lo : nnz);
      }
  return nd;
}

/* Q1 stiffness values: matrix_free_gpu/src/assembly.cu:41-99 with the
 * reference element of matrix_free_gpu/src/ElementFEM.cpp and the 2x2 Gauss
 * rule of GlobalAssembly.cu:37-64: val[elem2coo] += -(grad N_a . grad N_b)
 * det w_q, then val[dir2coo] = 1.  Summed sequentially in element, qp, a, b
 * order. */
void orc_fem_values(int nfx, int nfy, const double *CPf, const int *elem2coo,
                    const int *dir2coo, int ndir, int64_t nnz, double *val) {
  const int nex = nfx - 1, ney = nfy - 1;
  double g[2], w[2];
  orc_gauss_legendre01(2, g, w);
  for (int64_t k = 0; k < nnz; ++k) val[k] = 0.0;
  for (int ey = 0; ey < ney; ++ey)
    for (int ex = 0; ex < nex; ++ex) {
      const int64_t e = (int64_t)ey * nex + ex;
      int nodes[4] = {ey * nfx + ex, ey * nfx + ex + 1, (ey + 1) * nfx + ex,
                      (ey + 1) * nfx + ex + 1};
      for (int jj = 0; jj < 2; ++jj)
        for (int ii = 0; ii < 2; ++ii) {
          const double xi = g[ii], eta = g[jj];
          const double N1[2] = {1 - xi, xi}, N2[2] = {1 - eta, eta};
          const double d1[2] = {-1, 1}, d2[2] = {-1, 1};
          double dxi[4], deta[4];
          for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) {
              dxi[j * 2 + i] = d1[i] * N2[j];
              deta[j * 2 + i] = N1[i] * d2[j];
            }
          double xxi = 0, xeta = 0, yxi = 0, yeta = 0;
          for (int a = 0; a < 4; ++a) {
            xxi += CPf[2 * nodes[a]] * dxi[a];
            xeta += CPf[2 * nodes[a]] * deta[a];
            yxi += CPf[2 * nodes[a] + 1] * dxi[a];
            yeta += CPf[2 * nodes[a] + 1] * deta[a];
          }
          const double det = xxi * yeta - xeta * yxi;
          double dx[4], dy[4];
          for (int a = 0; a < 4; ++a) {
            dx[a] = (yeta * dxi[a] - yxi * deta[a]) / det;
            dy[a] = (-xeta * dxi[a] + xxi * deta[a]) / det;
          }
          const double wq = w[ii] * w[jj];
          for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
              const int pos = elem2coo[e * 16 + a * 4 + b];
              if (pos >= 0)
                val[pos] += -(dx[a] * dx[b] + dy[a] * dy[b]) * det * wq;
            }
        }
    }
  for (int i = 0; i < ndir; ++i) val[dir2coo[i]] = 1.0;
}

/* ------------------------------------------------------------------------ */
/* inverse LM                                                                */
/* ------------------------------------------------------------------------ */

/* function -> (element, local index) lists in CSR form:
 * matrix_free_gpu_dbd/src/InvLM.cpp:3-45 (counts, exclusive prefix offsets,
 * fill in element then local order).  Returns the total pair count. */
int64_t orc_invlm(int nloc, int nelem, int nfunc, const int *IEN, int *elemNum,
                  int *offset, int *elemIdx, int *baseIdx) {
  for (int i = 0; i < nfunc; ++i) elemNum[i] = 0;
  for (int64_t k = 0; k < (int64_t)nelem * nloc; ++k) elemNum[IEN[k]]++;
  int64_t s = 0;
  for (int i = 0; i < nfunc; ++i) { offset[i] = (int)s; s += elemNum[i]; }
  int *cnt = (int *)calloc(nfunc, sizeof(int));
  for (int e = 0; e < nelem; ++e)
    for (int l = 0; l < nloc; ++l) {
      const int f = IEN[(int64_t)e * nloc + l];
      elemIdx[offset[f] + cnt[f]] = e;
      baseIdx[offset[f] + cnt[f]] = l;
      cnt[f]++;
    }
  free(cnt);
  return s;
}
