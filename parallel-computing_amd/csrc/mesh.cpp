// mesh.cpp -- in-memory preprocessing of a tensor-product NURBS patch: knot
// vectors, Bezier extraction, IEN / ID / Dir, control points and weights.
//
// Replaces the reference's preprocessing chain (matrix_free/preprocess.cpp:
// 16-98 -> IENGenerator, IDGenerator, ControlPointGenerator,
// NURBSExtractionGenerator, Partition::GeneratePartitionSerial) without the
// file round trip (FileManager writes 16 significant digits,
// matrix_free/src/FileManager.cpp:46, which is not FP64 round-trip).
//
// Extraction operators are computed from blossoms (polar forms): the Bezier
// coefficient k of B-spline N_i on element [a,b] is the blossom of N_i at
// (a^(p-k), b^k), evaluated by de Boor's recurrence with one argument per
// level.  Geometry coefficients of polynomial maps (the Greville points of
// the identity map, the homogeneous quarter circle) come from Marsden's
// identity: coefficient i is the blossom at (t_{i+1}, ..., t_{i+p}).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "iga_mf.h"

struct iga_mesh {
  int p, q, nex, ney;
  std::vector<double> cp, w, C1, C2, h1, h2;
  std::vector<int> ien, id, dir;
  bool has_w = false;
};

namespace {

std::vector<double> open_uniform(int p, int nelem, double L) {
  std::vector<double> S;
  S.reserve(nelem + 2 * p + 1);
  const double h = L / nelem;
  for (int i = 0; i <= p; ++i) S.push_back(0.0);
  for (int i = 1; i < nelem; ++i) S.push_back(i * h);
  for (int i = 0; i <= p; ++i) S.push_back(L);
  return S;
}

// Blossom of the spline sum_i d[i] N_i restricted to knot span `span`
// (S[span] <= u < S[span+1]) at arguments v[0..p-1].
double blossom(int p, const std::vector<double> &S, int span, const double *d,
               const double *v) {
  double c[16];
  for (int r = 0; r <= p; ++r) c[r] = d[span - p + r];
  for (int lev = 1; lev <= p; ++lev) {
    for (int r = p; r >= lev; --r) {
      const int i = span - p + r;
      const double den = S[i + p + 1 - lev] - S[i];
      const double a = den != 0.0 ? (v[lev - 1] - S[i]) / den : 0.0;
      c[r] = (1.0 - a) * c[r - 1] + a * c[r];
    }
  }
  return c[p];
}

// C[e][r][k]: Bernstein coefficient k of the r-th nonzero B-spline on element e.
std::vector<double> extraction(int p, const std::vector<double> &S) {
  const int nfunc = (int)S.size() - p - 1;
  const int nelem = nfunc - p;
  const int a = p + 1;
  std::vector<double> C((size_t)nelem * a * a, 0.0);
  std::vector<double> unit(nfunc, 0.0);
  double v[16];
  for (int e = 0; e < nelem; ++e) {
    const int span = p + e;  // open knots with simple interior knots
    const double lo = S[span], hi = S[span + 1];
    for (int r = 0; r <= p; ++r) {
      const int i = span - p + r;
      unit[i] = 1.0;
      for (int k = 0; k <= p; ++k) {
        for (int t = 0; t < p; ++t) v[t] = (t < p - k) ? lo : hi;
        C[(size_t)e * a * a + r * a + k] = blossom(p, S, span, unit.data(), v);
      }
      unit[i] = 0.0;
    }
  }
  return C;
}

// Marsden coefficients of the linear map u -> u (Greville abscissae).
std::vector<double> greville(int p, const std::vector<double> &S) {
  const int n = (int)S.size() - p - 1;
  std::vector<double> g(n);
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 1; j <= p; ++j) s += S[i + j];
    g[i] = s / p;
  }
  return g;
}

uint64_t splitmix64(uint64_t &st) {
  uint64_t z = (st += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double u01(uint64_t &st) {
  return ((double)(splitmix64(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

void topology(iga_mesh &m) {
  const int p = m.p, q = m.q, nfx = m.nex + p, nfy = m.ney + q;
  const int nloc = (p + 1) * (q + 1);
  m.ien.resize((size_t)m.nex * m.ney * nloc);
  for (int ey = 0; ey < m.ney; ++ey)
    for (int ex = 0; ex < m.nex; ++ex) {
      int *row = &m.ien[((size_t)ey * m.nex + ex) * nloc];
      for (int b = 0; b <= q; ++b)
        for (int a = 0; a <= p; ++a) row[b * (p + 1) + a] = (ey + b) * nfx + ex + a;
    }
  m.id.resize((size_t)nfx * nfy);
  m.dir.clear();
  for (int j = 0; j < nfy; ++j)
    for (int i = 0; i < nfx; ++i) {
      const int A = j * nfx + i;
      const bool ring = i == 0 || j == 0 || i == nfx - 1 || j == nfy - 1;
      m.id[A] = ring ? -1 : A;
      if (ring) m.dir.push_back(A);
    }
}

}  // namespace

extern "C" {

int iga_rng_normal(uint64_t seed, int64_t n, double *out) {
  if (!out || n < 0) return IGA_ERR_INVALID_ARGUMENT;
  uint64_t st = seed;
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < n; i += 2) {
    const double u1 = u01(st), u2 = u01(st);
    const double r = std::sqrt(-2.0 * std::log(u1));
    out[i] = r * std::cos(two_pi * u2);
    if (i + 1 < n) out[i + 1] = r * std::sin(two_pi * u2);
  }
  return IGA_OK;
}

int iga_rng_uniform(uint64_t seed, int64_t n, double lo, double hi, double *out) {
  if (!out || n < 0) return IGA_ERR_INVALID_ARGUMENT;
  uint64_t st = seed;
  for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * u01(st);
  return IGA_OK;
}

int iga_mesh_build(const iga_mesh_params *P, iga_mesh **out) {
  if (!P || !out) return IGA_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  const int p = P->p;
  if (p < 1 || p > 8) return IGA_ERR_INVALID_DEGREE;
  if (P->nelem_x < p + 2 || P->nelem_y < p + 2) return IGA_ERR_TOO_FEW_ELEMENTS;
  iga_mesh *m = new (std::nothrow) iga_mesh;
  if (!m) return IGA_ERR_OUT_OF_MEMORY;
  m->p = m->q = p;
  m->nex = P->nelem_x;
  m->ney = P->nelem_y;
  const int nfx = m->nex + p, nfy = m->ney + p;
  const double L = (P->kind == IGA_GEOM_QUARTER_ANNULUS) ? 1.0 : P->L;
  if (!(L > 0.0)) { delete m; return IGA_ERR_INVALID_ARGUMENT; }
  const std::vector<double> S = open_uniform(p, m->nex, L);
  const std::vector<double> T = open_uniform(p, m->ney, L);
  m->C1 = extraction(p, S);
  m->C2 = extraction(p, T);
  m->h1.assign(m->nex, L / m->nex);
  m->h2.assign(m->ney, L / m->ney);
  topology(*m);
  m->cp.resize(2 * (size_t)nfx * nfy);
  const std::vector<double> gx = greville(p, S), gy = greville(p, T);
  switch (P->kind) {
    case IGA_GEOM_UNIT_SQUARE:
      for (int j = 0; j < nfy; ++j)
        for (int i = 0; i < nfx; ++i) {
          m->cp[2 * ((size_t)j * nfx + i)] = gx[i];
          m->cp[2 * ((size_t)j * nfx + i) + 1] = gy[j];
        }
      break;
    case IGA_GEOM_QUARTER_ANNULUS: {
      if (p < 2 || !(P->r1 > P->r0) || !(P->r0 > 0.0)) {
        delete m;
        return p < 2 ? IGA_ERR_INVALID_DEGREE : IGA_ERR_INVALID_ARGUMENT;
      }
      // rational quadratic arc: homogeneous control points (x w, y w, w)
      const double s = std::sqrt(0.5);
      const double Q[3][3] = {{1.0, 0.0, 1.0}, {s, s, s}, {0.0, 1.0, 1.0}};
      m->has_w = true;
      m->w.resize((size_t)nfx * nfy);
      const int npairs = p * (p - 1) / 2;
      for (int j = 0; j < nfy; ++j) {
        // degree-p blossom of the quadratic = mean of its degree-2 blossom
        // over all pairs of the p arguments t_{j+1..j+p}
        double hcoef[3] = {0.0, 0.0, 0.0};
        for (int a = 1; a <= p; ++a)
          for (int b = a + 1; b <= p; ++b) {
            const double u = T[j + a], v = T[j + b];
            const double c0 = (1 - u) * (1 - v), c1 = u * (1 - v) + v * (1 - u), c2 = u * v;
            for (int k = 0; k < 3; ++k) hcoef[k] += c0 * Q[0][k] + c1 * Q[1][k] + c2 * Q[2][k];
          }
        for (int k = 0; k < 3; ++k) hcoef[k] /= npairs;
        const double wj = hcoef[2], cx = hcoef[0] / wj, cy = hcoef[1] / wj;
        for (int i = 0; i < nfx; ++i) {
          const double r = P->r0 + (P->r1 - P->r0) * gx[i];
          m->cp[2 * ((size_t)j * nfx + i)] = r * cx;
          m->cp[2 * ((size_t)j * nfx + i) + 1] = r * cy;
          m->w[(size_t)j * nfx + i] = wj;
        }
      }
      break;
    }
    case IGA_GEOM_PERTURBED_SQUARE: {
      uint64_t st = P->seed;
      const double hx = L / m->nex, hy = L / m->ney;
      for (int j = 0; j < nfy; ++j)
        for (int i = 0; i < nfx; ++i) {
          double x = gx[i], y = gy[j];
          if (i > 0 && j > 0 && i < nfx - 1 && j < nfy - 1) {
            x += P->perturb * hx * (2.0 * u01(st) - 1.0);
            y += P->perturb * hy * (2.0 * u01(st) - 1.0);
          }
          m->cp[2 * ((size_t)j * nfx + i)] = x;
          m->cp[2 * ((size_t)j * nfx + i) + 1] = y;
        }
      if (P->w_hi > P->w_lo) {
        m->has_w = true;
        m->w.resize((size_t)nfx * nfy);
        for (size_t k = 0; k < m->w.size(); ++k) m->w[k] = P->w_lo + (P->w_hi - P->w_lo) * u01(st);
      }
      break;
    }
    default:
      delete m;
      return IGA_ERR_INVALID_ARGUMENT;
  }
  *out = m;
  return IGA_OK;
}

int iga_mesh_describe(const iga_mesh *m, iga_mesh_desc *d) {
  if (!m || !d) return IGA_ERR_INVALID_ARGUMENT;
  std::memset(d, 0, sizeof(*d));
  d->p = m->p;
  d->q = m->q;
  d->nelem_x = m->nex;
  d->nelem_y = m->ney;
  d->nqp_x = m->p + 1;
  d->nqp_y = m->q + 1;
  d->cp = m->cp.data();
  d->weights = m->has_w ? m->w.data() : nullptr;
  d->ien = m->ien.data();
  d->id = m->id.data();
  d->dir = m->dir.data();
  d->ndir = (int)m->dir.size();
  d->extraction_x = m->C1.data();
  d->extraction_y = m->C2.data();
  d->elem_size_x = m->h1.data();
  d->elem_size_y = m->h2.data();
  d->kappa = nullptr;
  return IGA_OK;
}

int iga_mesh_destroy(iga_mesh *m) {
  delete m;
  return IGA_OK;
}

}  // extern "C"
