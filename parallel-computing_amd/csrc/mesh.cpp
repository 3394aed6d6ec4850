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
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "iga_mf.h"

struct iga_mesh {
  int p, q, nex, ney;
  std::vector<double> cp, w, C1, C2, h1, h2;
  std::vector<int> ien, id, dir;
  bool has_w = false;
};

struct iga_fem_partition {
  int nlocalfunc = 0, nex = 0, ney = 0;
  std::vector<int> id, dir, ien;
  std::vector<double> cp;
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

int iga_mesh_eval(const iga_mesh *m, int64_t n, const double *uv, double *xy) {
  if (!m || n < 0 || (n > 0 && (!uv || !xy))) return IGA_ERR_INVALID_ARGUMENT;
  const int p = m->p, q = m->q, nfx = m->nex + p;
  // element boundaries along each direction: partial sums of the sizes
  std::vector<double> ex(m->nex + 1, 0.0), ey(m->ney + 1, 0.0);
  for (int i = 0; i < m->nex; ++i) ex[i + 1] = ex[i] + m->h1[i];
  for (int j = 0; j < m->ney; ++j) ey[j + 1] = ey[j] + m->h2[j];
  // B-splines of one direction at parameter u: element e and the p+1
  // functions e..e+p, N_a = sum_k C_e[a][k] B_k(t) with the Bernstein
  // polynomials B_k of the local coordinate t (the extraction identity)
  auto basis = [](const std::vector<double> &edge, const std::vector<double> &C, int deg,
                  double u, int &e, double *N) {
    const int ne = (int)edge.size() - 1;
    // the summed element sizes may end a few ulps short of the nominal end
    const double tol = 1e-12 * (edge[ne] - edge[0]);
    if (!(u >= edge[0] - tol && u <= edge[ne] + tol)) return false;
    u = std::min(std::max(u, edge[0]), edge[ne]);
    e = (int)(std::upper_bound(edge.begin(), edge.end(), u) - edge.begin()) - 1;
    if (e >= ne) e = ne - 1;
    const double t = (u - edge[e]) / (edge[e + 1] - edge[e]);
    double B[16];
    B[0] = 1.0;  // de Casteljau-style degree raising of the Bernstein basis
    for (int k = 1; k <= deg; ++k) {
      double prev = 0.0;
      for (int i = 0; i < k; ++i) {
        const double b = B[i];
        B[i] = prev + (1.0 - t) * b;
        prev = t * b;
      }
      B[k] = prev;
    }
    const double *Ce = C.data() + (size_t)e * (deg + 1) * (deg + 1);
    for (int a = 0; a <= deg; ++a) {
      double s = 0.0;
      for (int k = 0; k <= deg; ++k) s += Ce[a * (deg + 1) + k] * B[k];
      N[a] = s;
    }
    return true;
  };
  for (int64_t k = 0; k < n; ++k) {
    double Nx[16], Ny[16];
    int e1 = 0, e2 = 0;
    if (!basis(ex, m->C1, p, uv[2 * k], e1, Nx) || !basis(ey, m->C2, q, uv[2 * k + 1], e2, Ny))
      return IGA_ERR_INVALID_ARGUMENT;
    double X = 0.0, Y = 0.0, W = 0.0;
    for (int b = 0; b <= q; ++b)
      for (int a = 0; a <= p; ++a) {
        const size_t A = (size_t)(e2 + b) * nfx + (e1 + a);
        const double w = Nx[a] * Ny[b] * (m->has_w ? m->w[A] : 1.0);
        X += w * m->cp[2 * A];
        Y += w * m->cp[2 * A + 1];
        W += w;
      }
    xy[2 * k] = X / W;
    xy[2 * k + 1] = Y / W;
  }
  return IGA_OK;
}

}  // extern "C"

namespace {
// The lines of a text file (without line ends); false when it cannot be read.
bool read_lines(const char *filename, std::vector<std::string> &lines) {
  std::FILE *f = std::fopen(filename, "r");
  if (!f) return false;
  std::string text;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
  std::fclose(f);
  lines.clear();
  for (size_t a = 0; a <= text.size();) {
    size_t b = text.find('\n', a);
    if (b == std::string::npos) b = text.size();
    std::string l = text.substr(a, b - a);
    if (!l.empty() && l.back() == '\r') l.pop_back();
    lines.push_back(l);
    a = b + 1;
  }
  return true;
}

// "key: N" on line i
bool header_value(const std::vector<std::string> &lines, size_t i, const char *key, long &v) {
  const size_t k = std::strlen(key);
  if (i >= lines.size() || lines[i].compare(0, k, key) != 0) return false;
  char *end = nullptr;
  v = std::strtol(lines[i].c_str() + k, &end, 10);
  return end != lines[i].c_str() + k;
}

// The whitespace-separated numbers of a string, from offset `from`.
template <typename T>
void numbers(const std::string &l, size_t from, std::vector<T> &v) {
  v.clear();
  const char *c = l.c_str() + (from < l.size() ? from : l.size());
  for (;;) {
    char *end = nullptr;
    const double d = std::strtod(c, &end);
    if (end == c) break;
    v.push_back((T)d);
    c = end;
  }
}

// Named sections after the three header lines of a partition file: a label
// line, then one line of values (FileManager::WritePartition's layout).
std::map<std::string, std::vector<double>> sections(const std::vector<std::string> &lines) {
  std::map<std::string, std::vector<double>> sec;
  for (size_t i = 3; i + 1 < lines.size(); ++i) {
    const std::string &name = lines[i];
    if (name.empty()) continue;
    numbers(lines[i + 1], 0, sec[name]);
    ++i;
  }
  return sec;
}

bool all_finite(const std::vector<double> &v) {
  for (double d : v)
    if (!std::isfinite(d)) return false;
  return true;
}

void to_ints(const std::vector<double> &v, std::vector<int> &o) {
  o.resize(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = (int)v[i];
}
}  // namespace

extern "C" {

int iga_mesh_read_partition(const char *filename, int p, int q, iga_mesh **out) {
  if (!filename || !out) return IGA_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (p < 1 || p > 8 || q < 1 || q > 8) return IGA_ERR_INVALID_DEGREE;
  std::vector<std::string> lines;
  if (!read_lines(filename, lines)) return IGA_ERR_INVALID_ARGUMENT;
  long nfunc = 0, nex = 0, ney = 0;
  if (!header_value(lines, 0, "nlocalfunc: ", nfunc) ||
      !header_value(lines, 1, "nlocalelemx: ", nex) ||
      !header_value(lines, 2, "nlocalelemy: ", ney) || nfunc < 1 || nex < 1 || ney < 1)
    return IGA_ERR_BAD_MESH;
  std::map<std::string, std::vector<double>> sec = sections(lines);
  const char *need[] = {"ID", "Dir", "IEN", "ElemSize1", "ElemSize2", "CP", "NURBSExtraction1",
                        "NURBSExtraction2"};
  for (const char *k : need)
    if (!sec.count(k)) return IGA_ERR_BAD_MESH;
  const size_t nloc = (size_t)(p + 1) * (q + 1);
  if (sec["ID"].size() != (size_t)nfunc || sec["CP"].size() != 2 * (size_t)nfunc ||
      sec["IEN"].size() != (size_t)nex * ney * nloc || sec["ElemSize1"].size() != (size_t)nex ||
      sec["ElemSize2"].size() != (size_t)ney ||
      sec["NURBSExtraction1"].size() != (size_t)nex * (p + 1) * (p + 1) ||
      sec["NURBSExtraction2"].size() != (size_t)ney * (q + 1) * (q + 1))
    return IGA_ERR_BAD_MESH;
  for (const char *k : {"CP", "ElemSize1", "ElemSize2", "NURBSExtraction1", "NURBSExtraction2"})
    if (!all_finite(sec[k])) return IGA_ERR_NOT_FINITE;
  iga_mesh *m = new (std::nothrow) iga_mesh();
  if (!m) return IGA_ERR_OUT_OF_MEMORY;
  m->p = p;
  m->q = q;
  m->nex = (int)nex;
  m->ney = (int)ney;
  m->cp = sec["CP"];
  m->h1 = sec["ElemSize1"];
  m->h2 = sec["ElemSize2"];
  m->C1 = sec["NURBSExtraction1"];
  m->C2 = sec["NURBSExtraction2"];
  to_ints(sec["ID"], m->id);
  to_ints(sec["Dir"], m->dir);
  to_ints(sec["IEN"], m->ien);
  *out = m;
  return IGA_OK;
}

int iga_mesh_write_partition(const iga_mesh *m, const char *filename) {
  if (!m || !filename) return IGA_ERR_INVALID_ARGUMENT;
  if (m->has_w) return IGA_ERR_UNSUPPORTED;  // the format carries no NURBS weights
  std::FILE *f = std::fopen(filename, "w");
  if (!f) return IGA_ERR_INVALID_ARGUMENT;
  std::fprintf(f, "nlocalfunc: %d\nnlocalelemx: %d\nnlocalelemy: %d\n", (int)m->id.size(), m->nex,
               m->ney);
  auto ints = [&](const char *name, const std::vector<int> &v) {
    std::fprintf(f, "%s\n", name);
    for (int x : v) std::fprintf(f, "%d ", x);
    std::fprintf(f, "\n");
  };
  auto reals = [&](const char *name, const std::vector<double> &v) {
    std::fprintf(f, "%s\n", name);
    for (double x : v) std::fprintf(f, "%.16g ", x);
    std::fprintf(f, "\n");
  };
  ints("ID", m->id);
  ints("ghostID", {});
  ints("Dir", m->dir);
  ints("IEN", m->ien);
  reals("ElemSize1", m->h1);
  reals("ElemSize2", m->h2);
  reals("CP", m->cp);
  reals("NURBSExtraction1", m->C1);
  reals("NURBSExtraction2", m->C2);
  return std::fclose(f) == 0 ? IGA_OK : IGA_ERR_INVALID_ARGUMENT;
}

int iga_read_preprocess_info(const char *filename, iga_preprocess_info *info) {
  if (!filename || !info) return IGA_ERR_INVALID_ARGUMENT;
  std::memset(info, 0, sizeof(*info));
  std::FILE *f = std::fopen(filename, "r");
  if (!f) return IGA_ERR_INVALID_ARGUMENT;
  char line[512];
  int seen = 0;
  while (std::fgets(line, sizeof(line), f)) {
    std::string l(line);
    while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
    auto val = [&](const char *key) -> const char * {
      const size_t k = std::strlen(key);
      return l.compare(0, k, key) == 0 ? l.c_str() + k : nullptr;
    };
    const char *v;
    if ((v = val("p: "))) info->p = std::atoi(v), seen |= 1;
    else if ((v = val("q: "))) info->q = std::atoi(v), seen |= 2;
    else if ((v = val("Lx: "))) info->Lx = std::atof(v);
    else if ((v = val("Ly: "))) info->Ly = std::atof(v);
    else if ((v = val("nElemX: "))) info->nelem_x = std::atoi(v), seen |= 4;
    else if ((v = val("nElemY: "))) info->nelem_y = std::atoi(v), seen |= 8;
    else if ((v = val("part_num_1d: "))) info->part_num_1d = std::atoi(v);
    else if ((v = val("dim: "))) info->dim = std::atoi(v);
    else if ((v = val("base_name: ")))
      std::snprintf(info->base_name, sizeof(info->base_name), "%s", v);
  }
  std::fclose(f);
  return seen == 15 ? IGA_OK : IGA_ERR_BAD_MESH;
}

int iga_fem_read_partition(const char *filename, iga_fem_partition **out) {
  if (!filename || !out) return IGA_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  std::vector<std::string> lines;
  if (!read_lines(filename, lines)) return IGA_ERR_INVALID_ARGUMENT;
  long nfunc = 0, nex = 0, ney = 0;
  if (!header_value(lines, 0, "nlocalfunc: ", nfunc) ||
      !header_value(lines, 1, "nlocalelemx: ", nex) ||
      !header_value(lines, 2, "nlocalelemy: ", ney) || nfunc < 1 || nex < 1 || ney < 1)
    return IGA_ERR_BAD_MESH;
  std::map<std::string, std::vector<double>> sec = sections(lines);
  for (const char *k : {"ID", "Dir", "IEN", "CP"})
    if (!sec.count(k)) return IGA_ERR_BAD_MESH;
  // owned functions first, then ghosts: ID and CP cover both
  const size_t nid = sec["ID"].size();
  if (nid < (size_t)nfunc || sec["CP"].size() != 2 * nid ||
      sec["IEN"].size() != (size_t)nex * ney * 4)
    return IGA_ERR_BAD_MESH;
  // the reference's p = 2 preprocessing writes "nan" coordinates (a
  // degree-0 Greville average); its reader then stops at the first one
  if (!all_finite(sec["CP"])) return IGA_ERR_NOT_FINITE;
  iga_fem_partition *fp = new (std::nothrow) iga_fem_partition();
  if (!fp) return IGA_ERR_OUT_OF_MEMORY;
  fp->nlocalfunc = (int)nfunc;
  fp->nex = (int)nex;
  fp->ney = (int)ney;
  to_ints(sec["ID"], fp->id);
  to_ints(sec["Dir"], fp->dir);
  to_ints(sec["IEN"], fp->ien);
  fp->cp = sec["CP"];
  for (int a : fp->ien)
    if (a < 0 || (size_t)a >= nid) {
      delete fp;
      return IGA_ERR_BAD_MESH;
    }
  *out = fp;
  return IGA_OK;
}

int iga_fem_partition_view(const iga_fem_partition *fp, iga_fem_partition_desc *d) {
  if (!fp || !d) return IGA_ERR_INVALID_ARGUMENT;
  d->nlocalfunc = fp->nlocalfunc;
  d->nlocalelemx = fp->nex;
  d->nlocalelemy = fp->ney;
  d->nid = (int)fp->id.size();
  d->ndir = (int)fp->dir.size();
  d->id = fp->id.data();
  d->dir = fp->dir.data();
  d->ien = fp->ien.data();
  d->cp = fp->cp.data();
  return IGA_OK;
}

int iga_fem_partition_destroy(iga_fem_partition *fp) {
  delete fp;
  return IGA_OK;
}

int iga_read_nonzero_coordinate(const char *filename, int64_t cap, int *rows, int *cols,
                                int64_t *nnz) {
  if (!filename || !nnz || cap < 0 || (cap > 0 && (!rows || !cols)))
    return IGA_ERR_INVALID_ARGUMENT;
  std::vector<std::string> lines;
  if (!read_lines(filename, lines)) return IGA_ERR_INVALID_ARGUMENT;
  long n = -1;
  std::vector<long> r, c;
  bool have_r = false, have_c = false;
  for (size_t i = 0; i < lines.size(); ++i) {
    const std::string &l = lines[i];
    if (l.compare(0, 5, "nnz: ") == 0) (void)header_value(lines, i, "nnz: ", n);
    else if (l.compare(0, 6, "rows: ") == 0) numbers(l, 6, r), have_r = true;
    else if (l.compare(0, 6, "cols: ") == 0) numbers(l, 6, c), have_c = true;
  }
  if (n < 0 || !have_r || !have_c || r.size() != c.size() || (long)r.size() != n)
    return IGA_ERR_BAD_MESH;
  *nnz = n;
  if (n > cap) return IGA_ERR_SIZE_LIMIT;
  for (long k = 0; k < n; ++k) {
    rows[k] = (int)r[k];
    cols[k] = (int)c[k];
  }
  return IGA_OK;
}

int iga_write_nonzero_coordinate(const char *filename, int64_t nnz, const int *rows,
                                 const int *cols) {
  if (!filename || nnz < 0 || (nnz > 0 && (!rows || !cols))) return IGA_ERR_INVALID_ARGUMENT;
  std::FILE *f = std::fopen(filename, "w");
  if (!f) return IGA_ERR_INVALID_ARGUMENT;
  std::fprintf(f, "nnz: %lld\nrows: ", (long long)nnz);
  for (int64_t k = 0; k < nnz; ++k) std::fprintf(f, "%d ", rows[k]);
  std::fprintf(f, "\ncols: ");
  for (int64_t k = 0; k < nnz; ++k) std::fprintf(f, "%d ", cols[k]);
  std::fprintf(f, "\n");
  return std::fclose(f) == 0 ? IGA_OK : IGA_ERR_INVALID_ARGUMENT;
}

int iga_read_new_to_old(const char *filename, int64_t cap, int *map, int64_t *n) {
  if (!filename || !n || cap < 0 || (cap > 0 && !map)) return IGA_ERR_INVALID_ARGUMENT;
  std::vector<std::string> lines;
  if (!read_lines(filename, lines)) return IGA_ERR_INVALID_ARGUMENT;
  std::vector<long> v;
  if (!lines.empty()) numbers(lines[0], 0, v);  // the first line only, as the reference
  *n = (int64_t)v.size();
  if ((int64_t)v.size() > cap) return IGA_ERR_SIZE_LIMIT;
  for (size_t i = 0; i < v.size(); ++i) map[i] = (int)v[i];
  return IGA_OK;
}

int iga_write_new_to_old(const char *filename, int64_t n, const int *map) {
  if (!filename || n < 0 || (n > 0 && !map)) return IGA_ERR_INVALID_ARGUMENT;
  std::FILE *f = std::fopen(filename, "w");
  if (!f) return IGA_ERR_INVALID_ARGUMENT;
  for (int64_t i = 0; i < n; ++i) std::fprintf(f, "%d ", map[i]);
  std::fprintf(f, "\n");
  return std::fclose(f) == 0 ? IGA_OK : IGA_ERR_INVALID_ARGUMENT;
}

int iga_fem_pattern_from_files(const char *dir, int nparts, int64_t cap, int *rows, int *cols,
                               int64_t *nnz) {
  if (!dir || !nnz || nparts < 1 || cap < 0 || (cap > 0 && (!rows || !cols)))
    return IGA_ERR_INVALID_ARGUMENT;
  const std::string base(dir);
  auto path = [&](const std::string &name) { return base + "/" + name; };
  std::vector<int> r, c;
  for (int k = 0; k < nparts; ++k) {
    const std::string fn = path("coordinate_nonzero_" + std::to_string(k) + ".txt");
    int64_t n = 0;
    int rc = iga_read_nonzero_coordinate(fn.c_str(), 0, nullptr, nullptr, &n);
    if (rc != IGA_OK && rc != IGA_ERR_SIZE_LIMIT) return rc;
    const size_t at = r.size();
    r.resize(at + n);
    c.resize(at + n);
    rc = iga_read_nonzero_coordinate(fn.c_str(), n, r.data() + at, c.data() + at, &n);
    if (rc) return rc;
  }
  // new -> old numbering: the partitions' map; a serial preprocessing writes
  // none (its numbering is the natural one), so a missing file is the identity
  std::vector<int> map;
  const std::string mfn = path("new_to_old_mapping.txt");
  if (std::FILE *f = std::fopen(mfn.c_str(), "r")) {
    std::fclose(f);
    int64_t m = 0;
    int rc = iga_read_new_to_old(mfn.c_str(), 0, nullptr, &m);
    if (rc != IGA_OK && rc != IGA_ERR_SIZE_LIMIT) return rc;
    map.resize(m);
    rc = iga_read_new_to_old(mfn.c_str(), m, map.data(), &m);
    if (rc) return rc;
  }
  const int64_t total = (int64_t)r.size();
  std::vector<int64_t> key(total);
  const int64_t N = (int64_t)1 << 31;
  for (int64_t k = 0; k < total; ++k) {
    int a = r[k], b = c[k];
    if (!map.empty()) {
      if (a < 0 || b < 0 || (size_t)a >= map.size() || (size_t)b >= map.size())
        return IGA_ERR_BAD_MESH;
      a = map[a];
      b = map[b];
    }
    if (a < 0 || b < 0) return IGA_ERR_BAD_MESH;
    key[k] = (int64_t)a * N + b;
  }
  // lexicographic sort of the (row, col) pairs (duplicates kept, as std::sort
  // of the pairs does)
  std::sort(key.begin(), key.end());
  *nnz = total;
  if (total > cap) return IGA_ERR_SIZE_LIMIT;
  for (int64_t k = 0; k < total; ++k) {
    rows[k] = (int)(key[k] / N);
    cols[k] = (int)(key[k] % N);
  }
  return IGA_OK;
}

}  // extern "C"
