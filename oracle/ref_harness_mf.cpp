// ref_harness_mf.cpp -- C entry points over the reference's own CPU element
// numerics (matrix_free/src/*.cpp compiled from /root/reference by
// oracle/Makefile into oracle/_ref/libref_mf.so).  TEST INFRASTRUCTURE ONLY:
// it exists to produce golden vectors and to check the oracle restatement.
//
// What runs reference code: BSplineBasis, BernsteinBasis, QuadraturePoint,
// NURBSExtractionGenerator, IENGenerator, IDGenerator, AbscissaeGenerator,
// ElementMF::GenerateElement, ElementMFSF::GenerateElement,
// FileManager::WritePartition / ReadPartition.
// What is restated here because its translation unit needs PETSc:
//   - the element contraction of matrix_free/src/LocalAssemblyMF.cpp:31-67,
//   - the element load of LocalAssemblyMF.cpp:3-29,
//   - the SF element matrix of matrix_free/src/LocalAssemblyMFSF.cpp:43-102,
//   - the gather / scatter / Dirichlet steps of
//     matrix_free/src/GlobalAssemblyMF.cpp:141-210 (PETSc Vec calls),
//   - the serial renumbering of matrix_free_gpu/src/Partition.cpp:258-269.
#include <algorithm>
#include <cstring>
#include <vector>

#include "AbscissaeGenerator.hpp"
#include "BSplineBasis.hpp"
#include "BernsteinBasis.hpp"
#include "ElementMF.hpp"
#include "ElementMFSF.hpp"
#include "FileManager.hpp"
#include "IDGenerator.hpp"
#include "IENGenerator.hpp"
#include "NURBSExtractionGenerator.hpp"
#include "QuadraturePoint.hpp"

namespace {

std::vector<double> open_knots(int p, int nelem, double L) {
  // matrix_free/preprocess.cpp:50-56
  std::vector<double> S;
  const double h = L / nelem;
  for (int i = 0; i < p + 1; ++i) S.push_back(0.0);
  for (int i = 1; i < nelem; ++i) S.push_back(i * h);
  for (int i = 0; i < p + 1; ++i) S.push_back(L);
  return S;
}

}  // namespace

extern "C" {

int ref_quadrature(int n, double *x, double *w) {
  QuadraturePoint qp(n, 0, 1);
  std::vector<double> a = qp.GetQuadraturePoint(), b = qp.GetWeight();
  for (int i = 0; i < n; ++i) { x[i] = a[i]; w[i] = b[i]; }
  return 0;
}

int ref_bernstein_table(int p, int nq, double *B, double *dB) {
  QuadraturePoint qp(nq, 0, 1);
  BernsteinBasis bb(p);
  bb.GenerateBernsteinBasis(&qp);
  const std::vector<double> &b = bb.GetBernsteinBasis();
  const std::vector<double> &d = bb.GetBernsteinBasisDerivative();
  std::memcpy(B, b.data(), sizeof(double) * b.size());
  std::memcpy(dB, d.data(), sizeof(double) * d.size());
  return (int)b.size();
}

int ref_open_knots(int p, int nelem, double L, double *S) {
  std::vector<double> s = open_knots(p, nelem, L);
  std::memcpy(S, s.data(), sizeof(double) * s.size());
  return (int)s.size();
}

int ref_extraction(int p, int nelem, double L, double *C) {
  BSplineBasis basis(p, open_knots(p, nelem, L));
  NURBSExtractionGenerator neg;
  std::vector<double> c = neg.GenerateExtraction1D(&basis);
  std::memcpy(C, c.data(), sizeof(double) * c.size());
  return (int)c.size();
}

int ref_bspline_basis(int p, int nelem, double L, double u, double *N) {
  BSplineBasis basis(p, open_knots(p, nelem, L));
  int span = basis.FindSpan(u);
  std::vector<double> n = basis.BasisFuns(u, span);
  std::memcpy(N, n.data(), sizeof(double) * n.size());
  return span;
}

int ref_greville(int p, int nelem, double L, double *g) {
  AbscissaeGenerator ag;
  std::vector<double> a = ag.GrevilleAbscissae(open_knots(p, nelem, L), p);
  std::memcpy(g, a.data(), sizeof(double) * a.size());
  return (int)a.size();
}

// IEN and the serial ID / Dir of a p x q patch.
int ref_topology(int p, int q, int nex, int ney, int *IEN, int *ID, int *Dir) {
  BSplineBasis b1(p, open_knots(p, nex, 1.0)), b2(q, open_knots(q, ney, 1.0));
  IENGenerator ig;
  IDGenerator idg;
  std::vector<int> ien = ig.GenerateIEN2D(&b1, &b2);
  std::vector<int> id = idg.GenerateID2D(&b1, &b2);
  std::memcpy(IEN, ien.data(), sizeof(int) * ien.size());
  int nd = 0;
  for (size_t i = 0; i < id.size(); ++i) {
    if (id[i] == -1) { ID[i] = -1; Dir[nd++] = (int)i; }
    else ID[i] = (int)i;
  }
  return nd;
}

// Per-qp R, dR/dx, dR/dy, JxW of one element via ElementMF::GenerateElement
// (matrix_free/src/ElementMF.cpp:250-296).
int ref_element_qp(int p, int q, const double *C1e, const double *C2e,
                   double hx, double hy, const double *eCP, double *R,
                   double *dRdx, double *dRdy, double *JxW) {
  const int nloc = (p + 1) * (q + 1), nqp = (p + 1) * (q + 1);
  QuadraturePoint q1(p + 1, 0, 1), q2(q + 1, 0, 1);
  BernsteinBasis b1(p), b2(q);
  b1.GenerateBernsteinBasis(&q1);
  b2.GenerateBernsteinBasis(&q2);
  ElementMF el(p, q);
  el.SetElement(std::vector<double>(C1e, C1e + (p + 1) * (p + 1)),
                std::vector<double>(C2e, C2e + (q + 1) * (q + 1)), hx, hy);
  el.GenerateElement(&q1, &q2, &b1, &b2,
                     std::vector<double>(eCP, eCP + 2 * nloc));
  for (int k = 0; k < nqp; ++k) {
    for (int l = 0; l < nloc; ++l) {
      R[k * nloc + l] = el.get_R(k, l);
      dRdx[k * nloc + l] = el.get_dR_dx(k, l);
      dRdy[k * nloc + l] = el.get_dR_dy(k, l);
    }
    JxW[k] = el.get_JxW(k);
  }
  return nqp;
}

// y = P_int sum_e K_e x_e with the reference's element numerics.
int ref_apply_mf(int p, int q, int nex, int ney, int nfunc, const double *CP,
                 const int *IEN, const int *ID, const int *Dir, int ndir,
                 const double *C1, const double *C2, const double *h1,
                 const double *h2, const double *x, double *y) {
  const int nloc = (p + 1) * (q + 1), nqp = nloc;
  QuadraturePoint q1(p + 1, 0, 1), q2(q + 1, 0, 1);
  BernsteinBasis b1(p), b2(q);
  b1.GenerateBernsteinBasis(&q1);
  b2.GenerateBernsteinBasis(&q2);
  ElementMF el(p, q);
  std::vector<double> eCP(2 * nloc), xin(nloc), out(nloc), tx(nqp), ty(nqp);
  for (int i = 0; i < nfunc; ++i) y[i] = 0.0;
  for (int ey = 0; ey < ney; ++ey)
    for (int ex = 0; ex < nex; ++ex) {
      const int e = ey * nex + ex;
      for (int l = 0; l < nloc; ++l) {
        const int A = IEN[e * nloc + l];
        eCP[2 * l] = CP[2 * A];
        eCP[2 * l + 1] = CP[2 * A + 1];
        xin[l] = x[A];
      }
      el.SetElement(std::vector<double>(C1 + ex * (p + 1) * (p + 1),
                                        C1 + (ex + 1) * (p + 1) * (p + 1)),
                    std::vector<double>(C2 + ey * (q + 1) * (q + 1),
                                        C2 + (ey + 1) * (q + 1) * (q + 1)),
                    h1[ex], h2[ey]);
      el.GenerateElement(&q1, &q2, &b1, &b2, eCP);
      for (int k = 0; k < nqp; ++k) {
        tx[k] = 0.0; ty[k] = 0.0;
        for (int l = 0; l < nloc; ++l) {
          tx[k] += el.get_dR_dx(k, l) * xin[l];
          ty[k] += el.get_dR_dy(k, l) * xin[l];
        }
      }
      for (int k = 0; k < nqp; ++k) {
        tx[k] *= -el.get_JxW(k);
        ty[k] *= -el.get_JxW(k);
      }
      for (int l = 0; l < nloc; ++l) out[l] = 0.0;
      for (int k = 0; k < nqp; ++k)
        for (int l = 0; l < nloc; ++l) {
          out[l] += el.get_dR_dx(k, l) * tx[k];
          out[l] += el.get_dR_dy(k, l) * ty[k];
        }
      for (int l = 0; l < nloc; ++l) {
        const int I = ID[IEN[e * nloc + l]];
        if (I >= 0) y[I] += out[l];
      }
    }
  for (int d = 0; d < ndir; ++d) y[Dir[d]] = 0.0;
  return 0;
}

// mult_sf semantics: ElementMFSF::GenerateElement (matrix_free/src/
// ElementMFSF.cpp:237-297) feeding the element matrix of
// LocalAssemblyMFSF.cpp:43-102.
int ref_apply_sf(int p, int q, int nex, int ney, int nfunc, const double *CP,
                 const int *IEN, const int *ID, const int *Dir, int ndir,
                 const double *C1, const double *C2, const double *h1,
                 const double *h2, const double *x, double *y) {
  const int nloc = (p + 1) * (q + 1);
  const int nqp1 = p + 1, nqp2 = q + 1;
  QuadraturePoint q1(p + 1, 0, 1), q2(q + 1, 0, 1);
  const std::vector<double> qw1 = q1.GetWeight(), qw2 = q2.GetWeight();
  std::vector<int> map;
  for (int i = 0; i < p + 1; ++i)
    for (int j = 0; j < q + 1; ++j) { map.push_back(j); map.push_back(i); }
  ElementMFSF el(p, q);
  std::vector<double> eCP(2 * nloc), xin(nloc), K(nloc * nloc);
  std::vector<double> B1, B2, dB1, dB2, W, J, dWx, dWy;
  for (int i = 0; i < nfunc; ++i) y[i] = 0.0;
  for (int ey = 0; ey < ney; ++ey)
    for (int ex = 0; ex < nex; ++ex) {
      const int e = ey * nex + ex;
      for (int l = 0; l < nloc; ++l) {
        const int A = IEN[e * nloc + l];
        eCP[2 * l] = CP[2 * A];
        eCP[2 * l + 1] = CP[2 * A + 1];
        xin[l] = x[A];
      }
      el.SetElement(std::vector<double>(C1 + ex * (p + 1) * (p + 1),
                                        C1 + (ex + 1) * (p + 1) * (p + 1)),
                    std::vector<double>(C2 + ey * (q + 1) * (q + 1),
                                        C2 + (ey + 1) * (q + 1) * (q + 1)),
                    h1[ex], h2[ey]);
      el.GenerateElement(&q1, &q2, eCP, B1, B2, dB1, dB2, W, J, dWx, dWy);
      for (int l = 0; l < nloc * nloc; ++l) K[l] = 0.0;
      for (int i = 0; i < nloc; ++i) {
        const int i1 = map[i * 2], i2 = map[i * 2 + 1];
        for (int j = 0; j < nloc; ++j) {
          const int j1 = map[j * 2], j2 = map[j * 2 + 1];
          for (int ii = 0; ii < nqp1; ++ii) {
            double t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0, t6 = 0, t7 = 0;
            for (int jj = 0; jj < nqp2; ++jj) {
              const int r = ii * nqp2 + jj;
              const double w2 = W[r] * W[r], w3 = w2 * W[r];
              t1 += qw2[jj] * B2[nqp2 * jj + i2] * B2[nqp2 * jj + j2] * J[r] / w2;
              t2 += qw2[jj] * dB2[nqp2 * jj + i2] * dB2[nqp2 * jj + j2] * J[r] / w2;
              t3 += qw2[jj] * B2[nqp2 * jj + i2] * B2[nqp2 * jj + j2] * J[r] / w3 * dWx[r];
              t4 += qw2[jj] * B2[nqp2 * jj + i2] * B2[nqp2 * jj + j2] * J[r] / w3 * dWx[r];
              t5 += qw2[jj] * B2[nqp2 * jj + i2] * dB2[nqp2 * jj + j2] * J[r] / w3 * dWy[r];
              t6 += qw2[jj] * dB2[nqp2 * jj + i2] * B2[nqp2 * jj + j2] * J[r] / w3 * dWy[r];
              t7 += qw2[jj] * B2[nqp2 * jj + i2] * B2[nqp2 * jj + j2] *
                    (dWx[r] * dWx[r] + dWy[r] * dWy[r]) * J[r] / w2;
            }
            K[i * nloc + j] -= qw1[ii] * (dB1[nqp1 * ii + i1] * dB1[nqp1 * ii + j1] * t1 +
                                          B1[nqp1 * ii + i1] * B1[nqp1 * ii + j1] * t2 -
                                          B1[nqp1 * ii + i1] * dB1[nqp1 * ii + j1] * t3 -
                                          dB1[nqp1 * ii + i1] * B1[nqp1 * ii + j1] * t4 -
                                          B1[nqp1 * ii + i1] * B1[nqp1 * ii + j1] * t5 -
                                          B1[nqp1 * ii + i1] * B1[nqp1 * ii + j1] * t6 -
                                          B1[nqp1 * ii + i1] * B1[nqp1 * ii + j1] * t7);
          }
        }
      }
      for (int i = 0; i < nloc; ++i) {
        double s = 0.0;
        for (int j = 0; j < nloc; ++j) s += K[i * nloc + j] * xin[j];
        const int I = ID[IEN[e * nloc + i]];
        if (I >= 0) y[I] += s;
      }
    }
  for (int d = 0; d < ndir; ++d) y[Dir[d]] = 0.0;
  return 0;
}

// Load vector with the reference's f = x(1-x)y(1-y)
// (matrix_free/include/LocalAssemblyMF.hpp:52-55).
int ref_load(int p, int q, int nex, int ney, int nfunc, const double *CP,
             const int *IEN, const int *ID, const int *Dir, int ndir,
             const double *C1, const double *C2, const double *h1,
             const double *h2, double *F) {
  const int nloc = (p + 1) * (q + 1), nqp = nloc;
  QuadraturePoint q1(p + 1, 0, 1), q2(q + 1, 0, 1);
  BernsteinBasis b1(p), b2(q);
  b1.GenerateBernsteinBasis(&q1);
  b2.GenerateBernsteinBasis(&q2);
  ElementMF el(p, q);
  std::vector<double> eCP(2 * nloc), Fl(nloc);
  for (int i = 0; i < nfunc; ++i) F[i] = 0.0;
  for (int ey = 0; ey < ney; ++ey)
    for (int ex = 0; ex < nex; ++ex) {
      const int e = ey * nex + ex;
      for (int l = 0; l < nloc; ++l) {
        const int A = IEN[e * nloc + l];
        eCP[2 * l] = CP[2 * A];
        eCP[2 * l + 1] = CP[2 * A + 1];
      }
      el.SetElement(std::vector<double>(C1 + ex * (p + 1) * (p + 1),
                                        C1 + (ex + 1) * (p + 1) * (p + 1)),
                    std::vector<double>(C2 + ey * (q + 1) * (q + 1),
                                        C2 + (ey + 1) * (q + 1) * (q + 1)),
                    h1[ex], h2[ey]);
      el.GenerateElement(&q1, &q2, &b1, &b2, eCP);
      for (int l = 0; l < nloc; ++l) Fl[l] = 0.0;
      for (int k = 0; k < nqp; ++k) {
        double X = 0.0, Y = 0.0;
        for (int l = 0; l < nloc; ++l) {
          X += el.get_R(k, l) * eCP[2 * l];
          Y += el.get_R(k, l) * eCP[2 * l + 1];
        }
        const double f = X * (1 - X) * Y * (1 - Y);
        for (int l = 0; l < nloc; ++l) Fl[l] += el.get_R(k, l) * f * el.get_JxW(k);
      }
      for (int l = 0; l < nloc; ++l) {
        const int I = ID[IEN[e * nloc + l]];
        if (I >= 0) F[I] += Fl[l];
      }
    }
  for (int d = 0; d < ndir; ++d) F[Dir[d]] = 0.0;
  return 0;
}

// FileManager::WritePartition (full form) of the given arrays
int ref_write_partition(const char *fn, int nfunc, int nex, int ney, const double *h1,
                        const double *h2, const double *cp, const int *id, const int *dir,
                        int ndir, const int *ien, int nien, const double *c1, int nc1,
                        const double *c2, int nc2) {
  FileManager fm;
  fm.WritePartition(fn, nfunc, nex, ney, std::vector<double>(h1, h1 + nex),
                    std::vector<double>(h2, h2 + ney), std::vector<double>(cp, cp + 2 * nfunc),
                    std::vector<int>(id, id + nfunc), std::vector<int>(),
                    std::vector<int>(dir, dir + ndir), std::vector<int>(ien, ien + nien),
                    std::vector<double>(c1, c1 + nc1), std::vector<double>(c2, c2 + nc2));
  return 0;
}

// FileManager::ReadPartition (full form) into caller arrays of capacity cap
// each; counts[11] = nfunc, nex, ney, then the lengths read of h1, h2, cp,
// id, dir, ien, c1, c2 (-1 when one exceeds cap).
int ref_read_partition(const char *fn, int *counts, double *h1, double *h2, double *cp, int *id,
                       int *dir, int *ien, double *c1, double *c2, int cap) {
  FileManager fm;
  int nfunc = 0, nex = 0, ney = 0;
  std::vector<double> vh1, vh2, vcp, vc1, vc2;
  std::vector<int> vid, vghost, vdir, vien;
  fm.ReadPartition(fn, nfunc, nex, ney, vh1, vh2, vcp, vid, vghost, vdir, vien, vc1, vc2);
  const size_t sz[] = {vh1.size(), vh2.size(), vcp.size(), vid.size(), vdir.size(),
                       vien.size(), vc1.size(), vc2.size()};
  for (size_t k : sz)
    if (k > (size_t)cap) return -1;
  counts[0] = nfunc;
  counts[1] = nex;
  counts[2] = ney;
  for (int k = 0; k < 8; ++k) counts[3 + k] = (int)sz[k];
  std::copy(vh1.begin(), vh1.end(), h1);
  std::copy(vh2.begin(), vh2.end(), h2);
  std::copy(vcp.begin(), vcp.end(), cp);
  std::copy(vid.begin(), vid.end(), id);
  std::copy(vdir.begin(), vdir.end(), dir);
  std::copy(vien.begin(), vien.end(), ien);
  std::copy(vc1.begin(), vc1.end(), c1);
  std::copy(vc2.begin(), vc2.end(), c2);
  return 0;
}

}  // extern "C"
