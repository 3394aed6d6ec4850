// ref_harness_gpu.cpp -- C entry points over the reference's FEM-preconditioner
// and inverse-LM host code (matrix_free_gpu/src/{ElementFEM,Elem2COOGenerator,
// AbscissaeGenerator,QuadraturePoint,FileManager,Partition}.cpp and
// matrix_free_gpu_dbd/src/InvLM.cpp, compiled from /root/reference by
// oracle/Makefile into oracle/_ref/libref_gpu.so).  TEST INFRASTRUCTURE ONLY.
//
// Restated here (the reference runs them inside PETSc or as CUDA kernels):
//   - the FEM mesh set-up of matrix_free_gpu/matrixfreegpu.cu:104-137,
//   - the COO pattern that the dry run writes (GlobalAssemblyDR.cpp:33-94),
//   - the value loop of AssembleStiffnessKernel / DirichletBCKKernel
//     (matrix_free_gpu/src/assembly.cu:41-99), run sequentially.
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "AbscissaeGenerator.hpp"
#include "Elem2COOGenerator.hpp"
#include "ElementFEM.hpp"
#include "FileManager.hpp"
#include "IDGenerator.hpp"
#include "IENGenerator.hpp"
#include "InvLM.hpp"
#include "Partition.hpp"
#include "QuadraturePoint.hpp"

namespace {
struct FemMesh {
  std::vector<double> CP;
  std::vector<int> IEN, ID;
  int nfx, nfy, nex, ney;
};

FemMesh fem_mesh(int p, int q, int nElemX, int nElemY, double Lx, double Ly) {
  // matrix_free_gpu/matrixfreegpu.cu:104-137
  FemMesh m;
  const double hx = Lx / (nElemX + 2), hy = Ly / (nElemY + 2);
  m.nfx = p + nElemX;
  m.nfy = q + nElemY;
  std::vector<double> S, T;
  for (int i = 0; i < p - 1; ++i) S.push_back(0.0);
  for (int i = 1; i < nElemX + 2; ++i) S.push_back(i * hx);
  for (int i = 0; i < p - 1; ++i) S.push_back(Lx);
  for (int i = 0; i < q - 1; ++i) T.push_back(0.0);
  for (int i = 1; i < nElemY + 2; ++i) T.push_back(i * hy);
  for (int i = 0; i < q - 1; ++i) T.push_back(Ly);
  AbscissaeGenerator ag;
  m.CP = ag.GenerateAbscissae2D(S, T, p - 2, q - 2);
  m.nex = m.nfx - 1;
  m.ney = m.nfy - 1;
  IENGenerator ig;
  m.IEN = ig.GenerateIEN2D(m.nex, m.ney);
  IDGenerator idg;
  m.ID = idg.GenerateID2D(m.nfx, m.nfy);
  for (size_t i = 0; i < m.ID.size(); ++i)
    if (m.ID[i] != -1) m.ID[i] = (int)i;
  return m;
}
}  // namespace

extern "C" {

int ref_fem_nodes(int p, int q, int nex, int ney, double Lx, double Ly,
                  double *CPf) {
  FemMesh m = fem_mesh(p, q, nex, ney, Lx, Ly);
  std::memcpy(CPf, m.CP.data(), sizeof(double) * m.CP.size());
  return (int)(m.CP.size() / 2);
}

// Sorted unique COO pattern {(ID_a,ID_b) : both >= 0} U {(d,d)}, then
// Elem2COOGenerator::GenerateElem2COO / GenerateDir2COO and the value loop.
// Returns nnz; rows/cols/val need nnz entries, elem2coo 16*E, dir2coo ndir.
long ref_fem_assemble(int p, int q, int nex, int ney, double Lx, double Ly,
                      long cap, int *rows, int *cols, int *elem2coo,
                      int *dir2coo, double *val, int *ndir_out) {
  FemMesh m = fem_mesh(p, q, nex, ney, Lx, Ly);
  std::vector<std::pair<int, int>> pr;
  const int E = m.nex * m.ney;
  for (int e = 0; e < E; ++e)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        int I = m.ID[m.IEN[e * 4 + a]], J = m.ID[m.IEN[e * 4 + b]];
        if (I >= 0 && J >= 0) pr.emplace_back(I, J);
      }
  std::vector<int> Dir;
  for (size_t i = 0; i < m.ID.size(); ++i)
    if (m.ID[i] == -1) { pr.emplace_back((int)i, (int)i); Dir.push_back((int)i); }
  std::sort(pr.begin(), pr.end());
  pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
  const long nnz = (long)pr.size();
  if (nnz > cap) return -nnz;
  std::vector<int> r(nnz), c(nnz);
  for (long k = 0; k < nnz; ++k) { r[k] = pr[k].first; c[k] = pr[k].second; }
  Elem2COOGenerator gen(4, (int)nnz, m.nfx * m.nfy, m.nex, m.ney);
  std::vector<int> e2c, d2c;
  gen.GenerateElem2COO(m.IEN, m.ID, r, c, e2c);
  gen.GenerateDir2COO(m.ID, r, d2c);
  std::memcpy(rows, r.data(), sizeof(int) * nnz);
  std::memcpy(cols, c.data(), sizeof(int) * nnz);
  std::memcpy(elem2coo, e2c.data(), sizeof(int) * (size_t)E * 16);
  std::memcpy(dir2coo, d2c.data(), sizeof(int) * d2c.size());
  *ndir_out = (int)d2c.size();
  // values: AssembleStiffnessKernel's loop, sequential
  QuadraturePoint q1(2, 0, 1), q2(2, 0, 1);
  ElementFEM ef(1, 1);
  std::vector<double> g1 = q1.GetQuadraturePoint(), g2 = q2.GetQuadraturePoint();
  std::vector<double> w1 = q1.GetWeight(), w2 = q2.GetWeight();
  std::vector<double> N, dxi, deta, tabN, tabxi, tabeta, wq;
  for (int jj = 0; jj < 2; ++jj)
    for (int ii = 0; ii < 2; ++ii) {
      ef.GenerateRefElementSingleQP(g1[ii], g2[jj], N, dxi, deta);
      tabxi.insert(tabxi.end(), dxi.begin(), dxi.end());
      tabeta.insert(tabeta.end(), deta.begin(), deta.end());
      wq.push_back(w1[ii] * w2[jj]);
    }
  for (long k = 0; k < nnz; ++k) val[k] = 0.0;
  for (int e = 0; e < E; ++e) {
    double eCP[8];
    for (int j = 0; j < 4; ++j) {
      eCP[2 * j] = m.CP[2 * m.IEN[e * 4 + j]];
      eCP[2 * j + 1] = m.CP[2 * m.IEN[e * 4 + j] + 1];
    }
    for (int k = 0; k < 4; ++k) {
      const double *a = &tabxi[k * 4], *b = &tabeta[k * 4];
      double xxi = 0, xeta = 0, yxi = 0, yeta = 0;
      for (int i = 0; i < 4; ++i) {
        xxi += eCP[2 * i] * a[i];
        xeta += eCP[2 * i] * b[i];
        yxi += eCP[2 * i + 1] * a[i];
        yeta += eCP[2 * i + 1] * b[i];
      }
      const double J = xxi * yeta - xeta * yxi;
      const double i11 = yeta / J, i12 = -yxi / J, i21 = -xeta / J, i22 = xxi / J;
      double dx[4], dy[4];
      for (int i = 0; i < 4; ++i) {
        dx[i] = i11 * a[i] + i12 * b[i];
        dy[i] = i21 * a[i] + i22 * b[i];
      }
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          const int pos = e2c[e * 16 + i * 4 + j];
          if (pos >= 0) val[pos] += -(dx[i] * dx[j] + dy[i] * dy[j]) * J * wq[k];
        }
    }
  }
  for (size_t i = 0; i < d2c.size(); ++i) val[d2c[i]] = 1.0;
  return nnz;
}

long ref_invlm(int nloc, int nelem, int nfunc, const int *IEN, int *elemNum,
               int *offset, int *elemIdx, int *baseIdx) {
  InvLM inv(nloc, nelem, nfunc,
            std::vector<int>(IEN, IEN + (long)nelem * nloc));
  std::vector<int> a = inv.GetAllElemNum(), b = inv.GetAllOffset();
  std::vector<int> c = inv.GetAllElemIdx(), d = inv.GetAllBaseIdx();
  std::memcpy(elemNum, a.data(), sizeof(int) * a.size());
  std::memcpy(offset, b.data(), sizeof(int) * b.size());
  std::memcpy(elemIdx, c.data(), sizeof(int) * c.size());
  std::memcpy(baseIdx, d.data(), sizeof(int) * d.size());
  return (long)c.size();
}

// The FEM preprocessing of matrix_free_gpu/preprocess_fem.cpp:15-70 on
// [0,1]^2 (Greville nodes of degree p-2, Q1 IEN, ID with -1 on the boundary),
// partitioned by the reference's Partition::GeneratePartition into
// part_fem_<r>.txt and new_to_old_mapping.txt in directory dir (both written
// relative to the working directory, hence the chdir).  Returns the node
// count per direction in nf[0..1].
int ref_fem_write_partitions(const char *dir, int p, int q, int nElemX, int nElemY,
                             int part_num_1d, int *nf) {
  char old[4096];
  if (!getcwd(old, sizeof(old)) || chdir(dir) != 0) return -1;
  const double Lx = 1.0, Ly = 1.0;
  const double hx = Lx / (nElemX + 2), hy = Ly / (nElemY + 2);
  const int nFuncX = p + nElemX, nFuncY = q + nElemY;
  std::vector<double> S, T;
  for (int i = 0; i < p - 1; ++i) S.push_back(0.0);
  for (int i = 1; i < nElemX + 2; ++i) S.push_back(i * hx);
  for (int i = 0; i < p - 1; ++i) S.push_back(Lx);
  for (int i = 0; i < q - 1; ++i) T.push_back(0.0);
  for (int i = 1; i < nElemY + 2; ++i) T.push_back(i * hy);
  for (int i = 0; i < q - 1; ++i) T.push_back(Ly);
  AbscissaeGenerator ag;
  std::vector<double> CP = ag.GenerateAbscissae2D(S, T, p - 2, q - 2);
  IENGenerator ig;
  std::vector<int> IEN = ig.GenerateIEN2D(nFuncX - 1, nFuncY - 1);
  IDGenerator idg;
  std::vector<int> ID = idg.GenerateID2D(nFuncX, nFuncY);
  Partition part(part_num_1d, part_num_1d, 2, "part_fem");
  part.GeneratePartition(nFuncX - 1, nFuncY - 1, CP, IEN, ID);
  nf[0] = nFuncX;
  nf[1] = nFuncY;
  return chdir(old) == 0 ? 0 : -1;
}

// FileManager::ReadPartition (short form).  counts: nlocalfunc, nlocalelemx,
// nlocalelemy, |ID|, |Dir|, |IEN|, |CP|; arrays copied when they fit cap.
int ref_fem_read_partition(const char *fn, long cap, int *counts, int *id, int *dir,
                           int *ien, double *cp) {
  FileManager fm;
  int nf = 0, nx = 0, ny = 0;
  std::vector<double> CP;
  std::vector<int> ID, Dir, IEN;
  fm.ReadPartition(fn, nf, nx, ny, CP, ID, Dir, IEN);
  const long sz[7] = {nf, nx, ny, (long)ID.size(), (long)Dir.size(), (long)IEN.size(),
                      (long)CP.size()};
  for (int k = 0; k < 7; ++k) counts[k] = (int)sz[k];
  if ((long)ID.size() > cap || (long)Dir.size() > cap || (long)IEN.size() > cap ||
      (long)CP.size() > cap)
    return -1;
  std::copy(ID.begin(), ID.end(), id);
  std::copy(Dir.begin(), Dir.end(), dir);
  std::copy(IEN.begin(), IEN.end(), ien);
  std::copy(CP.begin(), CP.end(), cp);
  return 0;
}

void ref_write_nonzero(const char *fn, int nnz, const int *rows, const int *cols) {
  FileManager fm;
  fm.WriteNonZeroCoordinate(fn, nnz, std::vector<int>(rows, rows + nnz),
                            std::vector<int>(cols, cols + nnz));
}

// FileManager::ReadNonZeroCoordinate; returns nnz (negative when over cap)
long ref_read_nonzero(const char *fn, long cap, int *rows, int *cols) {
  FileManager fm;
  int nnz = 0;
  std::vector<int> r, c;
  fm.ReadNonZeroCoordinate(fn, nnz, r, c);
  if ((long)r.size() > cap || r.size() != c.size()) return -(long)r.size() - 1;
  std::copy(r.begin(), r.end(), rows);
  std::copy(c.begin(), c.end(), cols);
  return nnz;
}

void ref_write_new_to_old(const char *fn, int n, const int *map) {
  FileManager fm;
  fm.WriteNewToOldMapping(fn, std::vector<int>(map, map + n));
}

long ref_read_new_to_old(const char *fn, long cap, int *map) {
  FileManager fm;
  std::vector<int> v;
  fm.ReadNewToOldMapping(fn, v);
  if ((long)v.size() > cap) return -(long)v.size() - 1;
  std::copy(v.begin(), v.end(), map);
  return (long)v.size();
}

}  // extern "C"
