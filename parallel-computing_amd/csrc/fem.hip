// fem.hip -- the low-order (Q1) FEM preconditioner matrix, assembled on the
// device by a deterministic sort-reduce.
//
// Replaces the reference's elem2coo -> coo2aij chain: dry-run sparsity
// pattern (matrix_free_gpu/src/GlobalAssemblyDR.cpp:33-94), host
// re-sort and Elem2COOGenerator (matrix_free_gpu/matrixfreegpu.cu:149-196,
// matrix_free_gpu/src/Elem2COOGenerator.cpp:3-69), atomic value assembly
// (matrix_free_gpu/src/assembly.cu:41-99) and MatSetPreallocationCOO /
// MatSetValuesCOO (matrix_free_gpu/src/GlobalAssembly.cu:3-108).
//
// Every element contribution (e, a, b) gets the key row*N + col (a sentinel
// when either node is Dirichlet); Dirichlet diagonals are appended.  A stable
// radix sort of the keys gives the reference's lexicographically sorted,
// unique pattern (segment heads), elem2coo / dir2coo (segment id of each
// contribution), and values summed in contribution order inside each segment:
// indices are bit-exact and values do not depend on thread timing.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cmath>
#include <new>
#include <vector>

#include "iga_mf.h"

struct iga_fem {
  int device = 0;
  int nfx = 0, nfy = 0, nrows = 0, ndir = 0;
  int64_t nnz = 0;
  int *rowptr = nullptr, *col = nullptr, *coo_row = nullptr;
  int *elem2coo = nullptr, *dir2coo = nullptr;
  double *val = nullptr;
  std::vector<void *> allocs;
};

namespace {

__device__ __forceinline__ bool ring(int i, int j, int nfx, int nfy) {
  return i == 0 || j == 0 || i == nfx - 1 || j == nfy - 1;
}

// contributions of Q1 element e: -(grad N_a . grad N_b) det w over 2x2 Gauss
// points (the element of matrix_free_gpu/src/ElementFEM.cpp with the
// Jacobian of assembly.cu:3-39)
__global__ void k_fem_contrib(const double *nodes, int nfx, int nfy, int64_t N,
                              uint64_t sentinel, uint64_t *keys, int *idx, double *cval) {
  const int nex = nfx - 1, ney = nfy - 1;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nex * ney) return;
  const int ex = e % nex, ey = e / nex;
  const int nd[4] = {ey * nfx + ex, ey * nfx + ex + 1, (ey + 1) * nfx + ex, (ey + 1) * nfx + ex + 1};
  const int ni[4] = {ex, ex + 1, ex, ex + 1}, nj[4] = {ey, ey, ey + 1, ey + 1};
  double X[4], Y[4];
  for (int a = 0; a < 4; ++a) {
    X[a] = nodes[2 * (size_t)nd[a]];
    Y[a] = nodes[2 * (size_t)nd[a] + 1];
  }
  const double g0 = 0.5 - 0.5 / sqrt(3.0), g1 = 0.5 + 0.5 / sqrt(3.0);
  const double gp[2] = {g0, g1};
  double K[4][4] = {};
  for (int jj = 0; jj < 2; ++jj)
    for (int ii = 0; ii < 2; ++ii) {
      const double xi = gp[ii], eta = gp[jj];
      const double dxi[4] = {-(1 - eta), (1 - eta), -eta, eta};
      const double deta[4] = {-(1 - xi), -xi, (1 - xi), xi};
      double xxi = 0, xeta = 0, yxi = 0, yeta = 0;
      for (int a = 0; a < 4; ++a) {
        xxi += X[a] * dxi[a];
        xeta += X[a] * deta[a];
        yxi += Y[a] * dxi[a];
        yeta += Y[a] * deta[a];
      }
      const double det = xxi * yeta - xeta * yxi;
      double dx[4], dy[4];
      for (int a = 0; a < 4; ++a) {
        dx[a] = (yeta * dxi[a] - yxi * deta[a]) / det;
        dy[a] = (-xeta * dxi[a] + xxi * deta[a]) / det;
      }
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) K[a][b] += -(dx[a] * dx[b] + dy[a] * dy[b]) * det * 0.25;
    }
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      const size_t k = (size_t)e * 16 + a * 4 + b;
      const bool keep = !ring(ni[a], nj[a], nfx, nfy) && !ring(ni[b], nj[b], nfx, nfy);
      keys[k] = keep ? (uint64_t)nd[a] * (uint64_t)N + (uint64_t)nd[b] : sentinel;
      idx[k] = (int)k;
      cval[k] = K[a][b];
    }
}

// Dirichlet diagonals, ascending node order (the order of the reference's
// Dir list / GenerateDir2COO loop)
__global__ void k_fem_dir(const int *dir, int ndir, int64_t N, int64_t base, uint64_t *keys,
                          int *idx, double *cval) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndir) return;
  const int64_t k = base + i;
  keys[k] = (uint64_t)dir[i] * (uint64_t)N + (uint64_t)dir[i];
  idx[k] = (int)k;
  cval[k] = 1.0;
}

__global__ void k_heads(const uint64_t *keys, int n, uint64_t sentinel, int *head) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  head[k] = (keys[k] != sentinel && (k == 0 || keys[k] != keys[k - 1])) ? 1 : 0;
}

// seg = inclusive_scan(head): segment id + 1
__global__ void k_scatter(const uint64_t *keys, const int *perm, const int *seg, const int *head,
                          int n, uint64_t sentinel, int64_t N, int64_t nelemc, int *rows,
                          int *cols, int *starts, int *elem2coo, int *dir2coo) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int orig = perm[k];
  if (keys[k] == sentinel) {
    if (orig < nelemc) elem2coo[orig] = -1;
    return;
  }
  const int s = seg[k] - 1;  // inclusive scan of the heads is 1-based
  if (head[k]) {
    rows[s] = (int)(keys[k] / (uint64_t)N);
    cols[s] = (int)(keys[k] % (uint64_t)N);
    starts[s] = k;
  }
  if (orig < nelemc)
    elem2coo[orig] = s;
  else
    dir2coo[orig - nelemc] = s;
}

__global__ void k_segsum(const int *starts, int nnz, int nvalid, const int *perm,
                         const double *cval, const int *rows, int nrows, double *val,
                         int *rowptr) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  const int b = starts[s], e = (s + 1 < nnz) ? starts[s + 1] : nvalid;
  double acc = 0.0;
  for (int j = b; j < e; ++j) acc += cval[perm[j]];
  val[s] = acc;
  if (s == 0 || rows[s] != rows[s - 1]) rowptr[rows[s]] = s;
  if (s == nnz - 1) rowptr[nrows] = nnz;
}

}  // namespace

extern "C" {

int iga_fem_reference_nodes(int p, int q, int nex, int ney, double Lx, double Ly,
                            double *nodes) {
  // matrix_free_gpu/matrixfreegpu.cu:104-122: knots with p-1 repeated ends and
  // interior knots i*L/(nelem+2); nodes at the degree-(p-2) Greville points.
  // For degree 2 the reference's degree-0 average divides by zero; the
  // uniform grid (what degree 3 yields) is used instead.
  if (!nodes) return IGA_ERR_INVALID_ARGUMENT;
  if (p < 2 || q < 2 || p > 8 || q > 8) return IGA_ERR_INVALID_DEGREE;
  if (nex < p + 2 || ney < q + 2) return IGA_ERR_TOO_FEW_ELEMENTS;
  const int nfx = p + nex, nfy = q + ney;
  std::vector<double> g[2];
  for (int d = 0; d < 2; ++d) {
    const int pd = d ? q : p, ne = d ? ney : nex, nf = d ? nfy : nfx;
    const double L = d ? Ly : Lx;
    g[d].resize(nf);
    if (pd == 2) {
      for (int i = 0; i < nf; ++i) g[d][i] = (i == nf - 1) ? L : i * (L / (nf - 1));
      continue;
    }
    const double h = L / (ne + 2);
    std::vector<double> S;
    for (int i = 0; i < pd - 1; ++i) S.push_back(0.0);
    for (int i = 1; i < ne + 2; ++i) S.push_back(i * h);
    for (int i = 0; i < pd - 1; ++i) S.push_back(L);
    const int deg = pd - 2;
    for (int i = 0; i < nf; ++i) {
      double s = 0.0;
      for (int j = 1; j <= deg; ++j) s += S[i + j];
      g[d][i] = s / deg;
    }
  }
  for (int j = 0; j < nfy; ++j)
    for (int i = 0; i < nfx; ++i) {
      nodes[2 * ((size_t)j * nfx + i)] = g[0][i];
      nodes[2 * ((size_t)j * nfx + i) + 1] = g[1][j];
    }
  return IGA_OK;
}

int iga_fem_destroy(iga_fem *f) {
  if (!f) return IGA_OK;
  int cur = -1;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(f->device);
  for (void *p : f->allocs) (void)hipFree(p);
  if (cur >= 0) (void)hipSetDevice(cur);
  delete f;
  return IGA_OK;
}

int iga_fem_assemble(int nfx, int nfy, const double *nodes_h, int device, iga_fem **out) {
  if (!nodes_h || !out || nfx < 3 || nfy < 3) return IGA_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  const int64_t N = (int64_t)nfx * nfy;
  const int64_t ne = (int64_t)(nfx - 1) * (nfy - 1);
  const int64_t nc = ne * 16;
  std::vector<int> dir;
  for (int j = 0; j < nfy; ++j)
    for (int i = 0; i < nfx; ++i)
      if (i == 0 || j == 0 || i == nfx - 1 || j == nfy - 1) dir.push_back(j * nfx + i);
  const int64_t n = nc + (int64_t)dir.size();
  if (n >= (int64_t)1 << 31) return IGA_ERR_SIZE_LIMIT;
  for (int64_t k = 0; k < 2 * N; ++k)
    if (!std::isfinite(nodes_h[k])) return IGA_ERR_BAD_MESH;
  iga_fem *f = new (std::nothrow) iga_fem;
  if (!f) return IGA_ERR_OUT_OF_MEMORY;
  f->device = device;
  f->nfx = nfx;
  f->nfy = nfy;
  f->nrows = (int)N;
  f->ndir = (int)dir.size();
  std::vector<void *> tmp;
  auto cleanup = [&](int rc) {
    for (void *p : tmp) (void)hipFree(p);
    if (rc) iga_fem_destroy(f);
    return rc;
  };
  if (hipSetDevice(device) != hipSuccess) return cleanup(IGA_ERR_HIP);
  auto keep_alloc = [&](void **p, size_t bytes) {
    if (hipMalloc(p, bytes ? bytes : 256) != hipSuccess) { *p = nullptr; return false; }
    f->allocs.push_back(*p);
    return true;
  };
  auto tmp_alloc = [&](void **p, size_t bytes) {
    if (hipMalloc(p, bytes ? bytes : 256) != hipSuccess) { *p = nullptr; return false; }
    tmp.push_back(*p);
    return true;
  };
  double *nodes = nullptr, *cval = nullptr;
  uint64_t *keys = nullptr, *keys_s = nullptr;
  int *idx = nullptr, *perm = nullptr, *head = nullptr, *seg = nullptr, *ddir = nullptr;
  int *starts = nullptr;
  bool ok = tmp_alloc((void **)&nodes, 2 * N * 8) && tmp_alloc((void **)&cval, n * 8) &&
            tmp_alloc((void **)&keys, n * 8) && tmp_alloc((void **)&keys_s, n * 8) &&
            tmp_alloc((void **)&idx, n * 4) && tmp_alloc((void **)&perm, n * 4) &&
            tmp_alloc((void **)&head, n * 4) && tmp_alloc((void **)&seg, n * 4) &&
            tmp_alloc((void **)&ddir, dir.size() * 4) && tmp_alloc((void **)&starts, n * 4) &&
            keep_alloc((void **)&f->elem2coo, nc * 4) &&
            keep_alloc((void **)&f->dir2coo, dir.size() * 4) &&
            keep_alloc((void **)&f->rowptr, (N + 1) * 4);
  if (!ok) return cleanup(IGA_ERR_OUT_OF_MEMORY);
  hipStream_t s = nullptr;
  if (hipMemcpy(nodes, nodes_h, 2 * N * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(ddir, dir.data(), dir.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(IGA_ERR_HIP);
  const uint64_t sentinel = (uint64_t)N * (uint64_t)N;
  int end_bit = 1;
  while (end_bit < 64 && (sentinel >> end_bit) != 0) ++end_bit;
  const int bs = 256;
  hipLaunchKernelGGL(k_fem_contrib, dim3((unsigned)((ne + bs - 1) / bs)), dim3(bs), 0, s, nodes,
                     nfx, nfy, N, sentinel, keys, idx, cval);
  if (!dir.empty())
    hipLaunchKernelGGL(k_fem_dir, dim3((unsigned)((dir.size() + bs - 1) / bs)), dim3(bs), 0, s,
                       ddir, (int)dir.size(), N, nc, keys, idx, cval);
  if (hipGetLastError() != hipSuccess) return cleanup(IGA_ERR_HIP);
  size_t tb = 0, tb2 = 0;
  void *temp = nullptr;
  // rocPRIM's LSD radix sort is stable: equal keys keep their contribution
  // order, so the sums below run in a fixed order (deterministic values)
  if (rocprim::radix_sort_pairs(nullptr, tb, keys, keys_s, idx, perm, (size_t)n, 0u,
                                (unsigned)end_bit, s) != hipSuccess ||
      rocprim::inclusive_scan(nullptr, tb2, head, seg, (size_t)n, rocprim::plus<int>(), s) !=
          hipSuccess)
    return cleanup(IGA_ERR_HIP);
  if (!tmp_alloc(&temp, tb > tb2 ? tb : tb2)) return cleanup(IGA_ERR_OUT_OF_MEMORY);
  if (rocprim::radix_sort_pairs(temp, tb, keys, keys_s, idx, perm, (size_t)n, 0u,
                                (unsigned)end_bit, s) != hipSuccess)
    return cleanup(IGA_ERR_HIP);
  hipLaunchKernelGGL(k_heads, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, keys_s, (int)n,
                     sentinel, head);
  if (rocprim::inclusive_scan(temp, tb2, head, seg, (size_t)n, rocprim::plus<int>(), s) !=
      hipSuccess)
    return cleanup(IGA_ERR_HIP);
  // nnz = last inclusive-scan value; valid entries precede the sentinels
  int last = 0;
  if (hipMemcpy(&last, seg + (n - 1), 4, hipMemcpyDeviceToHost) != hipSuccess)
    return cleanup(IGA_ERR_HIP);
  const int nnz = last;
  f->nnz = nnz;
  // number of non-sentinel sorted entries
  // sorted entries before the sentinels (contributions with a Dirichlet node)
  int64_t nvalid = 0;
  {
    int64_t nsent = 0;
    for (int j = 0; j < nfy - 1; ++j)
      for (int i = 0; i < nfx - 1; ++i) {
        const int ni[4] = {i, i + 1, i, i + 1}, nj[4] = {j, j, j + 1, j + 1};
        int inter = 0;
        for (int a = 0; a < 4; ++a)
          inter += !(ni[a] == 0 || nj[a] == 0 || ni[a] == nfx - 1 || nj[a] == nfy - 1);
        nsent += 16 - inter * inter;
      }
    nvalid = n - nsent;
  }
  ok = keep_alloc((void **)&f->col, (size_t)nnz * 4) &&
       keep_alloc((void **)&f->val, (size_t)nnz * 8) &&
       keep_alloc((void **)&f->coo_row, (size_t)nnz * 4);
  if (!ok) return cleanup(IGA_ERR_OUT_OF_MEMORY);
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, keys_s, perm,
                     seg, head, (int)n, sentinel, N, nc, f->coo_row, f->col, starts, f->elem2coo,
                     f->dir2coo);
  hipLaunchKernelGGL(k_segsum, dim3((unsigned)((nnz + bs - 1) / bs)), dim3(bs), 0, s, starts, nnz,
                     (int)nvalid, perm, cval, f->coo_row, (int)N, f->val, f->rowptr);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return cleanup(IGA_ERR_HIP);
  cleanup(IGA_OK);
  *out = f;
  return IGA_OK;
}

int iga_fem_view(const iga_fem *f, iga_csr_view *v) {
  if (!f || !v) return IGA_ERR_INVALID_ARGUMENT;
  v->nrows = f->nrows;
  v->nnz = f->nnz;
  v->rowptr = f->rowptr;
  v->col = f->col;
  v->val = f->val;
  v->coo_row = f->coo_row;
  v->elem2coo = f->elem2coo;
  v->dir2coo = f->dir2coo;
  v->ndir = f->ndir;
  return IGA_OK;
}

}  // extern "C"
