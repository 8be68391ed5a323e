// Device matvec plan and CDNA4 kernels for the flattened k-tree traversal.
//
// The reference computes y = A x by a recursive top-down/bottom-up traversal of
// the node list (fem::matvec_rec, FEM/include/matvec.h:243-548).  Here the
// traversal is flattened once per mesh (mesh.cpp) into per-leaf gather maps,
// and each matvec is two launches on the caller's stream:
//
//   leaf_kernel   one work-item per traversal leaf: gather the leaf's nodal
//                 values (hanging ranks through the hanging-gather lists),
//                 apply the elemental matrix, write the nPe outputs, and for
//                 hanging leaves the transposed interpolation of the missing
//                 ranks (matvec.h:482-522);
//   node_kernel   one work-item per node: sum the node's contributions in a
//                 fixed order (leaf SFC order, rank order) -- a deterministic
//                 segmented reduction, no atomics.
//
// Both are HBM-bound at p=1.  Layout of multi-RHS vectors: node-major,
// [node][nrhs] (the reference's dof-interleaved layout, include/oda.h:296-322).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dkt.h"
#include "capi_common.h"
#include "mesh.h"

namespace dkt {
const MeshParams &da_params(const dkt_da *da);
const NodeSet &da_nodes(const dkt_da *da);
const Traversal &da_traversal(const dkt_da *da);
}  // namespace dkt

using namespace dkt;

#define HIP_OK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      throw std::runtime_error(std::string(#expr " failed: ") + hipGetErrorString(e_));       \
  } while (0)

namespace {

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    n = count;
    if (count) HIP_OK(hipMalloc((void **)&p, count * sizeof(T)));
  }
  void upload(const std::vector<T> &v) {
    alloc(v.size());
    if (n) HIP_OK(hipMemcpy(p, v.data(), n * sizeof(T), hipMemcpyHostToDevice));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------
template <typename T, int NPE>
__global__ void __launch_bounds__(256) leaf_kernel(
    int L, int nrhs, const T *__restrict__ x, const int32_t *__restrict__ slot, const uint8_t *__restrict__ level,
    const int32_t *__restrict__ hang_of, const uint8_t *__restrict__ child, const int32_t *__restrict__ hg_off,
    const int32_t *__restrict__ hg_node, const T *__restrict__ hg_w, const T *__restrict__ kmat,
    int kmat_per_level, const T *__restrict__ wint, const uint8_t *__restrict__ dirichlet, int hp,
    T *__restrict__ vout, T *__restrict__ wout) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L) return;
  int32_t s[NPE];
#pragma unroll
  for (int r = 0; r < NPE; r++) s[r] = slot[(size_t)e * NPE + r];
  const int lev = level[e];
  const int h = hang_of[e];
  const T scale = kmat_per_level ? T(1) : (T)ldexp(1.0, -lev * hp);
  const T *K = kmat + (kmat_per_level ? (size_t)lev * NPE * NPE : 0);
  for (int q = 0; q < nrhs; q++) {
    T u[NPE];
#pragma unroll
    for (int r = 0; r < NPE; r++) {
      T val = T(0);
      const int n = s[r];
      if (n >= 0) {
        if (!(dirichlet && dirichlet[n])) val = x[(size_t)n * nrhs + q];
      } else if (h >= 0) {
        const int a = hg_off[(size_t)h * NPE + r], b = hg_off[(size_t)h * NPE + r + 1];
        for (int k = a; k < b; k++) {
          const int m = hg_node[k];
          if (!(dirichlet && dirichlet[m])) val += hg_w[k] * x[(size_t)m * nrhs + q];
        }
      }
      u[r] = val;
    }
    T v[NPE];
#pragma unroll
    for (int i = 0; i < NPE; i++) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < NPE; j++) acc += K[i * NPE + j] * u[j];
      v[i] = scale * acc;
    }
#pragma unroll
    for (int r = 0; r < NPE; r++) vout[((size_t)e * NPE + r) * nrhs + q] = v[r];
    if (h >= 0) {
      const T *W = wint + (size_t)child[e] * NPE * NPE;  // W[r][s]
#pragma unroll
      for (int t = 0; t < NPE; t++) {
        T acc = T(0);
#pragma unroll
        for (int r = 0; r < NPE; r++)
          if (s[r] < 0) acc += W[r * NPE + t] * v[r];
        wout[((size_t)h * NPE + t) * nrhs + q] = acc;
      }
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) node_kernel(int N, int nrhs, const int64_t *__restrict__ off,
                                                   const int64_t *__restrict__ src, const T *__restrict__ buf,
                                                   const uint8_t *__restrict__ dirichlet, T *__restrict__ y) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int64_t a = off[n], b = off[n + 1];
  const bool zero = dirichlet && dirichlet[n];
  for (int q = 0; q < nrhs; q++) {
    T acc = T(0);
    for (int64_t k = a; k < b; k++) acc += buf[src[k] * nrhs + q];
    y[(size_t)n * nrhs + q] = zero ? T(0) : acc;
  }
}

// ---------------------------------------------------------------------------
// Host-side operator matrices.
// ---------------------------------------------------------------------------
void one_d(int p, std::vector<double> &K, std::vector<double> &M) {
  if (p == 1) {
    K = {1, -1, -1, 1};
    M = {2.0 / 6, 1.0 / 6, 1.0 / 6, 2.0 / 6};
  } else {
    K = {7.0 / 3, -8.0 / 3, 1.0 / 3, -8.0 / 3, 16.0 / 3, -8.0 / 3, 1.0 / 3, -8.0 / 3, 7.0 / 3};
    M = {4.0 / 30, 2.0 / 30, -1.0 / 30, 2.0 / 30, 16.0 / 30, 2.0 / 30, -1.0 / 30, 2.0 / 30, 4.0 / 30};
  }
}

// Reference-element matrix (unit cube, axis 0 fastest): stiffness = sum_t of
// tensor products with the 1D stiffness on axis t and 1D mass elsewhere.
std::vector<double> ref_matrix(int dim, int p, int op) {
  const int n1 = p + 1;
  int npe = 1;
  for (int d = 0; d < dim; d++) npe *= n1;
  std::vector<double> K1, M1;
  one_d(p, K1, M1);
  std::vector<double> A((size_t)npe * npe, 0.0);
  for (int i = 0; i < npe; i++)
    for (int j = 0; j < npe; j++) {
      if (op == DKT_OP_IDENTITY) {
        A[i * npe + j] = i == j ? 1.0 : 0.0;
        continue;
      }
      const int terms = op == DKT_OP_STIFFNESS ? dim : 1;
      double s = 0.0;
      for (int t = 0; t < terms; t++) {
        double prod = 1.0;
        int ii = i, jj = j;
        for (int d = 0; d < dim; d++) {
          const double *m = (op == DKT_OP_STIFFNESS && d == t) ? K1.data() : M1.data();
          prod *= m[(ii % n1) * n1 + (jj % n1)];
          ii /= n1;
          jj /= n1;
        }
        s += prod;
      }
      A[i * npe + j] = s;
    }
  return A;
}

// W[c][r][s]: weight of parent rank s in child rank r for child Morton index c.
std::vector<double> interp_matrices(int dim, int p) {
  const int n1 = p + 1, nch = 1 << dim;
  int npe = 1;
  for (int d = 0; d < dim; d++) npe *= n1;
  double ip[2][9];
  interp_1d(p, 0, ip[0]);
  interp_1d(p, 1, ip[1]);
  std::vector<double> W((size_t)nch * npe * npe);
  for (int c = 0; c < nch; c++)
    for (int r = 0; r < npe; r++)
      for (int s = 0; s < npe; s++) {
        double w = 1.0;
        int rr = r, ss = s;
        for (int d = 0; d < dim; d++) {
          w *= ip[(c >> d) & 1][(ss % n1) * n1 + (rr % n1)];
          rr /= n1;
          ss /= n1;
        }
        W[((size_t)c * npe + r) * npe + s] = w;
      }
  return W;
}

template <typename T>
std::vector<T> cast(const std::vector<double> &v) {
  return std::vector<T>(v.begin(), v.end());
}

}  // namespace

struct dkt_plan {
  int dim = 0, order = 0, npe = 0, op = 0, hp = 0;
  size_t N = 0, L = 0, H = 0, nbuf = 0;
  bool per_level = false;
  DevBuf<int32_t> slot, hang_of, hg_off, hg_node;
  DevBuf<uint8_t> level, child, dirichlet;
  DevBuf<double> hg_w64, kmat64, wint64;
  DevBuf<float> hg_w32, kmat32, wint32;
  DevBuf<int64_t> csr_off, csr_src;
  DevBuf<double> scratch64;
  DevBuf<float> scratch32;
  size_t host_bytes_maps = 0;
};

namespace {

template <typename T, int NPE>
void launch_leaf(dkt_plan *pl, int nrhs, const T *x, const T *kmat, const T *hgw, const T *wint, T *buf,
                 bool masked, hipStream_t st) {
  const int blk = 256;
  const int grid = (int)((pl->L + blk - 1) / blk);
  if (!pl->L) return;
  T *vout = buf;
  T *wout = buf + pl->L * (size_t)NPE * nrhs;
  hipLaunchKernelGGL((leaf_kernel<T, NPE>), dim3(grid), dim3(blk), 0, st, (int)pl->L, nrhs, x, pl->slot.p,
                     pl->level.p, pl->hang_of.p, pl->child.p, pl->hg_off.p, pl->hg_node.p, hgw, kmat,
                     pl->per_level ? 1 : 0, wint, masked ? pl->dirichlet.p : nullptr, pl->hp, vout, wout);
  HIP_OK(hipGetLastError());
}

template <typename T>
void run_matvec(dkt_plan *pl, int nrhs, const T *x, T *y, bool masked, hipStream_t st) {
  const bool f64 = sizeof(T) == 8;
  const T *kmat = f64 ? (const T *)pl->kmat64.p : (const T *)pl->kmat32.p;
  const T *hgw = f64 ? (const T *)pl->hg_w64.p : (const T *)pl->hg_w32.p;
  const T *wint = f64 ? (const T *)pl->wint64.p : (const T *)pl->wint32.p;
  const size_t need = pl->nbuf * (size_t)nrhs;
  T *buf;
  if (f64) {
    if (pl->scratch64.n < need) pl->scratch64.alloc(need);
    buf = (T *)pl->scratch64.p;
  } else {
    if (pl->scratch32.n < need) pl->scratch32.alloc(need);
    buf = (T *)pl->scratch32.p;
  }
  switch (pl->npe) {
    case 4: launch_leaf<T, 4>(pl, nrhs, x, kmat, hgw, wint, buf, masked, st); break;
    case 8: launch_leaf<T, 8>(pl, nrhs, x, kmat, hgw, wint, buf, masked, st); break;
    case 9: launch_leaf<T, 9>(pl, nrhs, x, kmat, hgw, wint, buf, masked, st); break;
    case 16: launch_leaf<T, 16>(pl, nrhs, x, kmat, hgw, wint, buf, masked, st); break;
    case 27: launch_leaf<T, 27>(pl, nrhs, x, kmat, hgw, wint, buf, masked, st); break;
    default: throw std::invalid_argument("unsupported nodes per element");
  }
  const int blk = 256;
  const int grid = (int)((pl->N + blk - 1) / blk);
  hipLaunchKernelGGL((node_kernel<T>), dim3(grid), dim3(blk), 0, st, (int)pl->N, nrhs, pl->csr_off.p,
                     pl->csr_src.p, buf, masked ? pl->dirichlet.p : nullptr, y);
  HIP_OK(hipGetLastError());
}

}  // namespace

extern "C" int dkt_plan_create(const dkt_da *da, int op, const double *op_data, int flags, dkt_plan **out) {
  return guarded([&] {
    (void)flags;
    if (!da || !out) throw std::invalid_argument("null pointer");
    const MeshParams &mp = da_params(da);
    const NodeSet &ns = da_nodes(da);
    const Traversal &tr = da_traversal(da);
    auto pl = std::make_unique<dkt_plan>();
    pl->dim = mp.dim;
    pl->order = mp.order;
    pl->npe = tr.nPe;
    pl->op = op;
    pl->N = ns.nodes.size();
    pl->L = tr.leaves.size();
    pl->H = tr.pslot.size() / tr.nPe;
    const int npe = tr.nPe;
    if (pl->N >= (size_t)INT32_MAX || pl->L * (size_t)npe >= (size_t)INT32_MAX)
      throw std::invalid_argument("mesh too large for 32-bit indices");
    if (tr.hg_node.size() >= (size_t)INT32_MAX) throw std::invalid_argument("hanging gather too large");

    // Elemental matrices.
    std::vector<double> km;
    if (op == DKT_OP_CAPTURED) {
      if (!op_data) throw std::invalid_argument("captured operator needs per-level matrices");
      km.assign(op_data, op_data + (size_t)(mp.maxd + 1) * npe * npe);
      pl->per_level = true;
      pl->hp = 0;
    } else if (op == DKT_OP_STIFFNESS || op == DKT_OP_MASS || op == DKT_OP_IDENTITY) {
      km = ref_matrix(mp.dim, mp.order, op);
      pl->hp = op == DKT_OP_STIFFNESS ? mp.dim - 2 : (op == DKT_OP_MASS ? mp.dim : 0);
    } else {
      throw std::invalid_argument("unknown operator");
    }
    pl->kmat64.upload(km);
    pl->kmat32.upload(cast<float>(km));
    const std::vector<double> W = interp_matrices(mp.dim, mp.order);
    pl->wint64.upload(W);
    pl->wint32.upload(cast<float>(W));

    // Leaf maps.
    std::vector<uint8_t> lev(pl->L);
    for (size_t e = 0; e < pl->L; e++) lev[e] = (uint8_t)tr.leaves[e].lev;
    pl->slot.upload(tr.slot);
    pl->level.upload(lev);
    pl->hang_of.upload(tr.hang_of);
    pl->child.upload(tr.child);
    std::vector<int32_t> hoff(tr.hg_off.begin(), tr.hg_off.end());
    pl->hg_off.upload(hoff.empty() ? std::vector<int32_t>{0} : hoff);
    pl->hg_node.upload(tr.hg_node);
    pl->hg_w64.upload(tr.hg_w);
    pl->hg_w32.upload(cast<float>(tr.hg_w));
    std::vector<uint8_t> dir(pl->N, 0);
    for (u32 b : ns.bdy) dir[b] = 1;
    pl->dirichlet.upload(dir);

    // Node-side CSR over contribution positions, in leaf SFC order then rank:
    // direct outputs at e*nPe+r, hanging transposes at L*nPe + h*nPe + s.
    std::vector<int64_t> cnt(pl->N + 1, 0);
    for (size_t e = 0; e < pl->L; e++) {
      const int32_t *sl = &tr.slot[e * npe];
      for (int r = 0; r < npe; r++)
        if (sl[r] >= 0) cnt[sl[r] + 1]++;
      const int h = tr.hang_of[e];
      if (h >= 0) {
        const int32_t *ps = &tr.pslot[(size_t)h * npe];
        for (int s = 0; s < npe; s++)
          if (ps[s] >= 0 && sl[s] < 0) cnt[ps[s] + 1]++;
      }
    }
    for (size_t n = 0; n < pl->N; n++) cnt[n + 1] += cnt[n];
    std::vector<int64_t> src(cnt[pl->N]);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (size_t e = 0; e < pl->L; e++) {
      const int32_t *sl = &tr.slot[e * npe];
      for (int r = 0; r < npe; r++)
        if (sl[r] >= 0) src[fill[sl[r]]++] = (int64_t)e * npe + r;
      const int h = tr.hang_of[e];
      if (h >= 0) {
        const int32_t *ps = &tr.pslot[(size_t)h * npe];
        for (int s = 0; s < npe; s++)
          if (ps[s] >= 0 && sl[s] < 0) src[fill[ps[s]]++] = (int64_t)pl->L * npe + (int64_t)h * npe + s;
      }
    }
    pl->csr_off.upload(cnt);
    pl->csr_src.upload(src);
    pl->nbuf = (pl->L + pl->H) * (size_t)npe;
    pl->host_bytes_maps = tr.slot.size() * 4 + lev.size() + tr.hang_of.size() * 4;
    *out = pl.release();
  });
}

extern "C" int dkt_plan_destroy(dkt_plan *pl) {
  delete pl;
  return 0;
}

extern "C" int dkt_plan_matvec(dkt_plan *pl, int dtype, const void *x, void *y, int nrhs, int masked,
                               void *stream) {
  return guarded([&] {
    if (!pl || !x || !y) throw std::invalid_argument("null pointer");
    if (nrhs < 1 || nrhs > 64) throw std::invalid_argument("nrhs must be 1..64");
    if (x == y) throw std::invalid_argument("x and y must not alias");
    hipStream_t st = (hipStream_t)stream;
    if (dtype == 0)
      run_matvec<double>(pl, nrhs, (const double *)x, (double *)y, masked != 0, st);
    else if (dtype == 1)
      run_matvec<float>(pl, nrhs, (const float *)x, (float *)y, masked != 0, st);
    else
      throw std::invalid_argument("dtype must be 0 (f64) or 1 (f32)");
  });
}

// info[0..]: N, L, H, nPe, leaf-map bytes, contribution entries.
extern "C" int dkt_plan_info(const dkt_plan *pl, double *info, int n) {
  return guarded([&] {
    if (!pl || !info) throw std::invalid_argument("null pointer");
    const double v[] = {(double)pl->N, (double)pl->L, (double)pl->H, (double)pl->npe,
                        (double)pl->host_bytes_maps, (double)pl->csr_src.n};
    for (int i = 0; i < n && i < (int)(sizeof v / sizeof v[0]); i++) info[i] = v[i];
  });
}
