// Device matvec plan and CDNA4 kernels for the flattened k-tree traversal.
//
// The reference computes y = A x by a recursive top-down/bottom-up traversal of
// the node list (fem::matvec_rec, FEM/include/matvec.h:243-548).  Here the
// traversal is flattened once per mesh (mesh.cpp) into per-leaf maps and each
// matvec is a gather -> elemental apply -> deterministic reduction on device.
//
// Tiled path (default): leaves are cut into SFC-contiguous tiles, one tile per
// workgroup.  Per tile:
//   1. the tile's distinct nodes are gathered from x into LDS (x_loc);
//   2. one work-item per leaf builds its element vector from x_loc (16-bit
//      local indices), interpolating missing (hanging) ranks from its parent
//      group's nodes in registers (IKD_Parent2Child, FEM/include/refel.h:214-240)
//      plus the rare stale parent-buffer terms;
//   3. v = h^k K_ref u in registers;
//   4. direct outputs are added into an LDS accumulator y_loc in nPe phases, one
//      per rank (two disjoint leaves never share a node at the same rank, so a
//      phase has no write conflicts);
//   5. the transposed interpolation of missing ranks (matvec.h:482-522) is
//      summed per parent group and added in host-coloured phases;
//   6. nodes touched by this tile only are written to y; nodes shared with
//      other tiles are written as partials, summed in tile order by
//      fixup_kernel.
// The reduction order is fixed by the plan: results are bitwise reproducible,
// with no atomics.
//
// Simple path (flag DKT_PLAN_SIMPLE): a per-leaf kernel writing every
// contribution to a buffer, and a per-node kernel summing them through a CSR.
// It is kept as an independent device check of the tiled path.
//
// Multi-RHS vectors are node-major, [node][nrhs] (the reference's
// dof-interleaved layout, include/oda.h:296-322).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
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

#define HIP_OK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw std::runtime_error(std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr uint16_t kNone = 0xFFFF;
constexpr int kTileThreads = 256;  // one leaf per work-item

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
// Per-leaf tensor-product interpolation in registers.
// ip tables: IP[c][k][i] = l_k(child point i) for child half c (k parent rank).
// ---------------------------------------------------------------------------
template <int P>
struct Ip;
template <>
struct Ip<1> {
  __device__ static constexpr double w(int c, int k, int i) {
    // c=0: points {0,1/2}; c=1: points {1/2,1}.
    return (c == 0) ? (k == 0 ? (i == 0 ? 1.0 : 0.5) : (i == 0 ? 0.0 : 0.5))
                    : (k == 0 ? (i == 0 ? 0.5 : 0.0) : (i == 0 ? 0.5 : 1.0));
  }
};
template <>
struct Ip<2> {
  // Quadratic Lagrange on nodes {0,1/2,1} evaluated at child points.
  __device__ static constexpr double w(int c, int k, int i) {
    const double t = 0.5 * (c + 0.5 * i);  // child point in parent coords
    return k == 0 ? (t - 0.5) * (t - 1.0) / 0.5 : (k == 1 ? t * (t - 1.0) / -0.25 : t * (t - 0.5) / 0.5);
  }
};

template <int N, int D>
struct Pow {
  static constexpr int v = N * Pow<N, D - 1>::v;
};
template <int N>
struct Pow<N, 0> {
  static constexpr int v = 1;
};

// out[r] = sum_s W_c[r][s] in[s]   (forward: parent ranks -> child ranks)
template <typename T, int DIM, int P>
__device__ __forceinline__ void interp_fwd(const T *in, T *out, int c) {
  constexpr int n1 = P + 1, NPE = Pow<P + 1, DIM>::v;
  T a[NPE], b[NPE];
#pragma unroll
  for (int i = 0; i < NPE; i++) a[i] = in[i];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const int cb = (c >> d) & 1;
    constexpr int strides[4] = {1, n1, n1 * n1, n1 * n1 * n1};
    const int st = strides[d];
#pragma unroll
    for (int r = 0; r < NPE; r++) {
      const int id = (r / st) % n1;
      const int base = r - id * st;
      T acc = T(0);
#pragma unroll
      for (int k = 0; k < n1; k++) acc += (T)Ip<P>::w(cb, k, id) * a[base + k * st];
      b[r] = acc;
    }
#pragma unroll
    for (int i = 0; i < NPE; i++) a[i] = b[i];
  }
#pragma unroll
  for (int i = 0; i < NPE; i++) out[i] = a[i];
}

// out[s] = sum_r W_c[r][s] in[r]   (transpose: child ranks -> parent ranks)
template <typename T, int DIM, int P>
__device__ __forceinline__ void interp_bwd(const T *in, T *out, int c) {
  constexpr int n1 = P + 1, NPE = Pow<P + 1, DIM>::v;
  T a[NPE], b[NPE];
#pragma unroll
  for (int i = 0; i < NPE; i++) a[i] = in[i];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const int cb = (c >> d) & 1;
    constexpr int strides[4] = {1, n1, n1 * n1, n1 * n1 * n1};
    const int st = strides[d];
#pragma unroll
    for (int s = 0; s < NPE; s++) {
      const int kd = (s / st) % n1;
      const int base = s - kd * st;
      T acc = T(0);
#pragma unroll
      for (int i = 0; i < n1; i++) acc += (T)Ip<P>::w(cb, kd, i) * a[base + i * st];
      b[s] = acc;
    }
#pragma unroll
    for (int i = 0; i < NPE; i++) a[i] = b[i];
  }
#pragma unroll
  for (int i = 0; i < NPE; i++) out[i] = a[i];
}

// ---------------------------------------------------------------------------
// Tiled kernel.
// ---------------------------------------------------------------------------
struct TileHdr {
  int32_t leaf0, nleaf;   // leaf range
  int32_t node0, nnode;   // local node list range (tile_nodes / tile_out)
  int32_t grp0, ngrp;     // parent groups
  int32_t hph0, nhph;     // hanging-phase offsets range (hph_off has nhph+1 entries from hph0)
  int32_t st0, nst;       // stale terms
};

struct GroupRec {
  uint16_t first;  // first leaf (tile-local)
  uint8_t count;   // number of consecutive leaves
  uint8_t pad;
};

// Leaf aux word: bits 0-4 level, 5-8 child index, 9 hanging, 16-31 group (tile-local).
__device__ __forceinline__ int aux_level(uint32_t a) { return a & 31u; }
__device__ __forceinline__ int aux_child(uint32_t a) { return (a >> 5) & 15u; }
__device__ __forceinline__ bool aux_hang(uint32_t a) { return (a >> 9) & 1u; }
__device__ __forceinline__ int aux_group(uint32_t a) { return a >> 16; }

template <typename T, int DIM, int P>
__global__ void __launch_bounds__(kTileThreads) tile_kernel(
    const TileHdr *__restrict__ tiles, const int32_t *__restrict__ tile_nodes, const int32_t *__restrict__ tile_out,
    const uint16_t *__restrict__ leaf_loc, const uint32_t *__restrict__ leaf_aux, const GroupRec *__restrict__ groups,
    const uint16_t *__restrict__ group_ploc, const int32_t *__restrict__ hph_off, const uint32_t *__restrict__ hentries,
    const uint32_t *__restrict__ stale_key, const T *__restrict__ stale_w, const T *__restrict__ kref, int hp,
    const uint8_t *__restrict__ dirichlet, int nrhs, int ld, const T *__restrict__ x, T *__restrict__ y,
    T *__restrict__ partial) {
  constexpr int NPE = Pow<P + 1, DIM>::v;
  extern __shared__ unsigned char smem_raw[];
  const TileHdr th = tiles[blockIdx.x];
  T *x_loc = reinterpret_cast<T *>(smem_raw);
  T *y_loc = x_loc + (size_t)th.nnode * nrhs;
  T *w_loc = y_loc + (size_t)th.nnode * nrhs;  // [nleaf][NPE] transposed hanging outputs (one rhs at a time)
  const int tid = threadIdx.x;

  // 1. Stage the tile's nodes.
  for (int j = tid; j < th.nnode; j += blockDim.x) {
    const int g = tile_nodes[th.node0 + j];
    const bool z = dirichlet && dirichlet[g];
    for (int q = 0; q < nrhs; q++) {
      x_loc[j * nrhs + q] = z ? T(0) : x[(size_t)g * ld + q];
      y_loc[j * nrhs + q] = T(0);
    }
  }
  __syncthreads();

  const bool active = tid < th.nleaf;
  uint16_t loc[NPE];
  uint32_t aux = 0;
  if (active) {
    const uint16_t *lp = leaf_loc + (size_t)(th.leaf0 + tid) * NPE;
#pragma unroll
    for (int r = 0; r < NPE; r++) loc[r] = lp[r];
    aux = leaf_aux[th.leaf0 + tid];
  } else {
#pragma unroll
    for (int r = 0; r < NPE; r++) loc[r] = kNone;
  }
  const bool hang = active && aux_hang(aux);
  const int child = aux_child(aux);
  const T scale = (T)ldexp(1.0, -aux_level(aux) * hp);
  const uint16_t *ploc = hang ? group_ploc + (size_t)(th.grp0 + aux_group(aux)) * NPE : nullptr;

  for (int q = 0; q < nrhs; q++) {
    T u[NPE], v[NPE];
    if (active) {
      if (hang) {
        T xp[NPE], xi[NPE];
#pragma unroll
        for (int s = 0; s < NPE; s++) xp[s] = ploc[s] != kNone ? x_loc[ploc[s] * nrhs + q] : T(0);
        interp_fwd<T, DIM, P>(xp, xi, child);
        for (int k = 0; k < th.nst; k++) {  // stale parent-buffer terms (rare)
          const uint32_t key = stale_key[th.st0 + k];
          if ((int)(key >> 24) == tid) {
            const int r = (key >> 16) & 255, j = key & 0xFFFF;
#pragma unroll
            for (int rr = 0; rr < NPE; rr++)
              if (rr == r) xi[rr] += stale_w[th.st0 + k] * x_loc[j * nrhs + q];
          }
        }
#pragma unroll
        for (int r = 0; r < NPE; r++) u[r] = loc[r] != kNone ? x_loc[loc[r] * nrhs + q] : xi[r];
      } else {
#pragma unroll
        for (int r = 0; r < NPE; r++) u[r] = loc[r] != kNone ? x_loc[loc[r] * nrhs + q] : T(0);
      }
#pragma unroll
      for (int i = 0; i < NPE; i++) {
        T acc = T(0);
#pragma unroll
        for (int j = 0; j < NPE; j++) acc += kref[i * NPE + j] * u[j];
        v[i] = scale * acc;
      }
    }
    // 4. Direct outputs, one phase per rank.
#pragma unroll
    for (int r = 0; r < NPE; r++) {
      __syncthreads();
      if (loc[r] != kNone) y_loc[loc[r] * nrhs + q] += v[r];
    }
    // 5. Transposed interpolation of the missing ranks.
    if (th.nhph > 0) {
      if (active) {
        T vm[NPE], wv[NPE];
#pragma unroll
        for (int r = 0; r < NPE; r++) vm[r] = (hang && loc[r] == kNone) ? v[r] : T(0);
        interp_bwd<T, DIM, P>(vm, wv, child);
#pragma unroll
        for (int s = 0; s < NPE; s++) w_loc[tid * NPE + s] = (hang && loc[s] == kNone) ? wv[s] : T(0);
      }
      for (int ph = 0; ph < th.nhph; ph++) {
        __syncthreads();
        const int a = hph_off[th.hph0 + ph], b = hph_off[th.hph0 + ph + 1];
        for (int k = a + tid; k < b; k += blockDim.x) {
          const uint32_t ent = hentries[k];
          const int g = ent >> 8, s = ent & 255;
          const GroupRec gr = groups[th.grp0 + g];
          T acc = T(0);
          for (int l = 0; l < gr.count; l++) acc += w_loc[(gr.first + l) * NPE + s];
          const int j = group_ploc[(size_t)(th.grp0 + g) * NPE + s];
          y_loc[j * nrhs + q] += acc;
        }
      }
    }
  }
  __syncthreads();
  // 6. Write out: owned nodes to y, shared nodes to their partial slot.
  for (int j = tid; j < th.nnode; j += blockDim.x) {
    const int o = tile_out[th.node0 + j];
    if (o == INT32_MIN) continue;  // gathered only, no contribution from this tile
    if (o >= 0) {
      const bool z = dirichlet && dirichlet[o];
      for (int q = 0; q < nrhs; q++) y[(size_t)o * ld + q] = z ? T(0) : y_loc[j * nrhs + q];
    } else {
      const size_t slot = (size_t)(-o - 1);
      for (int q = 0; q < nrhs; q++) partial[slot * nrhs + q] = y_loc[j * nrhs + q];
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) fixup_kernel(int ns, const int32_t *__restrict__ snode,
                                                    const int32_t *__restrict__ soff, const T *__restrict__ partial,
                                                    const uint8_t *__restrict__ dirichlet, int nrhs, int ld,
                                                    T *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const int n = snode[i];
  const int a = soff[i], b = soff[i + 1];
  const bool z = dirichlet && dirichlet[n];
  for (int q = 0; q < nrhs; q++) {
    T acc = T(0);
    for (int k = a; k < b; k++) acc += partial[(size_t)k * nrhs + q];
    y[(size_t)n * ld + q] = z ? T(0) : acc;
  }
}

// ---------------------------------------------------------------------------
// Simple path kernels.
// ---------------------------------------------------------------------------
template <typename T, int DIM, int P>
__global__ void __launch_bounds__(256) leaf_kernel(
    int L, int nrhs, const T *__restrict__ x, const int32_t *__restrict__ slot, const uint8_t *__restrict__ level,
    const int32_t *__restrict__ hang_of, const uint8_t *__restrict__ child, const int32_t *__restrict__ pslot,
    const int32_t *__restrict__ hc_off, const uint8_t *__restrict__ hc_rank, const int32_t *__restrict__ hc_node,
    const T *__restrict__ hc_w, const T *__restrict__ kmat, int kmat_per_level, const uint8_t *__restrict__ dirichlet,
    int hp, T *__restrict__ vout, T *__restrict__ wout) {
  constexpr int NPE = Pow<P + 1, DIM>::v;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L) return;
  int32_t s[NPE];
#pragma unroll
  for (int r = 0; r < NPE; r++) s[r] = slot[(size_t)e * NPE + r];
  const int lev = level[e];
  const int h = hang_of[e];
  const int c = child[e];
  const T scale = kmat_per_level ? T(1) : (T)ldexp(1.0, -lev * hp);
  const T *K = kmat + (kmat_per_level ? (size_t)lev * NPE * NPE : 0);
  auto xv = [&](int n, int q) -> T { return (dirichlet && dirichlet[n]) ? T(0) : x[(size_t)n * nrhs + q]; };
  for (int q = 0; q < nrhs; q++) {
    T u[NPE];
    if (h >= 0) {
      T xp[NPE], xi[NPE];
#pragma unroll
      for (int t = 0; t < NPE; t++) {
        const int n = pslot[(size_t)h * NPE + t];
        xp[t] = n >= 0 ? xv(n, q) : T(0);
      }
      interp_fwd<T, DIM, P>(xp, xi, c);
      for (int k = hc_off[h]; k < hc_off[h + 1]; k++) {
        const int r = hc_rank[k];
#pragma unroll
        for (int rr = 0; rr < NPE; rr++)
          if (rr == r) xi[rr] += hc_w[k] * xv(hc_node[k], q);
      }
#pragma unroll
      for (int r = 0; r < NPE; r++) u[r] = s[r] >= 0 ? xv(s[r], q) : xi[r];
    } else {
#pragma unroll
      for (int r = 0; r < NPE; r++) u[r] = s[r] >= 0 ? xv(s[r], q) : T(0);
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
      T vm[NPE], wv[NPE];
#pragma unroll
      for (int r = 0; r < NPE; r++) vm[r] = s[r] < 0 ? v[r] : T(0);
      interp_bwd<T, DIM, P>(vm, wv, c);
#pragma unroll
      for (int t = 0; t < NPE; t++) wout[((size_t)h * NPE + t) * nrhs + q] = wv[t];
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) node_kernel(int N, int nrhs, const int64_t *__restrict__ off,
                                                   const int32_t *__restrict__ src, const T *__restrict__ buf,
                                                   const uint8_t *__restrict__ dirichlet, T *__restrict__ y) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int64_t a = off[n], b = off[n + 1];
  const bool zero = dirichlet && dirichlet[n];
  for (int q = 0; q < nrhs; q++) {
    T acc = T(0);
    for (int64_t k = a; k < b; k++) acc += buf[(size_t)src[k] * nrhs + q];
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

template <typename T>
std::vector<T> cast(const std::vector<double> &v) {
  return std::vector<T>(v.begin(), v.end());
}

}  // namespace

struct dkt_plan {
  int dim = 0, order = 0, npe = 0, op = 0, hp = 0, maxd = 0;
  size_t N = 0, L = 0, H = 0;
  bool per_level = false, simple = false;
  DevBuf<uint8_t> dirichlet;
  DevBuf<double> kmat64;
  DevBuf<float> kmat32;
  // simple path
  DevBuf<int32_t> slot, hang_of, pslot, hc_off, hc_node, csr_src;
  DevBuf<uint8_t> level, child, hc_rank;
  DevBuf<double> hc_w64;
  DevBuf<float> hc_w32;
  DevBuf<int64_t> csr_off;
  size_t nbuf = 0;
  // tiled path
  size_t ntiles = 0, nshared = 0, npartial = 0, max_tile_nodes = 0;
  int tile_leaves = kTileThreads;
  DevBuf<TileHdr> tiles;
  DevBuf<int32_t> tile_nodes, tile_out, hph_off, snode, soff;
  DevBuf<uint16_t> leaf_loc, group_ploc;
  DevBuf<uint32_t> leaf_aux, hentries, stale_key;
  DevBuf<GroupRec> groups;
  DevBuf<double> stale_w64;
  DevBuf<float> stale_w32;
  // scratch
  DevBuf<double> scratch64;
  DevBuf<float> scratch32;
  double map_bytes = 0;
};

namespace {

// ---------------------------------------------------------------------------
// Plan construction.
// ---------------------------------------------------------------------------
void build_simple(dkt_plan *pl, const Traversal &tr) {
  const int npe = pl->npe;
  std::vector<uint8_t> lev(pl->L);
  for (size_t e = 0; e < pl->L; e++) lev[e] = (uint8_t)tr.leaves[e].lev;
  pl->slot.upload(tr.slot);
  pl->level.upload(lev);
  pl->hang_of.upload(tr.hang_of);
  pl->child.upload(tr.child);
  pl->pslot.upload(tr.pslot.empty() ? std::vector<int32_t>(npe, -1) : tr.pslot);
  std::vector<int32_t> hco(tr.hc_off.begin(), tr.hc_off.end());
  pl->hc_off.upload(hco);
  pl->hc_rank.upload(tr.hc_rank.empty() ? std::vector<uint8_t>{0} : tr.hc_rank);
  pl->hc_node.upload(tr.hc_node.empty() ? std::vector<int32_t>{0} : tr.hc_node);
  pl->hc_w64.upload(tr.hc_w.empty() ? std::vector<double>{0} : tr.hc_w);
  pl->hc_w32.upload(cast<float>(tr.hc_w.empty() ? std::vector<double>{0} : tr.hc_w));
  // Node-side CSR over contribution positions, in leaf SFC order then rank:
  // direct outputs at e*nPe+r, hanging transposes at L*nPe + h*nPe + s.
  std::vector<int64_t> cnt(pl->N + 1, 0);
  for (size_t e = 0; e < pl->L; e++) {
    const int32_t *sl = &tr.slot[e * npe];
    for (int r = 0; r < npe; r++)
      if (sl[r] >= 0) cnt[sl[r] + 1]++;
    const int h = tr.hang_of[e];
    if (h >= 0)
      for (int s = 0; s < npe; s++)
        if (tr.pslot[(size_t)h * npe + s] >= 0 && sl[s] < 0) cnt[tr.pslot[(size_t)h * npe + s] + 1]++;
  }
  for (size_t n = 0; n < pl->N; n++) cnt[n + 1] += cnt[n];
  std::vector<int32_t> src(cnt[pl->N]);
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  for (size_t e = 0; e < pl->L; e++) {
    const int32_t *sl = &tr.slot[e * npe];
    for (int r = 0; r < npe; r++)
      if (sl[r] >= 0) src[fill[sl[r]]++] = (int32_t)(e * npe + r);
    const int h = tr.hang_of[e];
    if (h >= 0)
      for (int s = 0; s < npe; s++) {
        const int32_t n = tr.pslot[(size_t)h * npe + s];
        if (n >= 0 && sl[s] < 0) src[fill[n]++] = (int32_t)(pl->L * npe + (size_t)h * npe + s);
      }
  }
  pl->csr_off.upload(cnt);
  pl->csr_src.upload(src);
  pl->nbuf = (pl->L + pl->H) * (size_t)npe;
}

void build_tiled(dkt_plan *pl, const Traversal &tr) {
  const int npe = pl->npe;
  const size_t L = pl->L;
  // Leaves per tile: the per-leaf LDS rows (nPe values) must leave room for the
  // tile's nodes within 64 KiB.
  const int T = npe > 16 ? kTileThreads / 2 : kTileThreads;
  pl->tile_leaves = T;
  const size_t ntiles = (L + T - 1) / T;
  pl->ntiles = ntiles;
  const int dim = pl->dim;

  std::vector<TileHdr> hdr(ntiles);
  std::vector<int32_t> tnodes, tout, hph;
  std::vector<uint16_t> lloc(L * (size_t)npe, kNone), gploc;
  std::vector<uint32_t> laux(L, 0), hent, skey;
  std::vector<GroupRec> grps;
  std::vector<double> sw;
  // Tiles touching each node as a contribution target (ordered, de-duplicated).
  std::vector<int32_t> last_tile(pl->N, -1), ntarget(pl->N, 0);
  std::vector<int32_t> local(pl->N, -1);

  // Pass 1: count target tiles per node.
  for (size_t t = 0; t < ntiles; t++) {
    const size_t l0 = t * T, l1 = std::min(L, l0 + T);
    for (size_t e = l0; e < l1; e++) {
      const int32_t *sl = &tr.slot[e * npe];
      auto mark = [&](int32_t n) {
        if (last_tile[n] != (int32_t)t) {
          last_tile[n] = (int32_t)t;
          ntarget[n]++;
        }
      };
      for (int r = 0; r < npe; r++)
        if (sl[r] >= 0) mark(sl[r]);
      const int h = tr.hang_of[e];
      if (h >= 0)
        for (int s = 0; s < npe; s++) {
          const int32_t n = tr.pslot[(size_t)h * npe + s];
          if (n >= 0 && sl[s] < 0) mark(n);
        }
    }
  }
  // Partial slots for shared nodes, consecutive per node, in tile order.
  std::vector<int32_t> pbase(pl->N, -1), pfill(pl->N, 0);
  std::vector<int32_t> snode, soff{0};
  int64_t np = 0;
  for (size_t n = 0; n < pl->N; n++)
    if (ntarget[n] > 1) {
      pbase[n] = (int32_t)np;
      np += ntarget[n];
      snode.push_back((int32_t)n);
      soff.push_back((int32_t)np);
      if (np >= INT32_MAX) throw std::runtime_error("too many partial sums");
    }
  pl->nshared = snode.size();
  pl->npartial = (size_t)np;

  // Pass 2: tiles.
  std::fill(last_tile.begin(), last_tile.end(), -1);
  size_t max_nodes = 0;
  for (size_t t = 0; t < ntiles; t++) {
    const size_t l0 = t * T, l1 = std::min(L, l0 + T);
    TileHdr &H = hdr[t];
    H.leaf0 = (int32_t)l0;
    H.nleaf = (int32_t)(l1 - l0);
    H.node0 = (int32_t)tnodes.size();
    H.grp0 = (int32_t)grps.size();
    H.hph0 = (int32_t)hph.size();
    H.st0 = (int32_t)skey.size();
    std::vector<int32_t> touched;
    std::vector<uint8_t> is_target;
    auto loc_of = [&](int32_t n, bool target) -> uint16_t {
      if (last_tile[n] != (int32_t)t) {
        last_tile[n] = (int32_t)t;
        local[n] = (int32_t)touched.size();
        touched.push_back(n);
        is_target.push_back(0);
      }
      if (target) is_target[local[n]] = 1;
      if (local[n] >= 0xFFFF) throw std::runtime_error("tile has too many nodes");
      return (uint16_t)local[n];
    };
    // Parent groups: consecutive hanging leaves with the same parent cell.
    std::vector<std::pair<int, int>> gl;  // (group id, rank) hanging-transpose targets
    int cur_group = -1;
    Oct cur_parent{};
    for (size_t e = l0; e < l1; e++) {
      const int32_t *sl = &tr.slot[e * npe];
      for (int r = 0; r < npe; r++)
        if (sl[r] >= 0) lloc[e * npe + r] = loc_of(sl[r], true);
      const Oct &c = tr.leaves[e];
      uint32_t aux = (c.lev & 31u) | ((uint32_t)tr.child[e] << 5);
      const int h = tr.hang_of[e];
      if (h >= 0) {
        aux |= 1u << 9;
        // Parent cell: the leaf's anchor with its own level bit cleared.
        Oct par = c;
        par.lev = c.lev - 1;
        const uint32_t pmask = ~(uint32_t)(((uint64_t)1 << (pl->maxd - (int)par.lev)) - 1);
        for (int d = 0; d < dim; d++) par.x[d] = c.x[d] & pmask;
        bool same = cur_group >= 0 && par.lev == cur_parent.lev;
        for (int d = 0; same && d < dim; d++) same = par.x[d] == cur_parent.x[d];
        const int32_t *ps = &tr.pslot[(size_t)h * npe];
        if (!same || grps[H.grp0 + cur_group].count == 255) {
          cur_group = (int)(grps.size() - H.grp0);
          cur_parent = par;
          GroupRec g;
          g.first = (uint16_t)(e - l0);
          g.count = 0;
          g.pad = 0;
          grps.push_back(g);
          for (int s = 0; s < npe; s++) gploc.push_back(ps[s] >= 0 ? loc_of(ps[s], false) : kNone);
        }
        grps[H.grp0 + cur_group].count++;
        aux |= (uint32_t)cur_group << 16;
        uint16_t *gp = &gploc[(size_t)(H.grp0 + cur_group) * npe];
        for (int s = 0; s < npe; s++)
          if (ps[s] >= 0 && sl[s] < 0) {
            loc_of(ps[s], true);
            gl.push_back({cur_group, s});
            (void)gp;
          }
        for (int64_t k = tr.hc_off[h]; k < tr.hc_off[h + 1]; k++) {
          const uint16_t j = loc_of(tr.hc_node[k], false);
          skey.push_back(((uint32_t)(e - l0) << 24) | ((uint32_t)tr.hc_rank[k] << 16) | j);
          sw.push_back(tr.hc_w[k]);
        }
      } else {
        cur_group = -1;
      }
      laux[e] = aux;
    }
    if ((size_t)H.nleaf > 256) throw std::runtime_error("tile too large for the stale-term key");
    H.ngrp = (int32_t)(grps.size() - H.grp0);
    H.nst = (int32_t)(skey.size() - H.st0);
    // Hanging-transpose entries: de-duplicate (group, rank), colour by target node.
    std::sort(gl.begin(), gl.end());
    gl.erase(std::unique(gl.begin(), gl.end()), gl.end());
    std::vector<int> phase_of(gl.size());
    std::unordered_map<int, int> used;  // local node -> number of phases used
    int nph = 0;
    for (size_t i = 0; i < gl.size(); i++) {
      const int j = gploc[(size_t)(H.grp0 + gl[i].first) * npe + gl[i].second];
      int &u = used[j];
      phase_of[i] = u++;
      nph = std::max(nph, u);
    }
    H.nhph = nph;
    std::vector<std::vector<uint32_t>> byph(nph);
    for (size_t i = 0; i < gl.size(); i++)
      byph[phase_of[i]].push_back(((uint32_t)gl[i].first << 8) | (uint32_t)gl[i].second);
    int32_t base = (int32_t)hent.size();
    hph.push_back(base);
    for (int ph = 0; ph < nph; ph++) {
      for (uint32_t v : byph[ph]) hent.push_back(v);
      hph.push_back((int32_t)hent.size());
    }
    if (nph == 0) hph.pop_back();
    // Output classification of the tile's nodes.
    H.nnode = (int32_t)touched.size();
    max_nodes = std::max(max_nodes, touched.size());
    for (size_t j = 0; j < touched.size(); j++) {
      const int32_t n = touched[j];
      tnodes.push_back(n);
      if (!is_target[j])
        tout.push_back(INT32_MIN);
      else if (ntarget[n] == 1)
        tout.push_back(n);
      else
        tout.push_back(-(pbase[n] + pfill[n]++) - 1);
    }
  }
  pl->max_tile_nodes = max_nodes;
  pl->tiles.upload(hdr);
  pl->tile_nodes.upload(tnodes);
  pl->tile_out.upload(tout);
  pl->leaf_loc.upload(lloc);
  pl->leaf_aux.upload(laux);
  pl->groups.upload(grps.empty() ? std::vector<GroupRec>(1) : grps);
  pl->group_ploc.upload(gploc.empty() ? std::vector<uint16_t>(npe, kNone) : gploc);
  pl->hph_off.upload(hph.empty() ? std::vector<int32_t>{0} : hph);
  pl->hentries.upload(hent.empty() ? std::vector<uint32_t>{0} : hent);
  pl->stale_key.upload(skey.empty() ? std::vector<uint32_t>{0} : skey);
  pl->stale_w64.upload(sw.empty() ? std::vector<double>{0} : sw);
  pl->stale_w32.upload(cast<float>(sw.empty() ? std::vector<double>{0} : sw));
  pl->snode.upload(snode.empty() ? std::vector<int32_t>{0} : snode);
  pl->soff.upload(soff);
  pl->map_bytes = (double)(hdr.size() * sizeof(TileHdr) + tnodes.size() * 8 + lloc.size() * 2 + laux.size() * 4 +
                           grps.size() * sizeof(GroupRec) + gploc.size() * 2 + hent.size() * 4 + skey.size() * 12 +
                           snode.size() * 8);
}

// ---------------------------------------------------------------------------
// Launch.
// ---------------------------------------------------------------------------
template <typename T, int DIM, int P>
void run_simple(dkt_plan *pl, int nrhs, const T *x, T *y, bool masked, hipStream_t st) {
  constexpr int NPE = Pow<P + 1, DIM>::v;
  const bool f64 = sizeof(T) == 8;
  const T *kmat = f64 ? (const T *)pl->kmat64.p : (const T *)pl->kmat32.p;
  const T *hcw = f64 ? (const T *)pl->hc_w64.p : (const T *)pl->hc_w32.p;
  const size_t need = pl->nbuf * (size_t)nrhs;
  T *buf;
  if (f64) {
    if (pl->scratch64.n < need) pl->scratch64.alloc(need);
    buf = (T *)pl->scratch64.p;
  } else {
    if (pl->scratch32.n < need) pl->scratch32.alloc(need);
    buf = (T *)pl->scratch32.p;
  }
  const uint8_t *dir = masked ? pl->dirichlet.p : nullptr;
  const int blk = 256;
  if (pl->L) {
    hipLaunchKernelGGL((leaf_kernel<T, DIM, P>), dim3((pl->L + blk - 1) / blk), dim3(blk), 0, st, (int)pl->L, nrhs, x,
                       pl->slot.p, pl->level.p, pl->hang_of.p, pl->child.p, pl->pslot.p, pl->hc_off.p,
                       pl->hc_rank.p, pl->hc_node.p, hcw, kmat, pl->per_level ? 1 : 0, dir, pl->hp, buf,
                       buf + pl->L * (size_t)NPE * nrhs);
    HIP_OK(hipGetLastError());
  }
  hipLaunchKernelGGL((node_kernel<T>), dim3((pl->N + blk - 1) / blk), dim3(blk), 0, st, (int)pl->N, nrhs,
                     pl->csr_off.p, pl->csr_src.p, buf, dir, y);
  HIP_OK(hipGetLastError());
}

template <typename T, int DIM, int P>
void run_tiled(dkt_plan *pl, int nrhs, const T *x, T *y, bool masked, hipStream_t st) {
  constexpr int NPE = Pow<P + 1, DIM>::v;
  const bool f64 = sizeof(T) == 8;
  const T *kmat = f64 ? (const T *)pl->kmat64.p : (const T *)pl->kmat32.p;
  const T *sw = f64 ? (const T *)pl->stale_w64.p : (const T *)pl->stale_w32.p;
  const size_t need = std::max<size_t>(1, pl->npartial * (size_t)nrhs);
  T *part;
  if (f64) {
    if (pl->scratch64.n < need) pl->scratch64.alloc(need);
    part = (T *)pl->scratch64.p;
  } else {
    if (pl->scratch32.n < need) pl->scratch32.alloc(need);
    part = (T *)pl->scratch32.p;
  }
  // Columns are processed in chunks whose LDS footprint stays within 64 KiB.
  const size_t fixed = (size_t)pl->tile_leaves * NPE * sizeof(T);
  const size_t per_col = 2 * pl->max_tile_nodes * sizeof(T);
  if (fixed + per_col > 64 * 1024) throw std::invalid_argument("tile does not fit in LDS");
  const int chunk = (int)std::min<size_t>((size_t)nrhs, (64 * 1024 - fixed) / per_col);
  const uint8_t *dir = masked ? pl->dirichlet.p : nullptr;
  for (int q0 = 0; q0 < nrhs; q0 += chunk) {
    const int nq = std::min(chunk, nrhs - q0);
    const size_t lds = fixed + per_col * nq;
    if (pl->ntiles) {
      hipLaunchKernelGGL((tile_kernel<T, DIM, P>), dim3(pl->ntiles), dim3(kTileThreads), lds, st, pl->tiles.p,
                         pl->tile_nodes.p, pl->tile_out.p, pl->leaf_loc.p, pl->leaf_aux.p, pl->groups.p,
                         pl->group_ploc.p, pl->hph_off.p, pl->hentries.p, pl->stale_key.p, sw, kmat, pl->hp, dir,
                         nq, nrhs, x + q0, y + q0, part);
      HIP_OK(hipGetLastError());
    }
    if (pl->nshared) {
      const int blk = 256;
      hipLaunchKernelGGL((fixup_kernel<T>), dim3((pl->nshared + blk - 1) / blk), dim3(blk), 0, st,
                         (int)pl->nshared, pl->snode.p, pl->soff.p, part, dir, nq, nrhs, y + q0);
      HIP_OK(hipGetLastError());
    }
  }
}

template <typename T>
void run_matvec(dkt_plan *pl, int nrhs, const T *x, T *y, bool masked, hipStream_t st) {
  const bool simple = pl->simple || pl->per_level;
#define DKT_DISPATCH(D, P)                                   \
  if (pl->dim == D && pl->order == P) {                      \
    if (simple)                                              \
      run_simple<T, D, P>(pl, nrhs, x, y, masked, st);       \
    else                                                     \
      run_tiled<T, D, P>(pl, nrhs, x, y, masked, st);        \
    return;                                                  \
  }
  DKT_DISPATCH(2, 1)
  DKT_DISPATCH(3, 1)
  DKT_DISPATCH(4, 1)
  DKT_DISPATCH(2, 2)
  DKT_DISPATCH(3, 2)
#undef DKT_DISPATCH
  throw std::invalid_argument("unsupported (dim, order)");
}

}  // namespace

extern "C" int dkt_plan_create(const dkt_da *da, int op, const double *op_data, int flags, dkt_plan **out) {
  return guarded([&] {
    if (!da || !out) throw std::invalid_argument("null pointer");
    const MeshParams &mp = da_params(da);
    const NodeSet &ns = da_nodes(da);
    const Traversal &tr = da_traversal(da);
    auto pl = std::make_unique<dkt_plan>();
    pl->dim = mp.dim;
    pl->order = mp.order;
    pl->maxd = mp.maxd;
    pl->npe = tr.nPe;
    pl->op = op;
    pl->N = ns.nodes.size();
    pl->L = tr.leaves.size();
    pl->H = tr.pslot.size() / tr.nPe;
    pl->simple = (flags & 1) != 0;
    const int npe = tr.nPe;
    if (pl->N >= (size_t)INT32_MAX || (pl->L + pl->H) * (size_t)npe >= (size_t)INT32_MAX)
      throw std::invalid_argument("mesh too large for 32-bit indices");
    std::vector<double> km;
    if (op == DKT_OP_CAPTURED) {
      if (!op_data) throw std::invalid_argument("captured operator needs per-level matrices");
      km.assign(op_data, op_data + (size_t)(mp.maxd + 1) * npe * npe);
      pl->per_level = true;
    } else if (op == DKT_OP_STIFFNESS || op == DKT_OP_MASS || op == DKT_OP_IDENTITY) {
      km = ref_matrix(mp.dim, mp.order, op);
      pl->hp = op == DKT_OP_STIFFNESS ? mp.dim - 2 : (op == DKT_OP_MASS ? mp.dim : 0);
    } else {
      throw std::invalid_argument("unknown operator");
    }
    pl->kmat64.upload(km);
    pl->kmat32.upload(cast<float>(km));
    std::vector<uint8_t> dir(pl->N, 0);
    for (u32 b : ns.bdy) dir[b] = 1;
    pl->dirichlet.upload(dir);
    if (pl->simple || pl->per_level)
      build_simple(pl.get(), tr);
    else
      build_tiled(pl.get(), tr);
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

// info[0..]: nodes, leaves, hanging leaves, nPe, device map bytes, tiles, shared nodes, partial slots,
// max nodes per tile.
extern "C" int dkt_plan_info(const dkt_plan *pl, double *info, int n) {
  return guarded([&] {
    if (!pl || !info) throw std::invalid_argument("null pointer");
    const double v[] = {(double)pl->N,        (double)pl->L,       (double)pl->H,
                        (double)pl->npe,      pl->map_bytes,       (double)pl->ntiles,
                        (double)pl->nshared,  (double)pl->npartial, (double)pl->max_tile_nodes};
    for (int i = 0; i < n && i < (int)(sizeof v / sizeof v[0]); i++) info[i] = v[i];
  });
}
