// Device matvec plan: host construction of the tiled layout (tile_kernel.h),
// the simple check path, launch and the C ABI (include/dkt.h).
//
// The reference computes y = A x by a recursive top-down/bottom-up traversal of
// the node list (fem::matvec_rec, FEM/include/matvec.h:243-548).  The
// traversal is flattened once per mesh (mesh.cpp); each matvec is then the
// tile kernel plus a small fix-up kernel on the caller's stream.
//
// Simple path (DKT_PLAN_SIMPLE): a per-leaf kernel writing every contribution
// to a buffer and a per-node kernel summing them through a CSR in leaf order;
// an independent device check of the tiled path, also used for per-level
// captured operators.
//
// Multi-RHS vectors are node-major, [node][nrhs] (the reference's
// dof-interleaved layout, include/oda.h:296-322).
#include <algorithm>
#include <atomic>
#include <exception>
#include <mutex>
#include <thread>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <set>
#include <unordered_map>
#include <vector>

#include "../../include/dkt.h"
#include "capi_common.h"
#include "mesh.h"
#include "plan_common.h"
#include "plan_impl.h"
#include "tile_kernel.h"
#include "conf_kernel.h"

namespace dkt {
const MeshParams &da_params(const dkt_da *da);
const NodeSet &da_nodes(const dkt_da *da);
Traversal &da_traversal_mut(const dkt_da *da);
Traversal &da_conforming_mut(const dkt_da *da);
thread_local bool g_dry_run = false;
}  // namespace dkt

using namespace dkt;

namespace {

template <typename T>
std::vector<T> nonempty(std::vector<T> v, T fill = T()) {
  if (v.empty()) v.push_back(fill);
  return v;
}

template <typename T>
std::vector<T> cast(const std::vector<double> &v) {
  return std::vector<T>(v.begin(), v.end());
}

// ---------------------------------------------------------------------------
// Simple path kernels.
// ---------------------------------------------------------------------------
template <typename T, int DIM, int P>
__global__ void __launch_bounds__(256) leaf_kernel(
    int L, int nrhs, const T *__restrict__ x, const int32_t *__restrict__ slot, const uint8_t *__restrict__ level,
    const int32_t *__restrict__ hang_of, const uint8_t *__restrict__ child, const int32_t *__restrict__ pslot,
    const int32_t *__restrict__ hc_off, const uint8_t *__restrict__ hc_rank, const int32_t *__restrict__ hc_node,
    const T *__restrict__ hc_w, const T *__restrict__ kmat, int kmat_per_level, const uint8_t *__restrict__ dir_in,
    int hp, int conforming, T *__restrict__ vout, T *__restrict__ wout, const int *__restrict__ cancel) {
  constexpr int NPE = Pow<P + 1, DIM>::v;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L || (cancel && *cancel)) return;
  int32_t s[NPE];
#pragma unroll
  for (int r = 0; r < NPE; r++) s[r] = slot[(size_t)e * NPE + r];
  const int lev = level[e];
  const int h = hang_of[e];
  const int c = child[e];
  const T scale = kmat_per_level ? T(1) : (T)ldexp(1.0, -lev * hp);
  const T *K = kmat + (kmat_per_level ? (size_t)lev * NPE * NPE : 0);
  auto xv = [&](int n, int q) -> T { return (dir_in && dir_in[n]) ? T(0) : x[(size_t)n * nrhs + q]; };
  for (int q = 0; q < nrhs; q++) {
    T u[NPE];
    if (h >= 0) {
      T xi[NPE];
#pragma unroll
      for (int t = 0; t < NPE; t++) {
        const int n = pslot[(size_t)h * NPE + t];
        xi[t] = n >= 0 ? xv(n, q) : T(0);
      }
      interp<T, DIM, P, false>(xi, c);
      for (int k = hc_off[h]; k < hc_off[h + 1]; k++) {
        const int r = hc_rank[k];
        const T add = hc_w[k] * xv(hc_node[k], q);
#pragma unroll
        for (int rr = 0; rr < NPE; rr++)
          if (rr == r) xi[rr] += add;
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
    if (h >= 0 && conforming) {  // transposed constraint terms
      for (int k = hc_off[h]; k < hc_off[h + 1]; k++) {
        const int r = hc_rank[k];
        T vr = T(0);
#pragma unroll
        for (int rr = 0; rr < NPE; rr++)
          if (rr == r) vr = v[rr];
        wout[(size_t)k * nrhs + q] = hc_w[k] * vr;
      }
    } else if (h >= 0) {
      T w[NPE];
#pragma unroll
      for (int r = 0; r < NPE; r++) w[r] = s[r] < 0 ? v[r] : T(0);
      interp<T, DIM, P, true>(w, c);
#pragma unroll
      for (int t = 0; t < NPE; t++) wout[((size_t)h * NPE + t) * nrhs + q] = w[t];
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) node_kernel(int N, int nrhs, const int64_t *__restrict__ off,
                                                   const int32_t *__restrict__ src, const T *__restrict__ buf,
                                                   const uint8_t *__restrict__ dirichlet, T *__restrict__ y,
                                                   const int *__restrict__ cancel) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N || (cancel && *cancel)) return;
  const int64_t a = off[n], b = off[n + 1];
  const bool zero = dirichlet && dirichlet[n];
  for (int q = 0; q < nrhs; q++) {
    T acc = T(0);
    for (int64_t k = a; k < b; k++) acc += buf[(size_t)src[k] * nrhs + q];
    y[(size_t)n * nrhs + q] = zero ? T(0) : acc;
  }
}

// ---------------------------------------------------------------------------
// Reference-element matrices (unit cube, axis 0 fastest).
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

// Simultaneous diagonalisation of the 1D p = 2 matrices: B (3 x 3) and lambda
// with M1 = B^T B and K1 = B^T diag(lambda) B, so that the element mass is
// (B^T)^(x dim) (B)^(x dim) and the stiffness (B^T)^(x dim) diag(sum_d lambda)
// B^(x dim).  M1 = L L^T (Cholesky), C = L^-1 K1 L^-T = Q diag(lambda) Q^T
// (Jacobi rotations), B = Q^T L^T.  Returns B row-major, then lambda.
// The device table of q2_apply in precision T: B and lambda from q2_factors,
// then D[r] = sum_d lambda[i_d(r)] for the 3^dim ranks, summed in T from zero,
// axis 0 first.
template <typename T>
std::vector<T> q2_table(const std::vector<double> &kf, int dim) {
  std::vector<T> out(kf.begin(), kf.begin() + 12);
  int n = 1;
  for (int d = 0; d < dim; d++) n *= 3;
  for (int r = 0; r < n; r++) {
    T s = T(0);
    for (int d = 0, rr = r; d < dim; d++, rr /= 3) s += (T)kf[9 + rr % 3];
    out.push_back(s);
  }
  return out;
}

std::vector<double> q2_factors() {
  std::vector<double> K1, M1;
  one_d(2, K1, M1);
  double L[3][3] = {};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j <= i; j++) {
      double s = M1[i * 3 + j];
      for (int k = 0; k < j; k++) s -= L[i][k] * L[j][k];
      L[i][j] = i == j ? std::sqrt(s) : s / L[j][j];
    }
  // Li = L^-1 (lower triangular)
  double Li[3][3] = {};
  for (int i = 0; i < 3; i++) {
    Li[i][i] = 1.0 / L[i][i];
    for (int j = 0; j < i; j++) {
      double s = 0.0;
      for (int k = j; k < i; k++) s -= L[i][k] * Li[k][j];
      Li[i][j] = s / L[i][i];
    }
  }
  double C[3][3] = {}, Q[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) C[i][j] += Li[i][a] * K1[a * 3 + b] * Li[j][b];
  for (int sweep = 0; sweep < 50; sweep++) {
    double off = 0.0;
    for (int p = 0; p < 3; p++)
      for (int q = p + 1; q < 3; q++) off += C[p][q] * C[p][q];
    if (off < 1e-40) break;
    for (int p = 0; p < 3; p++)
      for (int q = p + 1; q < 3; q++) {
        if (std::fabs(C[p][q]) < 1e-300) continue;
        const double th = 0.5 * (C[q][q] - C[p][p]) / C[p][q];
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; k++) {  // C <- J^T C J, Q <- Q J
          const double ckp = C[k][p], ckq = C[k][q];
          C[k][p] = c * ckp - s * ckq;
          C[k][q] = s * ckp + c * ckq;
        }
        for (int k = 0; k < 3; k++) {
          const double cpk = C[p][k], cqk = C[q][k];
          C[p][k] = c * cpk - s * cqk;
          C[q][k] = s * cpk + c * cqk;
        }
        for (int k = 0; k < 3; k++) {
          const double qkp = Q[k][p], qkq = Q[k][q];
          Q[k][p] = c * qkp - s * qkq;
          Q[k][q] = s * qkp + c * qkq;
        }
      }
  }
  std::vector<double> out(12, 0.0);
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++)
      for (int a = 0; a < 3; a++) out[i * 3 + k] += Q[a][i] * L[k][a];  // (Q^T L^T)[i][k]
  for (int i = 0; i < 3; i++) out[9 + i] = C[i][i];
  // Check the factorisation against the 1D matrices.
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      double m = 0.0, k = 0.0;
      for (int i = 0; i < 3; i++) {
        m += out[i * 3 + a] * out[i * 3 + b];
        k += out[i * 3 + a] * out[9 + i] * out[i * 3 + b];
      }
      if (std::fabs(m - M1[a * 3 + b]) > 1e-14 || std::fabs(k - K1[a * 3 + b]) > 1e-13)
        throw std::runtime_error("p = 2 sum factorisation check failed");
    }
  return out;
}

// Stiffness = sum over axes t of tensor products with the 1D stiffness on t and
// the 1D mass elsewhere; mass = tensor product of 1D masses.
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

}  // namespace

namespace {

// ---------------------------------------------------------------------------
// Simple path construction.
// ---------------------------------------------------------------------------
// tr: the reference traversal, or for conforming plans the conforming element
// records (build_conforming).
void build_simple(dkt_plan *pl, const MeshParams &mp, Traversal &tr) {
  const int npe = pl->npe;
  const bool conf = pl->conforming;
  if (!conf) compute_stale_terms(mp, tr);
  const std::vector<int64_t> &t_off = conf ? tr.cf_off : tr.hc_off;
  const std::vector<uint8_t> &t_rank = conf ? tr.cf_rank : tr.hc_rank;
  const std::vector<int32_t> &t_node = conf ? tr.cf_node : tr.hc_node;
  const std::vector<double> &t_w = conf ? tr.cf_w : tr.hc_w;
  std::vector<uint8_t> lev(pl->L);
  for (size_t e = 0; e < pl->L; e++) lev[e] = (uint8_t)tr.leaves[e].lev;
  pl->slot.upload(tr.slot);
  pl->level.upload(lev);
  pl->hang_of.upload(tr.hang_of);
  pl->child.upload(tr.child);
  // Conforming: the constraint terms carry the whole value of a missing node,
  // so no parent-level interpolation is added.
  const size_t pslots = std::max<size_t>(conf ? (t_off.size() - 1) * npe : tr.pslot.size(), npe);
  pl->pslot.upload(tr.pslot.empty() || conf ? std::vector<int32_t>(pslots, -1) : tr.pslot);
  pl->hc_off.upload(std::vector<int32_t>(t_off.begin(), t_off.end()));
  pl->hc_rank.upload(nonempty(t_rank));
  pl->hc_node.upload(nonempty(t_node));
  pl->hc_w64.upload(nonempty(t_w));
  pl->hc_w32.upload(cast<float>(nonempty(t_w)));
  if ((pl->L + pl->H) * (size_t)npe + t_node.size() >= (size_t)INT32_MAX)
    throw std::invalid_argument("mesh too large for 32-bit contribution indices");
  // Node-side CSR over contribution positions, in leaf SFC order then rank:
  // direct outputs at e*nPe+r; hanging transposes at L*nPe + h*nPe + s, or
  // (conforming) the transposed constraint term k at L*nPe + k.
  std::vector<int64_t> cnt(pl->N + 1, 0);
  for (size_t e = 0; e < pl->L; e++) {
    const int32_t *sl = &tr.slot[e * npe];
    for (int r = 0; r < npe; r++)
      if (sl[r] >= 0) cnt[sl[r] + 1]++;
    const int h = tr.hang_of[e];
    if (h < 0) continue;
    if (conf) {
      for (int64_t k = t_off[h]; k < t_off[h + 1]; k++) cnt[t_node[k] + 1]++;
    } else {
      for (int s = 0; s < npe; s++)
        if (tr.pslot[(size_t)h * npe + s] >= 0 && sl[s] < 0) cnt[tr.pslot[(size_t)h * npe + s] + 1]++;
    }
  }
  for (size_t n = 0; n < pl->N; n++) cnt[n + 1] += cnt[n];
  std::vector<int32_t> src(cnt[pl->N]);
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  for (size_t e = 0; e < pl->L; e++) {
    const int32_t *sl = &tr.slot[e * npe];
    for (int r = 0; r < npe; r++)
      if (sl[r] >= 0) src[fill[sl[r]]++] = (int32_t)(e * npe + r);
    const int h = tr.hang_of[e];
    if (h < 0) continue;
    if (conf) {
      for (int64_t k = t_off[h]; k < t_off[h + 1]; k++) src[fill[t_node[k]]++] = (int32_t)(pl->L * npe + k);
    } else {
      for (int s = 0; s < npe; s++) {
        const int32_t n = tr.pslot[(size_t)h * npe + s];
        if (n >= 0 && sl[s] < 0) src[fill[n]++] = (int32_t)(pl->L * npe + (size_t)h * npe + s);
      }
    }
  }
  pl->csr_off.upload(cnt);
  pl->csr_src.upload(src);
  pl->nbuf = conf ? pl->L * (size_t)npe + t_node.size() : (pl->L + pl->H) * (size_t)npe;
  pl->map_bytes = (double)(tr.slot.size() * 4 + pl->L * 6 + pslots * 4 + t_off.size() * 4 + t_node.size() * 17 +
                           cnt.size() * 8 + src.size() * 4);
}

// ---------------------------------------------------------------------------
// Tiled path construction.
// ---------------------------------------------------------------------------
struct TiledHost {
  std::vector<TileHdr> hdr;
  std::vector<int32_t> tnodes, tout, snode, soff;
  std::vector<uint16_t> lloc, gploc;
  std::vector<uint32_t> laux, ckey;
  std::vector<uint32_t> hent;
  std::vector<uint64_t> blk;
  std::vector<uint16_t> rdr;
  std::vector<double> cw;
  int max_nodes = 0, max_grp = 0, max_he = 0, max_rd = 0, max_cr = 0, max_hang = 0;
  size_t nshared = 0, npartial = 0;
  std::vector<double> stats;
};

// Tile-local numbering of global node ids (first touch), an open-addressing
// map that is cleared in time proportional to its use.
struct LocalMap {
  std::vector<int32_t> key, val;
  std::vector<uint32_t> used;
  LocalMap() : key(1 << 15, -1), val(1 << 15, 0) {}
  void clear() {
    for (uint32_t u : used) key[u] = -1;
    used.clear();
  }
  // Returns the slot of n (key[slot] == n after the call) and whether it was new.
  uint32_t find(int32_t n, bool *fresh) {
    const uint32_t mask = (uint32_t)key.size() - 1;
    uint32_t h = ((uint32_t)n * 2654435761u) & mask;
    while (key[h] != -1 && key[h] != n) h = (h + 1) & mask;
    *fresh = key[h] == -1;
    if (*fresh) {
      if (used.size() * 2 >= key.size()) throw std::runtime_error("tile has too many nodes");
      key[h] = n;
      used.push_back(h);
    }
    return h;
  }
};

TiledHost build_tiled_host(const MeshParams &mp, const Traversal &tr, size_t N, int T, int n_groups) {
  const int npe = tr.nPe;
  const int dim = mp.dim;
  const size_t L = tr.leaves.size();
  StaleBuffer sb(mp, tr);
  // Tiles: runs of SFC-consecutive leaves, at most T leaves, kMaxReaders
  // stale-buffer readers (a reader is a hanging leaf with a missing rank that
  // interpolates from a parent rank with no parent-level node) and T / 4
  // sibling groups each.
  auto parent_of = [&](size_t e) {
    Oct par = tr.leaves[e];
    par.lev -= 1;
    const uint32_t pmask = ~(uint32_t)(((uint64_t)1 << (mp.maxd - (int)par.lev)) - 1);
    for (int d = 0; d < dim; d++) par.x[d] &= pmask;
    return par;
  };
  auto same_oct = [&](const Oct &a, const Oct &b) {
    bool same = a.lev == b.lev;
    for (int d = 0; same && d < dim; d++) same = a.x[d] == b.x[d];
    return same;
  };
  const int max_groups = std::max(1, T / 4);
  // wmask[c * nPe + r]: parent slots with a nonzero weight for child c, rank r.
  std::vector<uint32_t> wmask((size_t)(1 << dim) * npe, 0);
  for (int c = 0; c < (1 << dim); c++)
    for (int r = 0; r < npe; r++)
      for (const auto &tw : sb.weights(c, r)) wmask[(size_t)c * npe + r] |= 1u << tw.first;
  std::vector<size_t> bounds{0};
  {
    int nrd = 0, ngrp = 0;
    bool prev_hang = false;
    Oct prev_par{};
    for (size_t e = 0; e < L; e++) {
      bool reader = false;
      const int h = tr.hang_of[e];
      Oct par{};
      bool new_group = false;
      if (h >= 0) {
        const int32_t *sl = &tr.slot[e * npe];
        const int32_t *ps = &tr.pslot[(size_t)h * npe];
        uint32_t need = 0, abs_m = 0;
        for (int r = 0; r < npe; r++) {
          if (sl[r] < 0) need |= wmask[(size_t)tr.child[e] * npe + r];
          if (ps[r] < 0) abs_m |= 1u << r;
        }
        reader = (need & abs_m) != 0;
        par = parent_of(e);
        new_group = !(prev_hang && same_oct(par, prev_par) && (e - bounds.back()) % 64 != 0);
      }
      if (e - bounds.back() == (size_t)T || (reader && nrd == kMaxReaders) || (new_group && ngrp == max_groups)) {
        bounds.push_back(e);
        nrd = 0;
        ngrp = 0;
      }
      nrd += reader ? 1 : 0;
      ngrp += new_group ? 1 : 0;
      prev_hang = h >= 0;
      prev_par = par;
    }
    if (bounds.back() != L) bounds.push_back(L);
  }
  const size_t ntiles = bounds.size() - 1;
  TiledHost o;
  o.hdr.resize(ntiles);
  o.lloc.assign(L * (size_t)npe, kNone);
  o.laux.assign(L, 0);
  std::vector<int32_t> last_tile(N, -1), ntarget(N, 0);

  // Pass 1: tiles that write each node (as a direct or hanging-transpose target).
  for (size_t t = 0; t < ntiles; t++) {
    const size_t l0 = bounds[t], l1 = bounds[t + 1];
    auto mark = [&](int32_t n) {
      if (last_tile[n] != (int32_t)t) {
        last_tile[n] = (int32_t)t;
        ntarget[n]++;
      }
    };
    for (size_t e = l0; e < l1; e++) {
      const int32_t *sl = &tr.slot[e * npe];
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
  std::vector<int32_t> pbase(N, -1), pfill(N, 0);
  o.soff.push_back(0);
  int64_t np = 0;
  for (size_t n = 0; n < N; n++)
    if (ntarget[n] > 1) {
      pbase[n] = (int32_t)np;
      np += ntarget[n];
      o.snode.push_back((int32_t)n);
      o.soff.push_back((int32_t)np);
      if (np >= INT32_MAX) throw std::runtime_error("too many partial sums");
    }
  o.nshared = o.snode.size();
  o.npartial = (size_t)np;

  // Pass 2: tiles, built concurrently (each tile writes its own leaf rows and
  // a TileOut), then concatenated in tile order.
  struct TileOut {
    std::vector<uint16_t> gploc, rdr;
    std::vector<uint32_t> hent, ckey;
    std::vector<uint64_t> blk;
    std::vector<double> cw;
    std::vector<int32_t> touched;
    std::vector<uint8_t> is_target;
    int ngrp = 0, nhang = 0;
    double cone_sum = 0, cone_n = 0, span = -1;
  };
  // Tile-local node numbers: open addressing over global node ids.
  std::vector<TileOut> outs(ntiles);
  auto build_tile = [&](size_t t, TileOut &to, LocalMap &lm, std::unordered_map<uint64_t, Combo> &memo) {
    const size_t l0 = bounds[t], l1 = bounds[t + 1];
    lm.clear();
    std::vector<int32_t> &touched = to.touched;
    std::vector<uint8_t> &is_target = to.is_target;
    auto loc_of = [&](int32_t n, bool target) -> uint16_t {
      bool fresh;
      const uint32_t h = lm.find(n, &fresh);
      if (fresh) {
        lm.val[h] = (int32_t)touched.size();
        touched.push_back(n);
        is_target.push_back(0);
      }
      const int32_t l = lm.val[h];
      if (target) is_target[l] = 1;
      if (l >= 0xFFFF) throw std::runtime_error("tile has too many nodes");
      return (uint16_t)l;
    };
    std::vector<uint32_t> gl;  // per group: parent ranks that are hanging-transpose targets
    std::vector<size_t> hang_leaves;  // leaf ids of the tile's hanging leaves, in order
    int cur_group = -1;
    Oct cur_parent{};
    for (size_t e = l0; e < l1; e++) {
      const int32_t *sl = &tr.slot[e * npe];
      for (int r = 0; r < npe; r++)
        if (sl[r] >= 0) o.lloc[e * npe + r] = loc_of(sl[r], true);
      const Oct &c = tr.leaves[e];
      uint32_t aux = (c.lev & 31u) | ((uint32_t)tr.child[e] << 5);
      const int h = tr.hang_of[e];
      if (h < 0) {
        cur_group = -1;
        o.laux[e] = aux;
        continue;
      }
      aux |= 1u << 9;
      const Oct par = parent_of(e);
      // Groups end at 64-leaf boundaries: the kernel sums a group's transposed
      // values within one wavefront.
      const bool same = cur_group >= 0 && same_oct(par, cur_parent) && (e - l0) % 64 != 0;
      const int32_t *ps = &tr.pslot[(size_t)h * npe];
      if (!same) {
        cur_group = (int)gl.size();
        if (cur_group > 255) throw std::runtime_error("tile has too many sibling groups");
        aux |= 1u << 26;
        cur_parent = par;
        gl.push_back(0);
        for (int s = 0; s < npe; s++) to.gploc.push_back(ps[s] >= 0 ? loc_of(ps[s], false) : kNone);
      }
      aux |= ((uint32_t)hang_leaves.size() << 10) | ((uint32_t)cur_group << 18);
      o.laux[e] = aux;
      for (int s = 0; s < npe; s++)
        if (ps[s] >= 0 && sl[s] < 0) {
          loc_of(ps[s], true);
          gl[cur_group] |= 1u << s;
        }
      hang_leaves.push_back(e);
    }
    // Stale parent-buffer readers: leaves that read a parent rank with no node
    // at the parent level.  A reader's cone starts at the last hanging leaf
    // whose own parent values alone fix the slots the reader needs (start), or
    // before the tile (carry: the buffer state at the tile start is computed
    // from host-expanded terms).  Replaying from any earlier start is exact
    // as well, so readers are grouped into blocks, one replay per block.
    std::vector<std::pair<int, int>> rd;  // (start or -1, own hanging index)
    std::vector<uint8_t> rleaf;
    uint32_t entry = 0;  // slot sets as bit masks (nPe <= 27)
    auto absent = [&](size_t leaf) {
      const int32_t *pk = &tr.pslot[(size_t)tr.hang_of[leaf] * npe];
      uint32_t m = 0;
      for (int s = 0; s < npe; s++)
        if (pk[s] < 0) m |= 1u << s;
      return m;
    };
    for (size_t hi = 0; hi < hang_leaves.size(); hi++) {
      const size_t e = hang_leaves[hi];
      const int c = tr.child[e];
      const int32_t *sl = &tr.slot[e * npe];
      uint32_t cur = 0;
      for (int r = 0; r < npe; r++)
        if (sl[r] < 0) cur |= wmask[c * npe + r];
      cur &= absent(e);
      if (!cur) continue;
      int k = (int)hi - 1;
      int start = -1;
      while (true) {
        if (k < 0) {
          entry |= cur;
          break;
        }
        const size_t ek = hang_leaves[k];
        const int ck = tr.child[ek];
        uint32_t nxt = 0;
        for (uint32_t m = cur; m; m &= m - 1) nxt |= wmask[ck * npe + __builtin_ctz(m)];
        nxt &= absent(ek);
        if (!nxt) {
          start = k;
          break;
        }
        cur = nxt;
        k--;
      }
      rd.push_back({start, (int)hi});
      rleaf.push_back((uint8_t)(e - l0));
      to.cone_sum += (double)hi - start;
      to.cone_n += 1;
    }
    // Buffer state at the tile start on the slots carried readers need: a
    // header of nPe + 1 slot offsets, then the terms (slot << 16 | local node)
    // sorted by slot; the weights run parallel (zero under the header).
    size_t carry_terms = 0;
    if (entry && !hang_leaves.empty()) {
      const int before = tr.hang_of[hang_leaves[0]] - 1;  // last hanging leaf before the tile
      std::vector<uint32_t> ck;
      std::vector<double> cwl;
      if (before >= 0)
        for (uint32_t m = entry; m; m &= m - 1) {
          const int s = __builtin_ctz(m);
          for (const auto &e2 : sb.after((size_t)before, s, memo)) {
            if (e2.second == 0.0) continue;
            ck.push_back(((uint32_t)s << 16) | loc_of(e2.first, false));
            cwl.push_back(e2.second);
          }
        }
      if (!ck.empty()) {
        std::vector<uint32_t> hdr_off(npe + 1, 0);
        for (uint32_t k : ck) hdr_off[(k >> 16) + 1]++;
        for (int s = 0; s < npe; s++) hdr_off[s + 1] += hdr_off[s];
        for (int s = 0; s <= npe; s++) {
          to.ckey.push_back(hdr_off[s] + (uint32_t)npe + 1);
          to.cw.push_back(0.0);
        }
        carry_terms = to.ckey.size();
        to.ckey.insert(to.ckey.end(), ck.begin(), ck.end());
        to.cw.insert(to.cw.end(), cwl.begin(), cwl.end());
      }
    }
    // Blocks of consecutive readers (by own index) with the smallest maximum
    // replay length, at most one block per lane group.
    if (!rd.empty()) {
      auto cut = [&](int span, std::vector<std::pair<size_t, size_t>> *out) {
        size_t i = 0, blocks = 0;
        while (i < rd.size()) {
          int mn = rd[i].first;
          size_t j = i + 1;
          while (j < rd.size() && rd[j].second - std::min(mn, rd[j].first) <= span) mn = std::min(mn, rd[j++].first);
          if (out) out->push_back({i, j});
          blocks++;
          i = j;
        }
        return blocks;
      };
      int lo = 0, hi = T + 1;
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if ((int)cut(mid, nullptr) <= n_groups) hi = mid;
        else lo = mid + 1;
      }
      std::vector<std::pair<size_t, size_t>> blocks;
      cut(lo, &blocks);
      int mspan = 0;
      for (const auto &b : blocks) {
        int mn = rd[b.first].first, mx = 0;
        for (size_t r = b.first; r < b.second; r++) {
          mn = std::min(mn, rd[r].first);
          mx = std::max(mx, rd[r].second);
        }
        mspan = std::max(mspan, mx - mn);
        const uint64_t S = mn < 0 ? 0xFFFFu : (uint64_t)mn;
        to.blk.push_back(S | ((uint64_t)mx << 16) | ((uint64_t)b.first << 32) |
                         ((uint64_t)(b.second - b.first) << 48));
      }
      to.span = mspan;
      for (size_t r = 0; r < rd.size(); r++) to.rdr.push_back((uint16_t)(rd[r].second | (rleaf[r] << 8)));
    }
    to.ngrp = (int)gl.size();
    to.nhang = (int)hang_leaves.size();
    // Renumber the tile's local nodes in global node order, so that the x
    // gathers, the y writes and the partial-slot writes of consecutive
    // work-items go to increasing addresses.
    {
      const size_t nn = touched.size();
      std::vector<uint32_t> ord(nn);
      for (size_t k = 0; k < nn; k++) ord[k] = (uint32_t)k;
      std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return touched[a] < touched[b]; });
      std::vector<uint16_t> renum(nn);
      for (size_t k = 0; k < nn; k++) renum[ord[k]] = (uint16_t)k;
      std::vector<int32_t> t2(nn);
      std::vector<uint8_t> it2(nn);
      for (size_t k = 0; k < nn; k++) {
        t2[k] = touched[ord[k]];
        it2[k] = is_target[ord[k]];
      }
      touched.swap(t2);
      is_target.swap(it2);
      for (size_t i = l0 * npe; i < l1 * npe; i++)
        if (o.lloc[i] != kNone) o.lloc[i] = renum[o.lloc[i]];
      for (uint16_t &g : to.gploc)
        if (g != kNone) g = renum[g];
      for (size_t i = carry_terms; i < to.ckey.size(); i++)
        to.ckey[i] = (to.ckey[i] & 0xFFFF0000u) | renum[to.ckey[i] & 0xFFFFu];
    }
    // Hanging-transpose entries, sorted by target node for the segmented sum.
    std::vector<std::pair<uint32_t, uint32_t>> ent;
    for (size_t g = 0; g < gl.size(); g++)  // by group, then parent rank
      for (uint32_t m = gl[g]; m; m &= m - 1) {
        const int s = __builtin_ctz(m);
        const uint16_t j = to.gploc[g * npe + s];
        ent.push_back({(uint32_t)j, (uint32_t)g | ((uint32_t)s << 8) | ((uint32_t)j << 16)});
      }
    std::stable_sort(ent.begin(), ent.end(),
                     [](const std::pair<uint32_t, uint32_t> &a, const std::pair<uint32_t, uint32_t> &b) {
                       return a.first < b.first;
                     });
    for (const auto &e : ent) to.hent.push_back(e.second);
  };
  {
    const unsigned hw = std::thread::hardware_concurrency();
    const int nthreads = (int)std::max(1u, std::min(16u, hw ? hw : 1u));
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex err_mu;
    auto worker = [&]() {
      try {
        LocalMap lm;
        std::unordered_map<uint64_t, Combo> memo;
        while (true) {
          const size_t t0 = next.fetch_add(16);
          if (t0 >= ntiles) break;
          if (memo.size() > (1u << 20)) memo.clear();
          for (size_t t = t0; t < std::min(ntiles, t0 + 16); t++) build_tile(t, outs[t], lm, memo);
        }
      } catch (...) {
        std::lock_guard<std::mutex> g(err_mu);
        if (!err) err = std::current_exception();
        next.store(ntiles);
      }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nthreads; i++) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
    if (err) std::rethrow_exception(err);
  }
  // Concatenation in tile order; shared nodes get their partial slots in tile order.
  double s_rd = 0, m_rd = 0, s_cr = 0, m_cr = 0, s_gr = 0, s_nn = 0, s_cone = 0, n_cone = 0;
  double s_span = 0, m_span = 0, n_rtiles = 0;
  for (size_t t = 0; t < ntiles; t++) {
    TileOut &to = outs[t];
    TileHdr &H = o.hdr[t];
    H.leaf0 = (int32_t)bounds[t];
    H.nleaf = (int32_t)(bounds[t + 1] - bounds[t]);
    H.node0 = (int32_t)o.tnodes.size();
    H.grp0 = (int32_t)(o.gploc.size() / npe);
    H.he0 = (int32_t)o.hent.size();
    H.rd0 = (int32_t)o.rdr.size();
    H.bk0 = (int32_t)o.blk.size();
    H.cr0 = (int32_t)o.ckey.size();
    H.ngrp = to.ngrp;
    H.nrd = (int32_t)to.rdr.size();
    H.nbk = (int32_t)to.blk.size();
    H.ncr = (int32_t)to.ckey.size();
    H.nhe = (int32_t)to.hent.size();
    H.nnode = (int32_t)to.touched.size();
    o.gploc.insert(o.gploc.end(), to.gploc.begin(), to.gploc.end());
    o.hent.insert(o.hent.end(), to.hent.begin(), to.hent.end());
    o.rdr.insert(o.rdr.end(), to.rdr.begin(), to.rdr.end());
    o.blk.insert(o.blk.end(), to.blk.begin(), to.blk.end());
    o.ckey.insert(o.ckey.end(), to.ckey.begin(), to.ckey.end());
    o.cw.insert(o.cw.end(), to.cw.begin(), to.cw.end());
    // Output classification of the tile's nodes.
    for (size_t j = 0; j < to.touched.size(); j++) {
      const int32_t n = to.touched[j];
      o.tnodes.push_back(n);
      if (!to.is_target[j])
        o.tout.push_back(INT32_MIN);
      else if (ntarget[n] == 1)
        o.tout.push_back(n);
      else
        o.tout.push_back(-(pbase[n] + pfill[n]++) - 1);
    }
    o.max_nodes = std::max(o.max_nodes, H.nnode);
    o.max_grp = std::max(o.max_grp, H.ngrp);
    o.max_he = std::max(o.max_he, H.nhe);
    o.max_rd = std::max(o.max_rd, H.nrd);
    o.max_cr = std::max(o.max_cr, H.ncr);
    o.max_hang = std::max(o.max_hang, to.nhang);
    if (H.nrd > 0) n_rtiles += 1;
    s_rd += H.nrd;
    m_rd = std::max<double>(m_rd, H.nrd);
    const double nct = H.ncr ? H.ncr - (npe + 1) : 0;  // carry terms without the header
    s_cr += nct;
    m_cr = std::max<double>(m_cr, nct);
    s_gr += H.ngrp;
    s_nn += H.nnode;
    s_cone += to.cone_sum;
    n_cone += to.cone_n;
    if (to.span >= 0) {
      s_span += to.span;
      m_span = std::max(m_span, to.span);
    }
    TileOut().touched.swap(to.touched);  // release as we go
    std::vector<uint8_t>().swap(to.is_target);
  }
  const double nt = std::max<double>(1, (double)ntiles);
  o.stats = {s_rd / nt, m_rd, s_cr / nt, m_cr, s_gr / nt, s_nn / nt, (double)o.hent.size(),
             n_cone > 0 ? s_cone / n_cone : 0.0, n_rtiles > 0 ? s_span / n_rtiles : 0.0, m_span};
  return o;
}

template <int NPE>
size_t lds_f64(const TiledHost &o, int T) {
  return tile_lds_bytes<double, NPE>(TileDims{o.max_nodes, o.max_grp, o.max_he, o.max_rd, o.max_cr, T, o.max_hang}, 1);
}

void build_tiled(dkt_plan *pl, const MeshParams &mp, const Traversal &tr, int requested) {
  const int npe = pl->npe;
  const int G = npe <= 4 ? 4 : (npe <= 8 ? 8 : (npe <= 16 ? 16 : 32));
  // Leaves per tile: the requested size, else kDefaultTile (the fastest size for
  // the 4D p = 1 workload on MI355X; dendro-kt_amd/autotune.py times the others),
  // halved until the LDS footprint (FP64, one right-hand side) fits.
  int T = requested ? requested : kDefaultTile;
  TiledHost o;
  while (true) {
    o = build_tiled_host(mp, tr, pl->N, T, std::max(1, T / G));
    size_t lds = 0;
    switch (npe) {
      case 4: lds = lds_f64<4>(o, T); break;
      case 8: lds = lds_f64<8>(o, T); break;
      case 9: lds = lds_f64<9>(o, T); break;
      case 16: lds = lds_f64<16>(o, T); break;
      case 27: lds = lds_f64<27>(o, T); break;
      default: throw std::invalid_argument("unsupported nodes per element");
    }
    pl->lds_bytes = lds;
    if (lds <= kLdsLimit) break;
    if (requested || T <= 64) throw std::runtime_error("tile does not fit in LDS");
    T /= 2;
  }
  pl->tile_leaves = T;
  pl->ntiles = o.hdr.size();
  pl->nshared = o.nshared;
  pl->npartial = o.npartial;
  pl->max_nodes = o.max_nodes;
  pl->max_grp = o.max_grp;
  pl->max_he = o.max_he;
  pl->max_rd = o.max_rd;
  pl->max_cr = o.max_cr;
  pl->max_hang = o.max_hang;
  pl->stats = o.stats;
  pl->tiles.upload(o.hdr);
  pl->tile_nodes.upload(o.tnodes);
  pl->tile_out.upload(o.tout);
  pl->leaf_loc.upload(o.lloc);
  pl->leaf_aux.upload(o.laux);
  pl->group_ploc.upload(nonempty(o.gploc, kNone));
  pl->hentries.upload(nonempty(o.hent));
  pl->readers.upload(nonempty(o.rdr));
  pl->blocks.upload(nonempty(o.blk));
  pl->carry_key.upload(nonempty(o.ckey));
  pl->carry_w64.upload(nonempty(o.cw));
  pl->carry_w32.upload(cast<float>(nonempty(o.cw)));
  pl->snode.upload(nonempty(o.snode));
  pl->soff.upload(o.soff);
  pl->map_bytes = (double)(o.hdr.size() * sizeof(TileHdr) + o.tnodes.size() * 8 + o.lloc.size() * 2 +
                           o.laux.size() * 4 + o.gploc.size() * 2 + o.hent.size() * 4 + o.rdr.size() * 2 +
                           o.blk.size() * 8 + o.ckey.size() * 12 + o.snode.size() * 8);
}

// ---------------------------------------------------------------------------
// Tiled conforming operator construction (conf_kernel.h).
// ---------------------------------------------------------------------------
void build_conf_tiled(dkt_plan *pl, const MeshParams &mp, const Traversal &tr, int T) {
  const int npe = pl->npe, dim = mp.dim, n1 = mp.order + 1;
  const size_t L = tr.leaves.size(), N = pl->N;
  // A hanging point: an element node position that is not a mesh node, keyed
  // by its coordinates and the element's level (the constraint row of a
  // position depends on the level it is resolved from).
  using Key = std::array<u32, 5>;
  struct KeyHash {
    size_t operator()(const Key &k) const {
      u64 h = 0x9E3779B97F4A7C15ull;
      for (u32 v : k) h = (h ^ v) * 0xBF58476D1CE4E5B9ull;
      return (size_t)(h ^ (h >> 31));
    }
  };
  auto key_of = [&](size_t e, int r) {
    const Oct &c = tr.leaves[e];
    const u32 step = (1u << (mp.maxd - c.lev)) / (u32)mp.order;
    Key k{0, 0, 0, 0, c.lev};
    for (int d = 0; d < dim; d++, r /= n1) k[d] = c.x[d] + (u32)(r % n1) * step;
    return k;
  };
  // Terms of rank r of element e: cf rows hold each element's ranks in order.
  std::vector<int64_t> rk0(npe), rk1(npe);
  auto rank_ranges = [&](size_t e) {
    std::fill(rk0.begin(), rk0.end(), 0);
    std::fill(rk1.begin(), rk1.end(), 0);
    const int h = tr.hang_of[e];
    if (h < 0) return;
    for (int64_t k = tr.cf_off[h]; k < tr.cf_off[h + 1]; k++) {
      const int r = tr.cf_rank[k];
      if (rk1[r] == 0) rk0[r] = k;
      rk1[r] = k + 1;
    }
  };
  std::unordered_map<Key, int32_t, KeyHash> vmap;
  vmap.reserve(4096);
  // Tile bounds: at most T elements, kMaxConfTerms terms and kMaxConfVirt hanging points.
  std::vector<size_t> bounds{0};
  {
    int64_t nt = 0, nv = 0;
    std::vector<Key> fresh;
    for (size_t e = 0; e < L; e++) {
      rank_ranges(e);
      for (int pass = 0; pass < 2; pass++) {
        int64_t add_t = 0;
        fresh.clear();
        for (int r = 0; r < npe; r++)
          if (tr.slot[e * npe + r] < 0) {
            const Key k = key_of(e, r);
            if (!vmap.count(k) && std::find(fresh.begin(), fresh.end(), k) == fresh.end()) {
              fresh.push_back(k);
              add_t += rk1[r] - rk0[r];
            }
          }
        if (add_t > kMaxConfTerms || (int64_t)fresh.size() > kMaxConfVirt)
          throw std::runtime_error("element with too many constraint terms");
        if (pass == 0 && (e - bounds.back() == (size_t)T || nt + add_t > kMaxConfTerms ||
                          nv + (int64_t)fresh.size() > kMaxConfVirt)) {
          bounds.push_back(e);
          vmap.clear();
          nt = nv = 0;
          continue;
        }
        for (const Key &k : fresh) vmap.emplace(k, 0);
        nt += add_t;
        nv += (int64_t)fresh.size();
        break;
      }
    }
    if (bounds.back() != L) bounds.push_back(L);
  }
  const size_t ntiles = bounds.size() - 1;
  const int rs = conf_row(npe);  // leaf_loc row: npe local points (odd npe: then the level)
  std::vector<uint16_t> lloc(L * (size_t)rs, 0);
  std::vector<uint8_t> llev(L, 0);
  // The plan's weight table: the distinct constraint weights, ordered by their
  // bits; a term stores its weight's index.
  std::vector<uint64_t> wbits;
  {
    const size_t nw_terms = tr.cf_w.size(), chunk = 1 << 16;
    const size_t nchunks = (nw_terms + chunk - 1) / chunk;
    std::vector<std::vector<uint64_t>> part(nchunks);
    std::atomic<size_t> next_chunk{0};
    auto scan = [&]() {
      for (size_t c = next_chunk++; c < nchunks; c = next_chunk++) {
        std::vector<uint64_t> &d = part[c];
        uint64_t last = 0;
        for (size_t k = c * chunk; k < std::min(nw_terms, (c + 1) * chunk); k++) {
          uint64_t b;
          std::memcpy(&b, &tr.cf_w[k], sizeof b);
          if ((!d.empty() && b == last) || std::find(d.begin(), d.end(), b) != d.end()) continue;
          d.push_back(b);
          last = b;
          if (d.size() > (size_t)kMaxConfWeights) break;
        }
      }
    };
    {
      const unsigned hw = std::thread::hardware_concurrency();
      std::vector<std::thread> pool;
      for (unsigned i = 1; i < std::max(1u, std::min(16u, hw ? hw : 1u)); i++) pool.emplace_back(scan);
      scan();
      for (auto &th : pool) th.join();
    }
    for (const auto &d : part) wbits.insert(wbits.end(), d.begin(), d.end());
    std::sort(wbits.begin(), wbits.end());
    wbits.erase(std::unique(wbits.begin(), wbits.end()), wbits.end());
    if ((int)wbits.size() > kMaxConfWeights) throw std::runtime_error("too many distinct constraint weights");
  }
  std::vector<double> wtab(wbits.size());
  for (size_t i = 0; i < wbits.size(); i++) std::memcpy(&wtab[i], &wbits[i], sizeof(double));
  auto weight_index = [&](double w) -> uint16_t {
    uint64_t b;
    std::memcpy(&b, &w, sizeof b);
    return (uint16_t)(std::lower_bound(wbits.begin(), wbits.end(), b) - wbits.begin());
  };
  // Tiles are independent given their bounds: built in parallel into per-tile
  // records, then concatenated in tile order.
  struct ConfTileOut {
    std::vector<int32_t> nodes;  // tile-local node -> mesh node (first touch)
    std::vector<uint16_t> voff;  // per hanging point: first term (tile-relative)
    std::vector<uint32_t> tpk, tspk, sg;
  };
  struct Scratch {
    LocalMap lm;
    std::unordered_map<Key, int32_t, KeyHash> vm;
    std::vector<std::pair<size_t, int32_t>> vref;
    std::vector<int64_t> rk0, rk1;
    std::vector<uint16_t> tnode, tvirt, twi, perm;
  };
  std::vector<ConfTileOut> outs(ntiles);
  auto build_tile = [&](size_t t, ConfTileOut &to, Scratch &sc) {
    sc.lm.clear();
    sc.vm.clear();
    sc.vref.clear();
    sc.tnode.clear();
    sc.tvirt.clear();
    sc.twi.clear();
    sc.rk0.assign(npe, 0);
    sc.rk1.assign(npe, 0);
    int32_t nn = 0;
    auto loc_of = [&](int32_t n) -> uint16_t {
      bool fresh;
      const uint32_t h = sc.lm.find(n, &fresh);
      if (fresh) {
        sc.lm.val[h] = nn++;
        to.nodes.push_back(n);
      }
      return (uint16_t)sc.lm.val[h];
    };
    for (size_t e = bounds[t]; e < bounds[t + 1]; e++) {
      llev[e] = (uint8_t)tr.leaves[e].lev;
      if (rs > npe) lloc[e * rs + npe] = (uint16_t)tr.leaves[e].lev;
      // terms of rank r: cf rows hold each element's ranks in order
      std::fill(sc.rk0.begin(), sc.rk0.end(), 0);
      std::fill(sc.rk1.begin(), sc.rk1.end(), 0);
      if (const int h = tr.hang_of[e]; h >= 0)
        for (int64_t k = tr.cf_off[h]; k < tr.cf_off[h + 1]; k++) {
          const int r = tr.cf_rank[k];
          if (sc.rk1[r] == 0) sc.rk0[r] = k;
          sc.rk1[r] = k + 1;
        }
      for (int r = 0; r < npe; r++) {
        const int32_t n = tr.slot[e * npe + r];
        if (n >= 0) {
          lloc[e * rs + r] = loc_of(n);
          continue;
        }
        auto ins = sc.vm.emplace(key_of(e, r), (int32_t)to.voff.size());
        if (ins.second) {
          to.voff.push_back((uint16_t)sc.tnode.size());
          for (int64_t k = sc.rk0[r]; k < sc.rk1[r]; k++) {
            sc.tnode.push_back(loc_of(tr.cf_node[k]));
            sc.tvirt.push_back((uint16_t)ins.first->second);
            sc.twi.push_back(weight_index(tr.cf_w[k]));
          }
        }
        sc.vref.push_back({e * rs + r, ins.first->second});
      }
    }
    const int nv = (int)to.voff.size(), nt = (int)sc.tnode.size();
    if (nn + nv >= 0xFFFF) throw std::runtime_error("tile has too many points");
    for (const auto &vr : sc.vref) lloc[vr.first] = (uint16_t)(nn + vr.second);
    to.tpk.resize(nt);
    for (int k = 0; k < nt; k++) to.tpk[k] = (uint32_t)sc.tnode[k] | (uint32_t)sc.twi[k] << 16;
    // The terms again, sorted by node (row order within a node), one segment per node.
    sc.perm.resize(nt);
    for (int k = 0; k < nt; k++) sc.perm[k] = (uint16_t)k;
    std::stable_sort(sc.perm.begin(), sc.perm.end(),
                     [&](uint16_t a, uint16_t b) { return sc.tnode[a] < sc.tnode[b]; });
    to.tspk.resize(nt);
    for (int i = 0; i < nt; i++) {
      const int k = sc.perm[i];
      to.tspk[i] = (uint32_t)sc.tvirt[k] | (uint32_t)sc.twi[k] << 16;
      if (i == 0 || sc.tnode[sc.perm[i - 1]] != sc.tnode[k]) to.sg.push_back((uint32_t)sc.tnode[k] | (uint32_t)i << 16);
    }
  };
  {
    const unsigned hw = std::thread::hardware_concurrency();
    const int nthreads = (int)std::max(1u, std::min(16u, hw ? hw : 1u));
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex err_mu;
    auto worker = [&]() {
      try {
        Scratch sc;
        while (true) {
          const size_t t0 = next.fetch_add(16);
          if (t0 >= ntiles) break;
          for (size_t t = t0; t < std::min(ntiles, t0 + 16); t++) build_tile(t, outs[t], sc);
        }
      } catch (...) {
        std::lock_guard<std::mutex> g(err_mu);
        if (!err) err = std::current_exception();
        next.store(ntiles);
      }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nthreads; i++) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
    if (err) std::rethrow_exception(err);
  }
  std::vector<ConfTileHdr> hdr(ntiles);
  std::vector<int32_t> tnodes;
  std::vector<uint16_t> voff;
  std::vector<uint32_t> tpk, tspk, sg;
  int max_points = 0, max_virt = 0, max_terms = 0;
  for (size_t t = 0; t < ntiles; t++) {
    ConfTileOut &to = outs[t];
    ConfTileHdr &H = hdr[t];
    H.leaf0 = (int32_t)bounds[t];
    H.nleaf = (int32_t)(bounds[t + 1] - bounds[t]);
    H.node0 = (int32_t)tnodes.size();
    H.nnode = (int32_t)to.nodes.size();
    H.v0 = (int32_t)voff.size();
    H.nv = (int32_t)to.voff.size();
    H.t0 = (int32_t)tpk.size();
    H.nt = (int32_t)to.tpk.size();
    H.s0 = (int32_t)sg.size();
    H.ns = (int32_t)to.sg.size();
    tnodes.insert(tnodes.end(), to.nodes.begin(), to.nodes.end());
    voff.insert(voff.end(), to.voff.begin(), to.voff.end());
    tpk.insert(tpk.end(), to.tpk.begin(), to.tpk.end());
    tspk.insert(tspk.end(), to.tspk.begin(), to.tspk.end());
    sg.insert(sg.end(), to.sg.begin(), to.sg.end());
    max_points = std::max(max_points, H.nnode + H.nv);
    max_virt = std::max(max_virt, H.nv);
    max_terms = std::max(max_terms, H.nt);
    ConfTileOut().nodes.swap(to.nodes);  // release as we go
  }
  // Halve the tiles while one right-hand side in FP64 does not fit in LDS.
  pl->lds_bytes = conf_lds_bytes<double, 4>(max_points, 1, (int)wtab.size());
  if (pl->lds_bytes > (size_t)kLdsLimit && T > 16) return build_conf_tiled(pl, mp, tr, T / 2);
  // Owned nodes go to y; a node in several tiles gets a partial slot per tile
  // (node-major, tile order), summed by fixup_kernel.
  std::vector<int32_t> ntarget(N, 0);
  for (int32_t n : tnodes) ntarget[n]++;
  std::vector<int32_t> pbase(N, -1), pfill(N, 0), snode, soff{0};
  int64_t np = 0;
  for (size_t n = 0; n < N; n++)
    if (ntarget[n] > 1) {
      pbase[n] = (int32_t)np;
      np += ntarget[n];
      snode.push_back((int32_t)n);
      soff.push_back((int32_t)np);
      if (np >= INT32_MAX) throw std::runtime_error("too many partial sums");
    }
  std::vector<int32_t> tout(tnodes.size());
  for (size_t i = 0; i < tnodes.size(); i++) {
    const int32_t n = tnodes[i];
    tout[i] = ntarget[n] == 1 ? n : -(pbase[n] + pfill[n]++) - 1;
  }
  pl->tile_leaves = T;
  pl->ntiles = ntiles;
  pl->nshared = snode.size();
  pl->npartial = (size_t)np;
  pl->max_nodes = max_points;
  pl->max_virt = max_virt;
  pl->max_terms = max_terms;
  pl->ctiles.upload(hdr);
  pl->tile_nodes.upload(nonempty(tnodes));
  pl->tile_out.upload(nonempty(tout));
  pl->leaf_loc.upload(nonempty(lloc));
  if (rs == npe) pl->leaf_lev.upload(nonempty(llev));
  pl->virt_off.upload(nonempty(voff));
  pl->term_pk.upload(nonempty(tpk));
  pl->term_spk.upload(nonempty(tspk));
  pl->seg.upload(nonempty(sg));
  pl->nw = (int)wtab.size();
  pl->term_w64.upload(nonempty(wtab));
  pl->term_w32.upload(cast<float>(nonempty(wtab)));
  pl->snode.upload(nonempty(snode));
  pl->soff.upload(soff);
  pl->map_bytes = (double)(hdr.size() * sizeof(ConfTileHdr) + tnodes.size() * 8 + lloc.size() * 2 + (rs == npe ? llev.size() : 0) +
                           voff.size() * 2 + tpk.size() * 8 + wtab.size() * 8 + sg.size() * 4 +
                           snode.size() * 8);
}

// ---------------------------------------------------------------------------
// Launch.
// ---------------------------------------------------------------------------
// Scratch of `need` elements for one matvec: the caller's buffer when `ext`
// is given (too small: an error, nothing is allocated), else the plan's own,
// allocated or grown on first use.
template <typename T>
T *matvec_scratch(dkt_plan *pl, size_t need, const ExtScratch *ext) {
  if (ext) {
    if (need * sizeof(T) > ext->bytes)
      throw std::invalid_argument("matvec scratch of " + std::to_string(need * sizeof(T)) +
                                  " bytes exceeds the fixed scratch of " + std::to_string(ext->bytes) + " bytes");
    return (T *)ext->p;
  }
  if (sizeof(T) == 8) {
    if (pl->scratch64.n < need) pl->scratch64.alloc(need);
    return (T *)pl->scratch64.p;
  }
  if (pl->scratch32.n < need) pl->scratch32.alloc(need);
  return (T *)pl->scratch32.p;
}

template <typename T, int DIM, int P>
void run_simple(dkt_plan *pl, int nrhs, const T *x, T *y, int mask, hipStream_t st, const int *cancel,
                const ExtScratch *ext) {
  constexpr int NPE = Pow<P + 1, DIM>::v;
  const bool f64 = sizeof(T) == 8;
  const T *kmat = f64 ? (const T *)pl->kmat64.p : (const T *)pl->kmat32.p;
  const T *hcw = f64 ? (const T *)pl->hc_w64.p : (const T *)pl->hc_w32.p;
  T *buf = matvec_scratch<T>(pl, pl->nbuf * (size_t)nrhs, ext);
  const uint8_t *dir_in = mask == 1 ? pl->dirichlet.p : nullptr;
  const uint8_t *dir_out = mask ? pl->dirichlet.p : nullptr;
  const int blk = 256;
  if (pl->L) {
    hipLaunchKernelGGL((leaf_kernel<T, DIM, P>), dim3((pl->L + blk - 1) / blk), dim3(blk), 0, st, (int)pl->L, nrhs, x,
                       pl->slot.p, pl->level.p, pl->hang_of.p, pl->child.p, pl->pslot.p, pl->hc_off.p,
                       pl->hc_rank.p, pl->hc_node.p, hcw, kmat, pl->per_level ? 1 : 0, dir_in, pl->hp,
                       pl->conforming ? 1 : 0, buf,
                       buf + pl->L * (size_t)NPE * nrhs, cancel);
    HIP_OK(hipGetLastError());
  }
  hipLaunchKernelGGL((node_kernel<T>), dim3((pl->N + blk - 1) / blk), dim3(blk), 0, st, (int)pl->N, nrhs,
                     pl->csr_off.p, pl->csr_src.p, buf, dir_out, y, cancel);
  HIP_OK(hipGetLastError());
}

template <typename T, int DIM, int P>
void run_tiled(dkt_plan *pl, int nrhs, const T *x, T *y, int mask, hipStream_t st, const int *cancel,
               const ExtScratch *ext) {
  constexpr int NPE = Pow<P + 1, DIM>::v;
  const bool f64 = sizeof(T) == 8;
  const T *kmat = P == 2 ? (f64 ? (const T *)pl->kfac64.p : (const T *)pl->kfac32.p)
                         : (f64 ? (const T *)pl->kmat64.p : (const T *)pl->kmat32.p);
  const T *cw = f64 ? (const T *)pl->carry_w64.p : (const T *)pl->carry_w32.p;
  auto lds_for = [&](int nq) {
    return tile_lds_bytes<T, NPE>(
        TileDims{pl->max_nodes, pl->max_grp, pl->max_he, pl->max_rd, pl->max_cr, pl->tile_leaves, pl->max_hang}, nq);
  };
  // Columns are processed in chunks whose LDS footprint stays within budget.
  if (lds_for(1) > kLdsLimit) throw std::invalid_argument("tile does not fit in LDS");
  int chunk = 1;
  while (chunk < nrhs && lds_for(chunk + 1) <= kLdsLimit) chunk++;
  T *part = matvec_scratch<T>(pl, std::max<size_t>(1, pl->npartial * (size_t)chunk), ext);
  TileArgs A;
  A.tiles = pl->tiles.p;
  A.tile_nodes = pl->tile_nodes.p;
  A.tile_out = pl->tile_out.p;
  A.leaf_loc = pl->leaf_loc.p;
  A.leaf_aux = pl->leaf_aux.p;
  A.group_ploc = pl->group_ploc.p;
  A.hentries = pl->hentries.p;
  A.readers = pl->readers.p;
  A.carry_key = pl->carry_key.p;
  A.dir_in = mask == 1 ? pl->dirichlet.p : nullptr;
  A.dir_out = mask ? pl->dirichlet.p : nullptr;
  A.cancel = cancel;
  A.op = pl->op;
  A.hp = pl->hp;
  A.ld = nrhs;
  A.blocks = pl->blocks.p;
  A.dims = TileDims{pl->max_nodes, pl->max_grp, pl->max_he, pl->max_rd, pl->max_cr, pl->tile_leaves, pl->max_hang};
  for (int q0 = 0; q0 < nrhs; q0 += chunk) {
    const int nq = std::min(chunk, nrhs - q0);
    A.nq = nq;
    if (pl->profiling && !ext) HIP_OK(hipEventRecord(pl->next_event(), st));
    if (pl->ntiles) {
      if (nq == 1)
        hipLaunchKernelGGL((tile_kernel<T, DIM, P, true>), dim3(pl->ntiles), dim3(pl->tile_leaves), lds_for(nq), st, A,
                           kmat, cw, x + q0, y + q0, part);
      else
        hipLaunchKernelGGL((tile_kernel<T, DIM, P, false>), dim3(pl->ntiles), dim3(pl->tile_leaves), lds_for(nq), st,
                           A, kmat, cw, x + q0, y + q0, part);
      HIP_OK(hipGetLastError());
    }
    if (pl->profiling && !ext) HIP_OK(hipEventRecord(pl->next_event(), st));
    if (pl->nshared) {
      const int blk = 256;
      hipLaunchKernelGGL((fixup_kernel<T>), dim3((pl->nshared + blk - 1) / blk), dim3(blk), 0, st,
                         (int)pl->nshared, pl->snode.p, pl->soff.p, part, A.dir_out, nq, nrhs, y + q0, cancel);
      HIP_OK(hipGetLastError());
    }
    if (pl->profiling && !ext) HIP_OK(hipEventRecord(pl->next_event(), st));
  }
}

template <typename T, int DIM, int P>
void run_conf(dkt_plan *pl, int nrhs, const T *x, T *y, int mask, hipStream_t st, const int *cancel,
              const ExtScratch *ext) {
  constexpr int NPE = Pow<P + 1, DIM>::v;
  const bool f64 = sizeof(T) == 8;
  const T *kmat = P == 2 && !pl->per_level ? (f64 ? (const T *)pl->kfac64.p : (const T *)pl->kfac32.p)
                                           : (f64 ? (const T *)pl->kmat64.p : (const T *)pl->kmat32.p);
  const T *tw = f64 ? (const T *)pl->term_w64.p : (const T *)pl->term_w32.p;
  auto lds_for = [&](int nq) { return conf_lds_bytes<T, NPE>(pl->max_nodes, nq, pl->nw); };
  if (lds_for(1) > kLdsLimit) throw std::invalid_argument("tile does not fit in LDS");
  int chunk = 1;
  while (chunk < nrhs && lds_for(chunk + 1) <= kLdsLimit) chunk++;
  T *part = matvec_scratch<T>(pl, std::max<size_t>(1, pl->npartial * (size_t)chunk), ext);
  ConfArgs A;
  A.tiles = pl->ctiles.p;
  A.tile_nodes = pl->tile_nodes.p;
  A.tile_out = pl->tile_out.p;
  A.leaf_loc = pl->leaf_loc.p;
  A.leaf_lev = pl->leaf_lev.p;
  A.virt_off = pl->virt_off.p;
  A.term_pk = pl->term_pk.p;
  A.term_spk = pl->term_spk.p;
  A.seg = pl->seg.p;
  A.dir_in = mask == 1 ? pl->dirichlet.p : nullptr;
  A.dir_out = mask ? pl->dirichlet.p : nullptr;
  A.cancel = cancel;
  A.op = pl->op;
  A.hp = pl->hp;
  A.ld = nrhs;
  A.max_points = pl->max_nodes;
  A.tile_leaves = pl->tile_leaves;
  A.nw = pl->nw;
  for (int q0 = 0; q0 < nrhs; q0 += chunk) {
    const int nq = std::min(chunk, nrhs - q0);
    A.nq = nq;
    if (pl->profiling && !ext) HIP_OK(hipEventRecord(pl->next_event(), st));
    if (pl->ntiles) {
      if (pl->per_level)
        hipLaunchKernelGGL((conf_tile_kernel<T, DIM, P, true>), dim3(pl->ntiles), dim3(pl->tile_leaves), lds_for(nq),
                           st, A, kmat, tw, x + q0, y + q0, part);
      else
        hipLaunchKernelGGL((conf_tile_kernel<T, DIM, P, false>), dim3(pl->ntiles), dim3(pl->tile_leaves),
                           lds_for(nq), st, A, kmat, tw, x + q0, y + q0, part);
      HIP_OK(hipGetLastError());
    }
    if (pl->profiling && !ext) HIP_OK(hipEventRecord(pl->next_event(), st));
    if (pl->nshared) {
      hipLaunchKernelGGL((fixup_kernel<T>), dim3((pl->nshared + 255) / 256), dim3(256), 0, st, (int)pl->nshared,
                         pl->snode.p, pl->soff.p, part, A.dir_out, nq, nrhs, y + q0, cancel);
      HIP_OK(hipGetLastError());
    }
    if (pl->profiling && !ext) HIP_OK(hipEventRecord(pl->next_event(), st));
  }
}

template <typename T>
void run_matvec(dkt_plan *pl, int nrhs, const T *x, T *y, int mask, hipStream_t st, const int *cancel,
                const ExtScratch *ext) {
  const bool simple = pl->simple || (pl->per_level && !pl->conforming);
#define DKT_DISPATCH(D, P)                                         \
  if (pl->dim == D && pl->order == P) {                            \
    if (simple)                                                    \
      run_simple<T, D, P>(pl, nrhs, x, y, mask, st, cancel, ext);  \
    else if (pl->conforming)                                       \
      run_conf<T, D, P>(pl, nrhs, x, y, mask, st, cancel, ext);    \
    else                                                           \
      run_tiled<T, D, P>(pl, nrhs, x, y, mask, st, cancel, ext);   \
    return;                                                        \
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

namespace dkt {
void plan_apply(dkt_plan *pl, int dtype, const void *x, void *y, int nrhs, int mask, hipStream_t st,
                const int *cancel, const ExtScratch *ext) {
  if (mask < 0 || mask > 2) throw std::invalid_argument("masked must be 0, 1 or 2");
  if (dtype == 0)
    run_matvec<double>(pl, nrhs, (const double *)x, (double *)y, mask, st, cancel, ext);
  else if (dtype == 1)
    run_matvec<float>(pl, nrhs, (const float *)x, (float *)y, mask, st, cancel, ext);
  else
    throw std::invalid_argument("dtype must be 0 (f64) or 1 (f32)");
}
}  // namespace dkt

extern "C" int dkt_plan_create(const dkt_da *da, int op, const double *op_data, int flags, dkt_plan **out) {
  return guarded([&] {
    if (!da || !out) throw std::invalid_argument("null pointer");
    const MeshParams &mp = da_params(da);
    const NodeSet &ns = da_nodes(da);
    const bool conforming = (flags & DKT_PLAN_CONFORMING) != 0;
    Traversal &tr = conforming ? da_conforming_mut(da) : da_traversal_mut(da);
    auto pl = std::make_unique<dkt_plan>();
    pl->dim = mp.dim;
    pl->order = mp.order;
    pl->maxd = mp.maxd;
    pl->npe = tr.nPe;
    pl->op = op;
    pl->N = ns.nodes.size();
    pl->L = tr.leaves.size();
    pl->H = conforming ? tr.cf_off.size() - 1 : tr.pslot.size() / tr.nPe;
    pl->simple = (flags & DKT_PLAN_SIMPLE) != 0;
    pl->dry = (flags & DKT_PLAN_DRYRUN) != 0;
    pl->conforming = conforming;
    struct DryGuard {
      bool prev;
      explicit DryGuard(bool on) : prev(g_dry_run) { g_dry_run = on; }
      ~DryGuard() { g_dry_run = prev; }
    } dry_guard(pl->dry);
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
    if (mp.order == 2) {
      const std::vector<double> kf = q2_factors();
      pl->kfac64.upload(q2_table<double>(kf, mp.dim));
      pl->kfac32.upload(q2_table<float>(kf, mp.dim));
    }
    std::vector<uint8_t> dir(pl->N, 0);
    for (u32 b : ns.bdy) dir[b] = 1;
    pl->dirichlet.upload(dir);
    const int tshift = (flags >> DKT_PLAN_TILE_SHIFT) & 15;
    if (tshift && (tshift < 6 || tshift > 8)) throw std::invalid_argument("tile size must be 64, 128 or 256 leaves");
    if (pl->simple || (pl->per_level && !pl->conforming))
      build_simple(pl.get(), mp, tr);
    else if (pl->conforming)
      build_conf_tiled(pl.get(), mp, tr, tshift ? 1 << tshift : 256);
    else
      build_tiled(pl.get(), mp, tr, tshift ? 1 << tshift : 0);
    *out = pl.release();
  });
}

extern "C" int dkt_plan_profile(dkt_plan *pl, int enable) {
  return guarded([&] {
    if (!pl) throw std::invalid_argument("null pointer");
    pl->profiling = enable != 0;
    pl->ev_used = 0;
  });
}

extern "C" int dkt_plan_profile_read(dkt_plan *pl, double *tile_ms, double *total_ms, int *ncalls) {
  return guarded([&] {
    if (!pl || !tile_ms || !total_ms || !ncalls) throw std::invalid_argument("null pointer");
    double t = 0, tot = 0;
    const size_t n = pl->ev_used / 3;
    for (size_t i = 0; i < n; i++) {
      float a = 0, b = 0;
      HIP_OK(hipEventSynchronize(pl->ev_pool[3 * i + 2]));
      HIP_OK(hipEventElapsedTime(&a, pl->ev_pool[3 * i], pl->ev_pool[3 * i + 1]));
      HIP_OK(hipEventElapsedTime(&b, pl->ev_pool[3 * i], pl->ev_pool[3 * i + 2]));
      t += a;
      tot += b;
    }
    *tile_ms = t;
    *total_ms = tot;
    *ncalls = (int)n;
    pl->ev_used = 0;
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
    if (pl->dry) throw std::invalid_argument("a dry-run plan cannot run a matvec");
    if (nrhs < 1 || nrhs > 64) throw std::invalid_argument("nrhs must be 1..64");
    if (x == y) throw std::invalid_argument("x and y must not alias");
    plan_apply(pl, dtype, x, y, nrhs, masked, (hipStream_t)stream, nullptr);
  });
}

// ---------------------------------------------------------------------------
// Matvec without allocation: the partial sums go to a fixed device array.
// ---------------------------------------------------------------------------
namespace {
// One copy per device, held until the process ends.  Host code never passes
// `scratch` itself to a kernel (in host code the name is a host-side stand-in,
// not a device address): it uses the address hipGetSymbolAddress returns for
// the current device.
__device__ __align__(256) unsigned char scratch[DKT_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_dev;  // device address of `scratch`, per device (null: not looked up yet)

// Address of the current device's `scratch`, looked up once per device with
// that device current.
void *scratch_address() {
  int dev = -1;
  HIP_OK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((size_t)dev >= g_scratch_dev.size()) g_scratch_dev.resize(dev + 1, nullptr);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_dev[dev] = ptr;
  }
  return g_scratch_dev[dev];
}
}  // namespace

extern "C" int dkt_plan_matvec_noalloc(dkt_plan *pl, int dtype, const void *x, void *y, int nrhs, int masked,
                                       void *stream) {
  return guarded([&] {
    if (!pl || !x || !y) throw std::invalid_argument("null pointer");
    if (pl->dry) throw std::invalid_argument("a dry-run plan cannot run a matvec");
    if (nrhs < 1 || nrhs > 64) throw std::invalid_argument("nrhs must be 1..64");
    if (x == y) throw std::invalid_argument("x and y must not alias");
    const ExtScratch ext{scratch_address(), (size_t)DKT_NOALLOC_SCRATCH_BYTES};
    plan_apply(pl, dtype, x, y, nrhs, masked, (hipStream_t)stream, nullptr, &ext);
  });
}

extern "C" int dkt_plan_set_dirichlet(dkt_plan *pl, const uint32_t *ids, size_t n) {
  return guarded([&] {
    if (!pl || (n && !ids)) throw std::invalid_argument("null pointer");
    std::vector<uint8_t> dir(pl->N, 0);
    for (size_t i = 0; i < n; i++) {
      if (ids[i] >= pl->N) throw std::invalid_argument("Dirichlet node id out of range");
      dir[ids[i]] = 1;
    }
    if (!pl->dry) HIP_OK(hipDeviceSynchronize());  // no matvec in flight reads the old set
    pl->dirichlet.upload(dir);
  });
}

// info: nodes, leaves, hanging leaves, nPe, device map bytes, tiles, shared nodes,
// partial slots, max nodes per tile, the tile statistics, leaves per tile,
// max constraint terms per conforming tile, max hanging leaves per tile, LDS
// bytes per workgroup (FP64, one column).
extern "C" int dkt_plan_info(const dkt_plan *pl, double *info, int n) {
  return guarded([&] {
    if (!pl || !info) throw std::invalid_argument("null pointer");
    std::vector<double> v = {(double)pl->N,       (double)pl->L,       (double)pl->H,
                             (double)pl->npe,     pl->map_bytes,       (double)pl->ntiles,
                             (double)pl->nshared, (double)pl->npartial, (double)pl->max_nodes};
    v.insert(v.end(), pl->stats.begin(), pl->stats.end());
    v.push_back((double)pl->tile_leaves);
    v.push_back((double)pl->max_terms);
    v.push_back((double)pl->max_hang);
    v.push_back((double)pl->lds_bytes);
    for (int i = 0; i < n && i < (int)v.size(); i++) info[i] = v[i];
  });
}
