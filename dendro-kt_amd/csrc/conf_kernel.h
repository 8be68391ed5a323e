// Tiled matvec of the conforming operator sum_e C_e^T K_e C_e (mesh elements,
// hanging-node constraint terms from build_conforming; no counterpart in the
// reference, whose traversal semantics are tile_kernel's).
//
// A tile holds SFC-consecutive elements (one per work-item) and two kinds of
// local points: the mesh nodes its elements touch, and its hanging points
// (element node positions that are not mesh nodes, one per position and
// element level), each with its row of constraint terms over the tile's nodes.
// Every element rank then reads a local point, hanging or not:
//   1. stage x of the tile's nodes; a hanging point's value is its row applied
//      to them (terms in row order);
//   2. per element: u[r] = value at its local point of rank r, v = 2^(-level*hp) K_ref u,
//      v[r] added to that point's accumulator, one phase per rank;
//   3. transposed rows: each node adds sum_k w_k * acc[hanging point of term k]
//      over its terms in an order fixed by the plan (terms sorted by node);
//   4. write-out as tile_kernel (owned nodes to y, shared ones to partial slots
//      summed by fixup_kernel).
// Every sum runs in an order fixed by the plan: bitwise reproducible.
// LDS holds the local point values and the plan's weight table only.  The
// constraint terms are read from global memory where they are used: in row
// order for step 1, and a second copy sorted by node, cut into one segment per
// node, for step 3.  Staging the terms cost 14 bytes of LDS per term (28.6 KB
// of the 49.6 KB of a 256-element 3D p = 2 tile in FP64), which held the
// kernel to three workgroups per CU.
// A term is one 32-bit word: its point (tile-local node, or hanging point for
// the sorted copy) in the low half and the index of its weight in the plan's
// table of distinct weights in the high half.  The weights are products of the
// 1D interpolation weights (9 distinct values in 3D p = 2, 3 in 4D p = 1), so
// the table replaces 8 bytes of FP64 weight per term and copy.
#pragma once
#include "plan_common.h"

namespace dkt {

struct ConfTileHdr {
  int32_t leaf0, nleaf;  // elements
  int32_t node0, nnode;  // mesh nodes (tile_nodes / tile_out); local points [0, nnode)
  int32_t v0, nv;        // hanging points (virt_off), local points [nnode, nnode + nv)
  int32_t t0, nt;        // constraint terms of the hanging points' rows
  int32_t s0, ns;        // segments of the node-sorted terms (seg), one per node with terms
};

struct ConfArgs {
  const ConfTileHdr *tiles;
  const int32_t *tile_nodes, *tile_out;
  const uint16_t *leaf_loc;   // [E][conf_row(nPe)] local point of every rank (odd nPe: then the level)
  const uint8_t *leaf_lev;    // [E] element level (even nPe)
  const uint16_t *virt_off;  // [hanging point] tile-relative first term of its row
  const uint32_t *term_pk;   // per term, row order: tile-local node | weight index << 16
  const uint32_t *term_spk;  // per term, sorted by node: hanging point (0-based in the tile) | weight index << 16
  const uint32_t *seg;       // [segment] tile-local node | tile-relative first sorted term << 16
  const uint8_t *dir_in, *dir_out;
  const int *cancel;
  int op, hp, nq, ld;
  int max_points, tile_leaves;
  int nw;  // entries of the weight table (term_w)
};

// Largest weight table a conforming plan may have (LDS: 2 KB in FP64).
constexpr int kMaxConfWeights = 256;

// 16-bit entries per element in leaf_loc: nPe local points, and for odd nPe
// (p = 2) the element level in the one more entry that makes the row a whole
// number of 32-bit words; even nPe keep the level in leaf_lev.
__host__ __device__ constexpr int conf_row(int npe) { return npe % 2 ? npe + 1 : npe; }

// One tile of one plan (the workgroup may have more work-items than the
// plan's tile_leaves; they idle in the element phases and help in the loops).
// CAP: per-level captured element matrices (DKT_OP_CAPTURED), kref pointing
// to [level][nPe][nPe]; a separate instantiation, since the dense product's
// registers would otherwise count against the closed-form operators too.
template <typename T, int DIM, int P, bool CAP = false>
__device__ __forceinline__ void conf_tile_body(const ConfArgs &A, int tile, const T *__restrict__ kref,
                                              const T *__restrict__ term_w, const T *__restrict__ x,
                                              T *__restrict__ y, T *__restrict__ partial) {
  constexpr int NPE = Pow<P + 1, DIM>::v;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const ConfTileHdr th = A.tiles[tile];
  const int nq = A.nq;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    unsigned char *p = smem_raw + off;
    off += (bytes + 15) & ~(size_t)15;
    return p;
  };
  T *x_loc = reinterpret_cast<T *>(carve((size_t)A.max_points * nq * sizeof(T)));
  // One column: the accumulator reuses x_loc once every element has read it.
  T *y_loc = nq == 1 ? x_loc : reinterpret_cast<T *>(carve((size_t)A.max_points * nq * sizeof(T)));
  T *wt = reinterpret_cast<T *>(carve((size_t)A.nw * sizeof(T)));  // the weight table
  const int tid = threadIdx.x, nthr = blockDim.x;
  // The weight table's first batch, loaded with the stage loads and stored after them.
  const T w_first = tid < A.nw ? term_w[tid] : T(0);

  // 1. Stage the nodes, then evaluate the hanging points' rows.
  const bool active = tid < th.nleaf;
  // The element's local point per rank, two per register; loc_of(r) unpacks
  // one where it is used.  Kept unpacked, the nPe indices would each hold a
  // register from the gather through the rank phases (27 in 3D p = 2).
  // Rows are whole 32-bit words, read as such; for odd nPe the last word's
  // upper half is the level (level_of), so it needs no register of its own.
  constexpr int NW = conf_row(NPE) / 2;
  constexpr bool LEV_IN_ROW = NPE % 2 == 1;
  uint32_t lw[NW];
  int lev = 0;
  if (active) {
    const uint32_t *lp = reinterpret_cast<const uint32_t *>(A.leaf_loc + (size_t)(th.leaf0 + tid) * conf_row(NPE));
#pragma unroll
    for (int w = 0; w < NW; w++) lw[w] = lp[w];
    if constexpr (!LEV_IN_ROW) lev = A.leaf_lev[th.leaf0 + tid];
  } else {
#pragma unroll
    for (int w = 0; w < NW; w++) lw[w] = 0u;
  }
  auto loc_of = [&](int r) -> int { return (int)((lw[r >> 1] >> ((r & 1) * 16)) & 0xFFFFu); };
  auto level_of = [&]() -> int { return LEV_IN_ROW ? (int)(lw[NW - 1] >> 16) : lev; };
  const int npts = th.nnode + th.nv;
  // Staging in batches of kCB per work-item: every load of a batch is issued
  // before any of its values is used, so a batch costs one memory round trip
  // after its node ids, not kCB.  The first column's values are loaded before
  // the Dirichlet flags, and the flags unconditionally (entries past the
  // tile's nodes read node 0's): a flag tested inside a per-entry branch is
  // waited for there, one round trip per entry.
  constexpr int kCB = 4;
  for (int j0 = tid; j0 < th.nnode; j0 += kCB * nthr) {
    int g[kCB];
#pragma unroll
    for (int b = 0; b < kCB; b++) {
      const int j = j0 + b * nthr;
      g[b] = j < th.nnode ? A.tile_nodes[th.node0 + j] : -1;
    }
    T xv[kCB];
#pragma unroll
    for (int b = 0; b < kCB; b++) xv[b] = g[b] >= 0 ? x[(size_t)g[b] * A.ld] : T(0);
    uint32_t d[kCB];
#pragma unroll
    for (int b = 0; b < kCB; b++) d[b] = 0u;
    if (A.dir_in)
#pragma unroll
      for (int b = 0; b < kCB; b++) d[b] = A.dir_in[g[b] >= 0 ? g[b] : 0];
    for (int q = 0; q < nq; q++) {
      if (q > 0)
#pragma unroll
        for (int b = 0; b < kCB; b++) xv[b] = g[b] >= 0 ? x[(size_t)g[b] * A.ld + q] : T(0);
#pragma unroll
      for (int b = 0; b < kCB; b++)
        if (g[b] >= 0) x_loc[(j0 + b * nthr) * nq + q] = d[b] ? T(0) : xv[b];
    }
  }
  if (tid < A.nw) wt[tid] = w_first;
  for (int i = tid + nthr; i < A.nw; i += nthr) wt[i] = term_w[i];
  if (nq > 1)
    for (int j = tid; j < npts * nq; j += nthr) y_loc[j] = T(0);
  __syncthreads();
  // Hanging point values: each row over the staged nodes, terms in row order,
  // read kTB at a time with all loads of a batch issued together.  Entries
  // past the row's end repeat its last term with the value 0: they add an
  // exact +-0 to a sum that starts at +0 and so is never -0 (round to nearest
  // cancels to +0), which leaves its bits those of the row's own terms.  A
  // guarded add instead lets the compiler move the load into the guard and
  // wait for it there.
  constexpr int kTB = 4;
  const uint32_t *tpk = A.term_pk + th.t0;
  for (int v = tid; v < th.nv; v += nthr) {
    const int k0 = A.virt_off[th.v0 + v], k1 = v + 1 < th.nv ? A.virt_off[th.v0 + v + 1] : th.nt;
    for (int q = 0; q < nq; q++) {
      T s = T(0);
      for (int k = k0; k < k1; k += kTB) {
        uint32_t e[kTB];
#pragma unroll
        for (int b = 0; b < kTB; b++) e[b] = tpk[min(k + b, k1 - 1)];
#pragma unroll
        for (int b = 0; b < kTB; b++) {
          const T xv = x_loc[(int)(e[b] & 0xFFFFu) * nq + q];
          s += wt[e[b] >> 16] * (k + b < k1 ? xv : T(0));
        }
      }
      x_loc[(th.nnode + v) * nq + q] = s;
    }
  }
  __syncthreads();
  const T scale = (T)ldexp(1.0, -level_of() * A.hp);

  // 2. Elements.
  for (int q = 0; q < nq; q++) {
    T v[NPE];
    if (active) {
      T u[NPE];
#pragma unroll
      for (int r = 0; r < NPE; r++) u[r] = x_loc[loc_of(r) * nq + q];
      if constexpr (CAP) {
        // Per-level element matrices (row-major, v = K u; the level carries
        // the element size, so no scale): the lanes of a wavefront mostly
        // share a level, so each entry is one broadcast load.
        const T *K = kref + (size_t)level_of() * NPE * NPE;
#pragma unroll
        for (int i = 0; i < NPE; i++) {
          T acc = T(0);
#pragma unroll
          for (int j = 0; j < NPE; j++) acc += K[i * NPE + j] * u[j];
          v[i] = acc;
        }
      } else if constexpr (P == 1) {
        if (A.op == 0)
          q1_apply<T, DIM, 0>(u, v, scale);
        else if (A.op == 1)
          q1_apply<T, DIM, 1>(u, v, scale);
        else
#pragma unroll
          for (int i = 0; i < NPE; i++) v[i] = scale * u[i];
      } else {
        // The element scale 2^(-level hp) as an exponent (one register, not
        // an FP64 pair); scaling by a power of two gives the same bits in any
        // position.
        q2_apply<T, DIM>(u, kref, A.op, T(1));
        if constexpr (LEV_IN_ROW) asm volatile("" : "+v"(lw[NW - 1]));  // the level is read here, not kept
        const int ex = -level_of() * A.hp;
#pragma unroll
        for (int i = 0; i < NPE; i++) v[i] = ldexp(u[i], ex);
      }
    }
    // The packed indices are opaque from here on, so the rank phases unpack
    // them again instead of keeping the gather's unpacked copies alive.
#pragma unroll
    for (int w = 0; w < NW; w++) asm volatile("" : "+v"(lw[w]));
    if (nq == 1) {
      __syncthreads();
      for (int j = tid; j < npts; j += nthr) y_loc[j] = T(0);
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < NPE; r++) {
      if (active) y_loc[loc_of(r) * nq + q] += v[r];
      __syncthreads();
    }
  }

  // 3. Transposed rows: per node, its terms in sorted order (one segment per node).
  const uint32_t *spk = A.term_spk + th.t0;
  for (int g = tid; g < th.ns; g += nthr) {
    const uint32_t sg = A.seg[th.s0 + g];
    const int j = (int)(sg & 0xFFFFu), m0 = (int)(sg >> 16);
    const int m1 = g + 1 < th.ns ? (int)(A.seg[th.s0 + g + 1] >> 16) : th.nt;
    for (int q = 0; q < nq; q++) {
      T acc = T(0);
      for (int m = m0; m < m1; m += kTB) {
        uint32_t e[kTB];
#pragma unroll
        for (int b = 0; b < kTB; b++) e[b] = spk[min(m + b, m1 - 1)];
#pragma unroll
        for (int b = 0; b < kTB; b++) {  // padding adds an exact +-0, as in the rows
          const T yv = y_loc[(th.nnode + (int)(e[b] & 0xFFFFu)) * nq + q];
          acc += wt[e[b] >> 16] * (m + b < m1 ? yv : T(0));
        }
      }
      y_loc[j * nq + q] += acc;
    }
  }
  __syncthreads();

  // 4. Write out, the output slots of a batch of kCB loaded together.
  for (int j0 = tid; j0 < th.nnode; j0 += kCB * nthr) {
    int o[kCB];
#pragma unroll
    for (int b = 0; b < kCB; b++) {
      const int j = j0 + b * nthr;
      o[b] = j < th.nnode ? A.tile_out[th.node0 + j] : INT32_MIN;
    }
    uint32_t d[kCB];
#pragma unroll
    for (int b = 0; b < kCB; b++) d[b] = 0u;
    if (A.dir_out)
#pragma unroll
      for (int b = 0; b < kCB; b++) d[b] = A.dir_out[o[b] >= 0 ? o[b] : 0];
#pragma unroll
    for (int b = 0; b < kCB; b++) {
      const int j = j0 + b * nthr;
      if (o[b] == INT32_MIN) continue;
      if (o[b] >= 0) {
        for (int q = 0; q < nq; q++) y[(size_t)o[b] * A.ld + q] = d[b] ? T(0) : y_loc[j * nq + q];
      } else {
        for (int q = 0; q < nq; q++) partial[(size_t)(-o[b] - 1) * nq + q] = y_loc[j * nq + q];
      }
    }
  }
}

template <typename T, int DIM, int P, bool CAP = false>
__global__ void __launch_bounds__(kTileThreads) conf_tile_kernel(ConfArgs A, const T *__restrict__ kref,
                                                                 const T *__restrict__ term_w,
                                                                 const T *__restrict__ x, T *__restrict__ y,
                                                                 T *__restrict__ partial) {
  if (A.cancel && *A.cancel) return;
  conf_tile_body<T, DIM, P, CAP>(A, xcd_tile(blockIdx.x, gridDim.x), kref, term_w, x, y, partial);
}

// LDS bytes of conf_tile_kernel for nq right-hand sides and a weight table of nw entries.
template <typename T, int NPE>
inline size_t conf_lds_bytes(int max_points, int nq, int nw) {
  auto r16 = [](size_t b) { return (b + 15) & ~(size_t)15; };
  return (nq == 1 ? 1 : 2) * r16((size_t)max_points * nq * sizeof(T)) + r16((size_t)nw * sizeof(T));
}

}  // namespace dkt
