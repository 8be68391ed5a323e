// Tiled matvec kernels (one tile of SFC-consecutive traversal leaves per
// workgroup, one leaf per work-item) and the cross-tile fix-up.
//
// Per tile:
//   1. stage in LDS: the tile's distinct nodes from x (x_loc), the parent-node
//      maps of its sibling groups, its hanging-transpose entries, leaf words;
//   2. each hanging leaf interpolates its present parent nodes in registers
//      (IKD_Parent2Child, FEM/include/refel.h:214-240) into fbuf;
//   3. stale parent-buffer reads: the reference interpolates in a static buffer
//      in place and reads absent parent ranks from what the previous hanging
//      leaf left there (FEM/include/matvec.h:283,411-456).  For every leaf that
//      reads such a value ("reader"), the host gives the start of its
//      dependency cone; the reader's work-item replays the recurrence
//      B_k = F_k + W_ck (Q_k B_{k-1}) forward from there (exact on the slots
//      the reader needs), starting from host-computed carry terms when the
//      cone begins before the tile;
//   4. v = 2^(-level*hp) K_ref u in registers;
//   5. direct outputs are added into the LDS accumulator y_loc in nPe phases,
//      one per rank (two disjoint leaves never share a node at the same rank,
//      so a phase has no write conflicts);
//   6. the transposed interpolation of missing ranks (matvec.h:482-522) is
//      summed per (sibling group, parent rank), then per target node in a
//      segmented sum over entries sorted by target (no conflicts);
//   7. nodes touched by this tile only are written to y; nodes shared with
//      other tiles go to partial slots, summed in tile order by fixup_kernel.
// All reductions run in an order fixed by the plan: bitwise reproducible,
// no atomics.
#pragma once
#include "plan_common.h"

namespace dkt {

template <typename T, int DIM, int P>
__global__ void __launch_bounds__(kTileThreads) tile_kernel(TileArgs A, const T *__restrict__ kref,
                                                            const T *__restrict__ carry_w, const T *__restrict__ x,
                                                            T *__restrict__ y, T *__restrict__ partial) {
  constexpr int NPE = Pow<P + 1, DIM>::v;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const TileHdr th = A.tiles[blockIdx.x];
  const int nq = A.nq;
  const int TL = A.tile_leaves;
  const int ulen = A.max_he > A.max_rd * NPE ? A.max_he : A.max_rd * NPE;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    unsigned char *p = smem_raw + off;
    off += (bytes + 15) & ~(size_t)15;
    return p;
  };
  T *x_loc = reinterpret_cast<T *>(carve((size_t)A.max_nodes * nq * sizeof(T)));
  T *y_loc = reinterpret_cast<T *>(carve((size_t)A.max_nodes * nq * sizeof(T)));
  T *fbuf = reinterpret_cast<T *>(carve((size_t)TL * NPE * sizeof(T)));  // [leaf][NPE]
  T *ubuf = reinterpret_cast<T *>(carve((size_t)ulen * sizeof(T)));       // reader corrections, then group sums
  uint64_t *hent = reinterpret_cast<uint64_t *>(carve((size_t)A.max_he * 8));
  uint32_t *aux_s = reinterpret_cast<uint32_t *>(carve((size_t)TL * 4));
  uint16_t *gploc = reinterpret_cast<uint16_t *>(carve((size_t)A.max_grp * NPE * 2));
  uint16_t *ridx_s = reinterpret_cast<uint16_t *>(carve((size_t)TL * 2));
  uint8_t *hl_s = carve((size_t)TL);
  const int tid = threadIdx.x;

  // 1. Stage.
  const bool active = tid < th.nleaf;
  uint16_t loc[NPE];
  uint32_t aux = 0;
  if (active) {
    const uint16_t *lp = A.leaf_loc + (size_t)(th.leaf0 + tid) * NPE;
#pragma unroll
    for (int r = 0; r < NPE; r++) loc[r] = lp[r];
    aux = A.leaf_aux[th.leaf0 + tid];
  } else {
#pragma unroll
    for (int r = 0; r < NPE; r++) loc[r] = kNone;
  }
  if (tid < TL) {
    aux_s[tid] = aux;
    ridx_s[tid] = kNone;
  }
  for (int j = tid; j < th.nnode; j += kTileThreads) {
    const int g = A.tile_nodes[th.node0 + j];
    const bool z = A.dirichlet && A.dirichlet[g];
    for (int q = 0; q < nq; q++) {
      x_loc[j * nq + q] = z ? T(0) : x[(size_t)g * A.ld + q];
      y_loc[j * nq + q] = T(0);
    }
  }
  for (int k = tid; k < th.ngrp * NPE; k += kTileThreads) gploc[k] = A.group_ploc[(size_t)th.grp0 * NPE + k];
  for (int k = tid; k < th.nhe; k += kTileThreads) hent[k] = A.hentries[th.he0 + k];
  for (int k = tid; k < th.nhl; k += kTileThreads) hl_s[k] = A.hlist[th.hl0 + k];
  __syncthreads();

  const bool hang = active && aux_hang(aux);
  const int child = aux_child(aux);
  const T scale = (T)ldexp(1.0, -aux_level(aux) * A.hp);
  const uint16_t *ploc = gploc + (size_t)aux_group(aux) * NPE;

  for (int q = 0; q < nq; q++) {
    // 2. Interpolation of the present parent nodes.
    if (hang) {
      T f[NPE];
#pragma unroll
      for (int s = 0; s < NPE; s++) f[s] = ploc[s] != kNone ? x_loc[ploc[s] * nq + q] : T(0);
      interp<T, DIM, P, false>(f, child);
#pragma unroll
      for (int r = 0; r < NPE; r++) fbuf[tid * NPE + r] = f[r];
    }
    __syncthreads();
    // 3. Stale parent-buffer reads, one work-item per reader.
    for (int i = tid; i < th.nrd; i += kTileThreads) {
      const uint64_t rec = A.readers[th.rd0 + i];
      const int leaf = (int)(rec & 255u), start = (int)((rec >> 8) & 255u), own = (int)((rec >> 16) & 255u);
      const bool carry = (rec >> 24) & 1u;
      T B[NPE];
      int k;
      if (carry) {
#pragma unroll
        for (int s = 0; s < NPE; s++) B[s] = T(0);
        const int c0 = th.cr0 + (int)((rec >> 32) & 0xFFFFu), c1 = c0 + (int)(rec >> 48);
        for (int c = c0; c < c1; c++) {
          const uint32_t key = A.carry_key[c];
          const int s = (key >> 16) & 255, j = key & 0xFFFF;
          const T add = carry_w[c] * x_loc[j * nq + q];
#pragma unroll
          for (int ss = 0; ss < NPE; ss++)
            if (ss == s) B[ss] += add;
        }
        k = 0;
      } else {
        const int l = hl_s[start];
#pragma unroll
        for (int s = 0; s < NPE; s++) B[s] = fbuf[l * NPE + s];
        k = start + 1;
      }
      for (;; k++) {
        const int l = k < own ? hl_s[k] : leaf;
        const uint32_t la = aux_s[l];
        const uint16_t *pl = gploc + (size_t)aux_group(la) * NPE;
        T b[NPE];
#pragma unroll
        for (int s = 0; s < NPE; s++) b[s] = pl[s] == kNone ? B[s] : T(0);
        interp<T, DIM, P, false>(b, aux_child(la));
        if (k >= own) {  // the reader itself: keep only the stale part
#pragma unroll
          for (int s = 0; s < NPE; s++) ubuf[i * NPE + s] = b[s];
          ridx_s[leaf] = (uint16_t)i;
          break;
        }
#pragma unroll
        for (int s = 0; s < NPE; s++) B[s] = fbuf[l * NPE + s] + b[s];
      }
    }
    __syncthreads();
    // 4. Element vector and elemental operator.
    T v[NPE];
    if (active) {
      T u[NPE];
      const int ri = ridx_s[tid];
#pragma unroll
      for (int r = 0; r < NPE; r++) {
        T hv = T(0);
        if (hang) hv = fbuf[tid * NPE + r] + (ri != kNone ? ubuf[ri * NPE + r] : T(0));
        u[r] = loc[r] != kNone ? x_loc[loc[r] * nq + q] : hv;
      }
#pragma unroll
      for (int i = 0; i < NPE; i++) {
        T acc = T(0);
#pragma unroll
        for (int j = 0; j < NPE; j++) acc += kref[i * NPE + j] * u[j];
        v[i] = scale * acc;
      }
    }
    // 5. Direct outputs, one phase per rank.
#pragma unroll
    for (int r = 0; r < NPE; r++) {
      __syncthreads();
      if (loc[r] != kNone) y_loc[loc[r] * nq + q] += v[r];
    }
    // 6. Transposed interpolation of missing ranks.
    if (th.nhe > 0) {
      if (active) {
        T w[NPE];
#pragma unroll
        for (int r = 0; r < NPE; r++) w[r] = (hang && loc[r] == kNone) ? v[r] : T(0);
        interp<T, DIM, P, true>(w, child);
#pragma unroll
        for (int s = 0; s < NPE; s++) fbuf[tid * NPE + s] = (hang && loc[s] == kNone) ? w[s] : T(0);
      }
      __syncthreads();
      for (int k = tid; k < th.nhe; k += kTileThreads) {
        const uint64_t e = hent[k];
        const int f0 = (int)(e & 255u), cnt = (int)((e >> 8) & 255u), s = (int)((e >> 16) & 255u);
        T acc = T(0);
        for (int l = 0; l < cnt; l++) acc += fbuf[(f0 + l) * NPE + s];
        ubuf[k] = acc;
      }
      __syncthreads();
      for (int k = tid; k < th.nhe; k += kTileThreads) {
        const int j = (int)((hent[k] >> 24) & 0xFFFFu);
        if (k > 0 && (int)((hent[k - 1] >> 24) & 0xFFFFu) == j) continue;  // not a segment head
        T acc = T(0);
        for (int m = k; m < th.nhe && (int)((hent[m] >> 24) & 0xFFFFu) == j; m++) acc += ubuf[m];
        y_loc[j * nq + q] += acc;
      }
    }
    __syncthreads();
  }
  // 7. Write out: owned nodes to y, shared nodes to their partial slot.
  for (int j = tid; j < th.nnode; j += kTileThreads) {
    const int o = A.tile_out[th.node0 + j];
    if (o == INT32_MIN) continue;  // gathered only
    if (o >= 0) {
      const bool z = A.dirichlet && A.dirichlet[o];
      for (int q = 0; q < nq; q++) y[(size_t)o * A.ld + q] = z ? T(0) : y_loc[j * nq + q];
    } else {
      const size_t slot = (size_t)(-o - 1);
      for (int q = 0; q < nq; q++) partial[slot * nq + q] = y_loc[j * nq + q];
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) fixup_kernel(int ns, const int32_t *__restrict__ snode,
                                                    const int32_t *__restrict__ soff, const T *__restrict__ partial,
                                                    const uint8_t *__restrict__ dirichlet, int nq, int ld,
                                                    T *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const int n = snode[i];
  const int a = soff[i], b = soff[i + 1];
  const bool z = dirichlet && dirichlet[n];
  for (int q = 0; q < nq; q++) {
    T acc = T(0);
    for (int k = a; k < b; k++) acc += partial[(size_t)k * nq + q];
    y[(size_t)n * ld + q] = z ? T(0) : acc;
  }
}

// LDS bytes of one tile for nq right-hand sides (layout of tile_kernel).
template <typename T, int NPE>
inline size_t tile_lds_bytes(int max_nodes, int max_grp, int max_he, int max_rd, int tile_leaves, int nq) {
  const size_t ulen = (size_t)(max_he > max_rd * NPE ? max_he : max_rd * NPE);
  auto r16 = [](size_t b) { return (b + 15) & ~(size_t)15; };
  return 2 * r16((size_t)max_nodes * nq * sizeof(T)) + r16((size_t)tile_leaves * NPE * sizeof(T)) +
         r16(ulen * sizeof(T)) + r16((size_t)max_he * 8) + r16((size_t)tile_leaves * 4) +
         r16((size_t)max_grp * NPE * 2) + r16((size_t)tile_leaves * 2) + r16((size_t)tile_leaves);
}

}  // namespace dkt
