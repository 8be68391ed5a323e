// Tiled matvec kernels (one tile of SFC-consecutive traversal leaves per
// workgroup, one leaf per work-item) and the cross-tile fix-up.
//
// Per tile:
//   1. stage in LDS: the tile's distinct nodes from x (x_loc), the parent-node
//      maps of its sibling groups, its hanging-transpose entries, leaf words;
//   2. the parent-buffer state at the tile start, from host-expanded terms;
//   3. stale parent-buffer reads: the reference interpolates in a static buffer
//      in place and reads absent parent ranks from what the previous hanging
//      leaf left there (FEM/include/matvec.h:283,411-456).  For every leaf that
//      reads such a value ("reader"), the host gives the start of its
//      dependency cone; a group of lanes (one per buffer slot) replays the
//      recurrence B_k = W_ck (parent nodes where present, else B_(k-1))
//      forward from there, exact on the slots the reader needs;
//   4. each leaf forms its element vector (hanging leaves: parent nodes
//      interpolated in registers, IKD_Parent2Child, FEM/include/refel.h:214-240,
//      or the replayed buffer for readers) and applies
//      v = 2^(-level*hp) K_ref u (p = 1: in the Walsh-Hadamard basis that
//      diagonalises K_ref; p = 2: dense, K_ref staged in LDS);
//   5. direct outputs are added into the LDS accumulator y_loc in nPe phases,
//      one per rank (two disjoint leaves never share a node at the same rank,
//      so a phase has no write conflicts);
//   6. the transposed interpolation of missing ranks (matvec.h:482-522) is
//      summed over each sibling group with cross-lane moves inside the
//      wavefront (DPP row shifts, one permute across rows), then per target
//      node in a segmented sum over the (group, parent rank) entries sorted
//      by target (no conflicts);
//   7. nodes touched by this tile only are written to y; nodes shared with
//      other tiles go to partial slots, summed in tile order by fixup_kernel.
// All reductions run in an order fixed by the plan: bitwise reproducible,
// no atomics.
#pragma once
#include "plan_common.h"

namespace dkt {

// Nodes staged or written per work-item with all their loads in flight.
constexpr int kNB = 4;

// One leaf's rank-to-local-node row, with 16- or 8-byte loads where the row
// size allows (rows are nPe uint16 from a 256-byte aligned base).
template <int NPE>
__device__ __forceinline__ void load_row(const uint16_t *p, uint16_t *loc) {
  if constexpr (NPE % 8 == 0) {
#pragma unroll
    for (int c = 0; c < NPE / 8; c++) {
      const uint4 w = reinterpret_cast<const uint4 *>(p)[c];
      const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        loc[8 * c + 2 * k] = (uint16_t)(ws[k] & 0xFFFFu);
        loc[8 * c + 2 * k + 1] = (uint16_t)(ws[k] >> 16);
      }
    }
  } else if constexpr (NPE == 4) {
    const uint2 w = *reinterpret_cast<const uint2 *>(p);
    loc[0] = (uint16_t)(w.x & 0xFFFFu);
    loc[1] = (uint16_t)(w.x >> 16);
    loc[2] = (uint16_t)(w.y & 0xFFFFu);
    loc[3] = (uint16_t)(w.y >> 16);
  } else {
#pragma unroll
    for (int r = 0; r < NPE; r++) loc[r] = p[r];
  }
}

// Lanes that replay one stale-buffer reader together: one lane per buffer slot.
template <int NPE>
struct Group {
  static constexpr int v = NPE <= 4 ? 4 : (NPE <= 8 ? 8 : (NPE <= 16 ? 16 : 32));
};

// Elements of the tile kernel's y_loc: the accumulator, and with one column
// (the single-column kernel) the groups' parent values before it.
template <int NPE>
__host__ __device__ inline size_t tile_yloc_elems(const TileDims &D, int nq) {
  const size_t acc = (size_t)D.max_nodes * nq, par = (size_t)D.max_grp * NPE;
  return nq == 1 && par > acc ? par : acc;
}

// Bytes of the tile kernel's shared LDS region: reader values or per-group
// transposed sums, whichever is larger.
template <typename T, int NPE>
__host__ __device__ inline size_t tile_region_bytes(const TileDims &D) {
  const size_t readers = (size_t)D.max_rd * NPE * sizeof(T);
  const size_t groups = (size_t)D.max_grp * NPE * sizeof(T);
  const size_t m = readers > groups ? readers : groups;
  return m > 16 ? m : 16;
}

// Cross-lane moves inside a row of 16 lanes (DPP, a VALU operand modifier;
// no LDS round trip).  Controls: quad_perm [1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E,
// [3,2,1,0] = 0x1B, row_half_mirror = 0x141, row_ror:8 = 0x128.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long u = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xFFFFFFFFll), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
// Value of lane (lane ^ (1 << D)) within the same row of 16.
template <int D, typename T>
__device__ __forceinline__ T xor_lane(T a) {
  if constexpr (D == 0) return dpp_mov<0xB1>(a);
  else if constexpr (D == 1) return dpp_mov<0x4E>(a);
  else if constexpr (D == 2) return dpp_mov<0x1B>(dpp_mov<0x141>(a));  // (lane ^ 7) ^ 3
  else return dpp_mov<0x128>(a);
}

// One step of the sibling-group sum: w[s] += (value of lane wl + D in the same
// row of 16) when that lane is still in the run.  Products with the 0/1 factor
// are exact, so the fma adds an exact zero or rounds once like the plain sum.
// A step that no lane of the wavefront needs (every run within a row is at
// most D lanes long) is skipped: it would add exact zeros only.
template <typename T, int CTRL, int D, int N>
__device__ __forceinline__ void group_step(T (&w)[N], int wl, int run_end) {
  const bool in = wl + D < run_end && (wl & 15) + D < 16;
  if (!__ballot(in)) return;
  const T f = in ? T(1) : T(0);
#pragma unroll
  for (int s = 0; s < N; s++) w[s] = fma(dpp_mov<CTRL>(w[s]), f, w[s]);
}

// Forward parent-to-child interpolation with the buffer spread over a lane
// group (lane s holds slot s): the same tensor-product map as interp(), one
// axis at a time, with the other entries of an axis line fetched from their
// lanes.  p = 1: child node i of axis d is the parent node when i equals the
// child bit, else the midpoint 0.5 (a0 + a1).
template <typename T, int DIM, int P>
__device__ __forceinline__ T interp_lanes(T a, int lane, int c) {
  constexpr int G = Group<Pow<P + 1, DIM>::v>::v;
  if constexpr (P == 1) {
    const bool k0 = (lane & 1) == (c & 1);
    T b = xor_lane<0>(a);
    a = k0 ? a : T(0.5) * (a + b);
    if constexpr (DIM > 1) {
      const bool k1 = ((lane >> 1) & 1) == ((c >> 1) & 1);
      b = xor_lane<1>(a);
      a = k1 ? a : T(0.5) * (a + b);
    }
    if constexpr (DIM > 2) {
      const bool k2 = ((lane >> 2) & 1) == ((c >> 2) & 1);
      b = xor_lane<2>(a);
      a = k2 ? a : T(0.5) * (a + b);
    }
    if constexpr (DIM > 3) {
      const bool k3 = ((lane >> 3) & 1) == ((c >> 3) & 1);
      b = xor_lane<3>(a);
      a = k3 ? a : T(0.5) * (a + b);
    }
  } else {
    constexpr int n1 = P + 1;
    int st = 1;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      const int id = (lane / st) % n1;
      const int base = lane - id * st;
      const int cb = (c >> d) & 1;
      T acc = T(0);
#pragma unroll
      for (int k = 0; k < n1; k++) acc += (T)Ip<P>::w(cb, k, id) * __shfl(a, base + k * st, G);
      a = acc;
      st *= n1;
    }
  }
  return a;
}

// One tile of one plan (the workgroup may have more work-items than the
// plan's tile_leaves; they idle in the leaf phases and help in the loops).
// ONE: a single right-hand side per launch (nq = 1 at compile time, which
// keeps the leaf state out of a runtime column loop and halves the VGPRs).
template <typename T, int DIM, int P, bool ONE>
__device__ __forceinline__ void tile_body(const TileArgs &A, int tile, const T *__restrict__ kref,
                                         const T *__restrict__ carry_w, const T *__restrict__ x, T *__restrict__ y,
                                         T *__restrict__ partial) {
  constexpr int NPE = Pow<P + 1, DIM>::v;
  constexpr int G = Group<NPE>::v;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const TileHdr th = A.tiles[tile];
  const TileDims D = A.dims;
  const int nq = ONE ? 1 : A.nq;
  const int TL = D.tile_leaves;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    unsigned char *p = smem_raw + off;
    off += (bytes + 15) & ~(size_t)15;
    return p;
  };
  T *x_loc = reinterpret_cast<T *>(carve((size_t)D.max_nodes * nq * sizeof(T)));
  // ONE: y_loc first holds the parent values of the sibling groups [group][slot]
  // (replay and element phases), then the accumulator.
  T *y_loc = reinterpret_cast<T *>(carve(tile_yloc_elems<NPE>(D, nq) * sizeof(T)));
  T *pval = y_loc;
  // Shared region: reader values [reader][slot] (replay to element phase),
  // then group sums [group][slot] (transposed phase).
  T *ubuf = reinterpret_cast<T *>(carve(tile_region_bytes<T, NPE>(D)));
  T *cst = reinterpret_cast<T *>(carve((size_t)NPE * nq * sizeof(T)));  // parent buffer at the tile start
  uint32_t *hent = reinterpret_cast<uint32_t *>(carve((size_t)D.max_he * 4));
  uint32_t *cw_s = reinterpret_cast<uint32_t *>(carve((size_t)(D.max_hang + 1) * 4));  // chain word per hanging index
  uint16_t *gploc = reinterpret_cast<uint16_t *>(carve((size_t)D.max_grp * NPE * 2));
  uint16_t *ridx_s = reinterpret_cast<uint16_t *>(carve((size_t)TL * 2));
  uint16_t *rdl_s = reinterpret_cast<uint16_t *>(carve((size_t)D.max_rd * 2));
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;

  // 1. Stage.
  const bool active = tid < th.nleaf;
  uint16_t loc[NPE];
  uint32_t aux = 0;
  if (active) {
    load_row<NPE>(A.leaf_loc + (size_t)(th.leaf0 + tid) * NPE, loc);
    aux = A.leaf_aux[th.leaf0 + tid];
  } else {
#pragma unroll
    for (int r = 0; r < NPE; r++) loc[r] = kNone;
  }
  if (tid < TL) ridx_s[tid] = kNone;
  // Output slots of the first write-out batch, loaded with the stage loads so
  // that the write-out does not wait for them.
  int o_first[kNB];
#pragma unroll
  for (int b = 0; b < kNB; b++) {
    const int j = tid + b * nthr;
    o_first[b] = j < th.nnode ? A.tile_out[th.node0 + j] : INT32_MIN;  // INT32_MIN: gathered only
  }
  // Replay block of this lane group's first round, likewise.
  const uint64_t blk_first = (tid / G) < th.nbk ? A.blocks[th.bk0 + tid / G] : 0;
  // First two batches of the group parent maps and transpose entries, and the
  // readers' first batch: loads issued here, stored after the node gathers.
  const int ngp = th.ngrp * NPE;
  uint16_t gp_first[2];
  uint32_t he_first[2];
#pragma unroll
  for (int b = 0; b < 2; b++) {
    const int k = tid + b * nthr;
    gp_first[b] = k < ngp ? A.group_ploc[(size_t)th.grp0 * NPE + k] : kNone;
    he_first[b] = k < th.nhe ? A.hentries[th.he0 + k] : 0u;
  }
  const uint16_t rd_first = tid < th.nrd ? A.readers[th.rd0 + tid] : kNone;
  // Node gathers in batches of kNB per work-item, all loads issued before use:
  // the first column's values, then the Dirichlet flags, unconditionally
  // (entries past the tile's nodes read node 0's; a flag tested inside a
  // per-entry branch is waited for there, one round trip per entry).
  for (int j0 = tid; j0 < th.nnode; j0 += kNB * nthr) {
    int g[kNB];
#pragma unroll
    for (int b = 0; b < kNB; b++) {
      const int j = j0 + b * nthr;
      g[b] = j < th.nnode ? A.tile_nodes[th.node0 + j] : -1;
    }
    T xv[kNB];
#pragma unroll
    for (int b = 0; b < kNB; b++) xv[b] = g[b] >= 0 ? x[(size_t)g[b] * A.ld] : T(0);
    uint32_t d[kNB];
#pragma unroll
    for (int b = 0; b < kNB; b++) d[b] = 0u;
    if (A.dir_in)
#pragma unroll
      for (int b = 0; b < kNB; b++) d[b] = A.dir_in[g[b] >= 0 ? g[b] : 0];
    for (int q = 0; q < nq; q++) {
      if (q > 0)
#pragma unroll
        for (int b = 0; b < kNB; b++) xv[b] = g[b] >= 0 ? x[(size_t)g[b] * A.ld + q] : T(0);
#pragma unroll
      for (int b = 0; b < kNB; b++) {
        const int j = j0 + b * nthr;
        if (g[b] >= 0) {
          x_loc[j * nq + q] = d[b] ? T(0) : xv[b];
          if constexpr (!ONE) y_loc[j * nq + q] = T(0);
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < 2; b++) {
    const int k = tid + b * nthr;
    if (k < ngp) gploc[k] = gp_first[b];
    if (k < th.nhe) hent[k] = he_first[b];
  }
  if (tid < th.nrd) rdl_s[tid] = rd_first;
  for (int k = tid + 2 * nthr; k < ngp; k += nthr) gploc[k] = A.group_ploc[(size_t)th.grp0 * NPE + k];
  for (int k = tid + 2 * nthr; k < th.nhe; k += nthr) hent[k] = A.hentries[th.he0 + k];
  for (int k = tid + nthr; k < th.nrd; k += nthr) rdl_s[k] = A.readers[th.rd0 + k];
  const bool hang = active && aux_hang(aux);
  const int child = aux_child(aux);
  const int hidx = aux_hidx(aux);
  const int grp_id = aux_group(aux);
  const T scale = (T)ldexp(1.0, -aux_level(aux) * A.hp);
  if (hang) cw_s[hidx] = (uint32_t)child | ((uint32_t)grp_id << 8);
  __syncthreads();

  // 2. Parent-buffer state at the tile start, every column: slot s sums its
  //    terms (a header of slot offsets, then the terms by slot) in plan order,
  //    read kNB at a time with all loads of a batch issued together.  Entries
  //    past the slot's last term repeat it with the value 0: an exact +-0
  //    added to a sum that starts at +0 and so is never -0, which leaves its
  //    bits those of the slot's own terms.
  if (tid < NPE) {
    const uint32_t *ck = A.carry_key + th.cr0;
    const T *cwp = carry_w + th.cr0;
    const int lo = th.ncr > 0 ? (int)ck[tid] : 0, hi = th.ncr > 0 ? (int)ck[tid + 1] : 0;
    for (int q = 0; q < nq; q++) {
      T acc = T(0);
      for (int c = lo; c < hi; c += kNB) {
        uint32_t kk[kNB];
        T cwv[kNB];
#pragma unroll
        for (int b = 0; b < kNB; b++) {
          const int cc = min(c + b, hi - 1);
          kk[b] = ck[cc];
          cwv[b] = cwp[cc];
        }
#pragma unroll
        for (int b = 0; b < kNB; b++) {
          const T xv = x_loc[(kk[b] & 0xFFFFu) * nq + q];
          acc += cwv[b] * (c + b < hi ? xv : T(0));
        }
      }
      cst[tid * nq + q] = acc;
    }
  }
  if constexpr (ONE)
    for (int k = tid; k < th.ngrp * NPE; k += nthr) pval[k] = gploc[k] != kNone ? x_loc[gploc[k]] : T(0);
  __syncthreads();
  const uint16_t *ploc = gploc + (size_t)grp_id * NPE;
  const int lane = tid % G, grp = tid / G, ngroups = nthr / G;
  const int sl = lane < NPE ? lane : 0;
  const int wl = tid & 63;  // lane in the wavefront
  // Sibling-group runs in this wavefront: a run ends at the next group head or
  // non-hanging work-item.
  const uint64_t starts = __ballot(!hang || aux_ghead(aux));
  const uint64_t later = wl == 63 ? 0 : (starts >> (wl + 1));
  const int run_end = later ? wl + 1 + __builtin_ctzll(later) : 64;

  for (int q = 0; q < nq; q++) {
    // 3. Stale parent-buffer reads: a lane group per replay block runs the
    //    recurrence over the block's hanging leaves, B_k = W_ck (parent nodes
    //    where present, else B_(k-1)), and stores B_k of each reader it passes.
    for (int b0 = 0; b0 < th.nbk; b0 += ngroups) {
      const int bi = b0 + grp;
      if (bi >= th.nbk) continue;
      const uint64_t blk = b0 == 0 ? blk_first : A.blocks[th.bk0 + bi];
      const int S = (int)(blk & 0xFFFFu), E = (int)((blk >> 16) & 0xFFFFu);
      int p = (int)((blk >> 32) & 0xFFFFu);
      const int pend = p + (int)(blk >> 48);
      int nxt = rdl_s[p] & 255;
      // Parent value of this lane's slot for hanging index k (kNone: absent).
      auto parent = [&](uint32_t cw, T &xp) {
        const size_t gi = (size_t)(cw >> 8) * NPE + sl;
        const uint16_t pl = gploc[gi];
        if constexpr (ONE)
          xp = lane < NPE ? pval[gi] : T(0);
        else
          xp = (lane < NPE && pl != kNone) ? x_loc[pl * nq + q] : T(0);
        return lane < NPE && pl == kNone;
      };
      T B;
      int k;
      if (S == 0xFFFF) {
        B = cst[sl * nq + q];
        k = 0;
      } else {
        const uint32_t cw0 = cw_s[S];
        T xp;
        parent(cw0, xp);
        B = interp_lanes<T, DIM, P>(xp, lane, (int)(cw0 & 255u));
        k = S + 1;
      }
      // Chain words and parent values do not depend on B: loaded one step ahead.
      uint32_t cw = cw_s[k];
      T xp;
      bool ab = parent(cw, xp);
      for (;; k++) {
        const bool last = k >= E;
        const uint32_t cwn = last ? cw : cw_s[k + 1];
        T xpn;
        const bool abn = parent(cwn, xpn);
        B = interp_lanes<T, DIM, P>(ab ? B : xp, lane, (int)(cw & 255u));
        if (k == nxt) {
          if (lane < NPE) ubuf[p * NPE + lane] = B;
          if (lane == 0) ridx_s[rdl_s[p] >> 8] = (uint16_t)p;
          p++;
          nxt = p < pend ? (rdl_s[p] & 255) : -1;
        }
        if (last) break;
        cw = cwn;
        xp = xpn;
        ab = abn;
      }
    }
    __syncthreads();
    // 4. Element vector and elemental operator: a hanging leaf's missing ranks
    //    interpolate its parent nodes (readers: the replayed buffer B_k).
    T v[NPE];
    if (active) {
      // Missing ranks first (zero for a leaf without hanging nodes), then the
      // present ranks replace them: one LDS read per rank, no branches.
      T u[NPE];
      if (hang) {
        if constexpr (ONE) {
#pragma unroll
          for (int s = 0; s < NPE; s++) u[s] = pval[grp_id * NPE + s];
        } else {
#pragma unroll
          for (int s = 0; s < NPE; s++) u[s] = ploc[s] != kNone ? x_loc[ploc[s] * nq + q] : T(0);
        }
        interp<T, DIM, P, false>(u, child);
        const int ri = ridx_s[tid];
        if (ri != kNone)
#pragma unroll
          for (int r = 0; r < NPE; r++) u[r] = ubuf[ri * NPE + r];
      } else {
#pragma unroll
        for (int r = 0; r < NPE; r++) u[r] = T(0);
      }
#pragma unroll
      for (int r = 0; r < NPE; r++) {
        const bool present = loc[r] != kNone;
        const T xv = x_loc[(present ? loc[r] : 0) * nq + q];
        u[r] = present ? xv : u[r];
      }
      if constexpr (P == 1) {
        if (A.op == 0)
          q1_apply<T, DIM, 0>(u, v, scale);
        else if (A.op == 1)
          q1_apply<T, DIM, 1>(u, v, scale);
        else
#pragma unroll
          for (int i = 0; i < NPE; i++) v[i] = scale * u[i];
      } else {
        q2_apply<T, DIM>(u, kref, A.op, scale);
#pragma unroll
        for (int i = 0; i < NPE; i++) v[i] = u[i];
      }
    }
    if constexpr (ONE) {
      // parent values are consumed: y_loc becomes the accumulator
      __syncthreads();
      for (int j = tid; j < th.nnode; j += nthr) y_loc[j] = T(0);
    }
    // 5. Direct outputs, one phase per rank.
#pragma unroll
    for (int r = 0; r < NPE; r++) {
      __syncthreads();
      if (loc[r] != kNone) y_loc[loc[r] * nq + q] += v[r];
    }
    // 6. Transposed interpolation of missing ranks, summed over each sibling
    //    group inside the wavefront (pairwise, fixed order), then per target
    //    node over its (group, parent rank) entries sorted by target.
    if (th.nhe > 0) {
      if (__ballot(hang)) {  // wavefronts without hanging leaves have no groups
        T w[NPE];
#pragma unroll
        for (int r = 0; r < NPE; r++) w[r] = (hang && loc[r] == kNone) ? v[r] : T(0);
        interp<T, DIM, P, true>(w, child);
#pragma unroll
        for (int s = 0; s < NPE; s++) w[s] = (hang && loc[s] == kNone) ? w[s] : T(0);
        // Pairwise sums inside rows of 16 lanes with DPP row shifts (lane i
        // reads lane i + d of its row; a 0/1 factor keeps the run), then a run
        // that continues into the next row adds that row's first partial (a
        // group spans at most two rows).
        group_step<T, 0x101, 1>(w, wl, run_end);
        if constexpr (DIM > 1) group_step<T, 0x102, 2>(w, wl, run_end);
        if constexpr (DIM > 2) group_step<T, 0x104, 4>(w, wl, run_end);
        if constexpr (DIM > 3) group_step<T, 0x108, 8>(w, wl, run_end);
        const int nxt = (wl | 15) + 1;
        if (__ballot(nxt < run_end)) {
          const T fx = nxt < run_end ? T(1) : T(0);
#pragma unroll
          for (int s = 0; s < NPE; s++) w[s] = fma(__shfl(w[s], nxt & 63, 64), fx, w[s]);
        }
        if (hang && aux_ghead(aux))
#pragma unroll
          for (int s = 0; s < NPE; s++) ubuf[grp_id * NPE + s] = w[s];
      }
      __syncthreads();
      for (int k = tid; k < th.nhe; k += nthr) {
        const int j = (int)(hent[k] >> 16);
        if (k > 0 && (int)(hent[k - 1] >> 16) == j) continue;  // not a segment head
        T acc = T(0);
        for (int m = k; m < th.nhe && (int)(hent[m] >> 16) == j; m++) {
          const uint32_t e = hent[m];
          acc += ubuf[(int)(e & 255u) * NPE + (int)((e >> 8) & 255u)];
        }
        y_loc[j * nq + q] += acc;
      }
    }
    __syncthreads();
  }
  // 7. Write out: owned nodes to y, shared nodes to their partial slot.
  for (int j0 = tid; j0 < th.nnode; j0 += kNB * nthr) {
    int o[kNB];
#pragma unroll
    for (int b = 0; b < kNB; b++) {
      const int j = j0 + b * nthr;
      o[b] = j0 == tid ? o_first[b] : (j < th.nnode ? A.tile_out[th.node0 + j] : INT32_MIN);
    }
    uint32_t d[kNB];  // Dirichlet flags, loaded unconditionally as in the stage
#pragma unroll
    for (int b = 0; b < kNB; b++) d[b] = 0u;
    if (A.dir_out)
#pragma unroll
      for (int b = 0; b < kNB; b++) d[b] = A.dir_out[o[b] >= 0 ? o[b] : 0];
#pragma unroll
    for (int b = 0; b < kNB; b++) {
      const int j = j0 + b * nthr;
      if (o[b] == INT32_MIN) continue;
      for (int q = 0; q < nq; q++) {
        const T acc = y_loc[j * nq + q];
        if (o[b] >= 0)
          y[(size_t)o[b] * A.ld + q] = d[b] ? T(0) : acc;
        else
          partial[(size_t)(-o[b] - 1) * nq + q] = acc;
      }
    }
  }
}

// Waves per SIMD the compiler is asked to fit: 8, the most, for the p = 1
// single-column kernels whose vector registers allow it (their scalar
// registers, the uniform tile fields, otherwise hold them to 7); 1 leaves the
// register allocation free.
template <typename T, int DIM, int P, bool ONE>
constexpr int tile_waves() {
  return ONE && P == 1 && (sizeof(T) == 4 || DIM <= 3) ? 8 : 1;
}

template <typename T, int DIM, int P, bool ONE>
__global__ void __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(tile_waves<T, DIM, P, ONE>())))
tile_kernel(TileArgs A, const T *__restrict__ kref, const T *__restrict__ carry_w, const T *__restrict__ x,
            T *__restrict__ y, T *__restrict__ partial) {
  if (A.cancel && *A.cancel) return;
  tile_body<T, DIM, P, ONE>(A, xcd_tile(blockIdx.x, gridDim.x), kref, carry_w, x, y, partial);
}

// A plan of a batch: its maps (with nq and ld set for the batch) and vectors.
template <typename T>
struct BatchPlan {
  TileArgs A;
  const T *kref, *carry_w, *x;
  T *y, *partial;
  const int32_t *snode, *soff;
  int nshared;
};

// Index of the last start <= v (starts ascending, starts[0] = 0).  v is the
// same for the whole workgroup at every call site (a tile or block index), so
// the result is made wave-uniform explicitly: the record it selects is then
// read with scalar loads into scalar registers instead of occupying vector
// registers for its pointers (tile_kernel_batch<float, 3, 2, false> otherwise
// needs more VGPRs than the single-plan kernel).
__device__ __forceinline__ int find_range(const int *__restrict__ starts, int n, int v) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (starts[mid] <= v) lo = mid; else hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// All tiles of a batch of same-(dim, p) plans in one launch: batch tile
// t = xcd_tile(b) of workgroup b is tile t - tile_start[k] of plan k.
template <typename T, int DIM, int P, bool ONE>
__global__ void __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(tile_waves<T, DIM, P, ONE>())))
tile_kernel_batch(const BatchPlan<T> *__restrict__ plans, const int *__restrict__ tile_start, int nplans) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int k = find_range(tile_start, nplans, t);
  const BatchPlan<T> bp = plans[k];  // a copy: its fields live in scalar registers
  if (bp.A.cancel && *bp.A.cancel) return;
  tile_body<T, DIM, P, ONE>(bp.A, t - tile_start[k], bp.kref, bp.carry_w, bp.x, bp.y, bp.partial);
}

// Sum of the partial slots [a, b) of column q in slot order (a < b); the
// first 8 slots load together.
template <typename T>
__device__ __forceinline__ T sum_partials(const T *__restrict__ partial, int a, int b, int nq, int q) {
  T v[8];
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = a + j < b ? partial[(size_t)(a + j) * nq + q] : T(0);
  T acc = v[0];
#pragma unroll
  for (int j = 1; j < 8; j++)
    if (a + j < b) acc += v[j];
  for (int k = a + 8; k < b; k++) acc += partial[(size_t)k * nq + q];
  return acc;
}

template <typename T>
__global__ void __launch_bounds__(256) fixup_kernel(int ns, const int32_t *__restrict__ snode,
                                                    const int32_t *__restrict__ soff, const T *__restrict__ partial,
                                                    const uint8_t *__restrict__ dirichlet, int nq, int ld,
                                                    T *__restrict__ y, const int *__restrict__ cancel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns || (cancel && *cancel)) return;
  const int n = snode[i];
  const int a = soff[i], b = soff[i + 1];
  const bool z = dirichlet && dirichlet[n];
  for (int q = 0; q < nq; q++) y[(size_t)n * ld + q] = z ? T(0) : sum_partials(partial, a, b, nq, q);
}

// LDS bytes of one tile for nq right-hand sides (layout of tile_kernel).
template <typename T, int NPE>
inline size_t tile_lds_bytes(const TileDims &D, int nq) {
  auto r16 = [](size_t b) { return (b + 15) & ~(size_t)15; };
  return r16((size_t)D.max_nodes * nq * sizeof(T)) + r16(tile_yloc_elems<NPE>(D, nq) * sizeof(T)) +
         r16(tile_region_bytes<T, NPE>(D)) +
         r16((size_t)NPE * nq * sizeof(T)) + r16((size_t)D.max_he * 4) + r16((size_t)(D.max_hang + 1) * 4) +
         r16((size_t)D.max_grp * NPE * 2) + r16((size_t)D.tile_leaves * 2) + r16((size_t)D.max_rd * 2);
}

}  // namespace dkt
