// Shared pieces of the device plan: device buffers, tensor-product
// interpolation in registers, the tile layout and the plan object.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#define HIP_OK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw std::runtime_error(std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

namespace dkt {

constexpr uint16_t kNone = 0xFFFF;
constexpr int kTileThreads = 256;
constexpr int kDefaultTile = 128;
constexpr int kMaxConfTerms = 2048;  // constraint terms per conforming tile
constexpr int kMaxConfVirt = 1024;   // hanging points per conforming tile
constexpr int kMaxReaders = 32;  // stale-buffer readers per tile (bounds the replay outputs in LDS)
constexpr size_t kLdsLimit = 64 * 1024;

// Tile of workgroup b of an n-workgroup launch.  Workgroups go to the eight
// XCDs round-robin (b to XCD b % 8), each with its own L2; this map hands every
// XCD one contiguous run of SFC-consecutive tiles, taken in order, so the nodes
// that neighbouring tiles share are fetched into one L2 instead of eight.  A
// bijection of [0, n) for every n.
__host__ __device__ inline int xcd_tile(int b, int n) {
  constexpr int kXcd = 8;
  const int q = n / kXcd, r = n % kXcd, x = b % kXcd, i = b / kXcd;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// Host-only plan construction (DKT_PLAN_DRYRUN): sizes are recorded, nothing is
// allocated on the device.
extern thread_local bool g_dry_run;

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    n = count;
    if (count && !g_dry_run) HIP_OK(hipMalloc((void **)&p, count * sizeof(T)));
  }
  void upload(const std::vector<T> &v) {
    alloc(v.size());
    if (n && !g_dry_run) HIP_OK(hipMemcpy(p, v.data(), n * sizeof(T), hipMemcpyHostToDevice));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
};

template <int N, int D>
struct Pow {
  static constexpr int v = N * Pow<N, D - 1>::v;
};
template <int N>
struct Pow<N, 0> {
  static constexpr int v = 1;
};

// Ip<P>::w(c, k, i) = l_k(child point i): Lagrange basis of parent rank k at
// the i-th node of child half c (RefElement ip_1D_0/1, FEM/src/refel.cpp:158-165).
template <int P>
struct Ip;
template <>
struct Ip<1> {
  __host__ __device__ static constexpr double w(int c, int k, int i) {
    return (c == 0) ? (k == 0 ? (i == 0 ? 1.0 : 0.5) : (i == 0 ? 0.0 : 0.5))
                    : (k == 0 ? (i == 0 ? 0.5 : 0.0) : (i == 0 ? 0.5 : 1.0));
  }
};
template <>
struct Ip<2> {
  __host__ __device__ static constexpr double w(int c, int k, int i) {
    const double t = 0.5 * (c + 0.5 * i);
    return k == 0 ? (t - 0.5) * (t - 1.0) / 0.5 : (k == 1 ? t * (t - 1.0) / -0.25 : t * (t - 0.5) / 0.5);
  }
};

// In place, axis 0 fastest (KroneckerProduct, FEM/include/tensor.h:181-203):
//   forward  a[r] <- sum_s W_c[r][s] a[s]   (parent ranks -> child ranks)
//   backward a[s] <- sum_r W_c[r][s] a[r]   (child ranks -> parent ranks)
template <typename T, int DIM, int P, bool BWD>
__device__ __forceinline__ void interp(T *a, int c) {
  constexpr int n1 = P + 1, NPE = Pow<P + 1, DIM>::v;
  if constexpr (P == 1) {
    // In place, one axis at a time over the pairs (a0, a1) of that axis:
    // forward, the child's node at the child bit keeps the parent value and the
    // other is the midpoint; backward is the transpose,
    //   cb = 0: (a0 + a1/2, a1/2),  cb = 1: (a0/2, a1 + a0/2),
    // written with 0/1 factors: every product with them is exact, so each
    // fma adds an exact zero or rounds once like the plain sum.
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      const bool cb = (c >> d) & 1;
      const T one_c = cb ? T(1) : T(0), one_n = T(1) - one_c;
      const T half_c = T(0.5) * one_c, half_n = T(0.5) * one_n;
#pragma unroll
      for (int r = 0; r < NPE; r++) {
        if ((r >> d) & 1) continue;
        const T a0 = a[r], a1 = a[r | (1 << d)];
        if constexpr (!BWD) {
          const T m = T(0.5) * (a0 + a1);
          a[r] = cb ? m : a0;
          a[r | (1 << d)] = cb ? a1 : m;
        } else {
          const T h = fma(a0, half_c, a1 * half_n);
          a[r] = fma(a0, one_n, h);
          a[r | (1 << d)] = fma(a1, one_c, h);
        }
      }
    }
    return;
  }
  if constexpr (P == 2) {
    // In place, one axis line (3 values) at a time.
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      const int cb = (c >> d) & 1;
      constexpr int strides[4] = {1, 3, 9, 27};
      const int st = strides[d];
      T w[3][3];  // w[k][i]: parent node k -> child node i (forward)
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int i = 0; i < 3; i++) w[k][i] = cb ? (T)Ip<2>::w(1, k, i) : (T)Ip<2>::w(0, k, i);
#pragma unroll
      for (int r = 0; r < NPE; r++) {
        if ((r / st) % 3 != 0) continue;
        const T a0 = a[r], a1 = a[r + st], a2 = a[r + 2 * st];
        if constexpr (!BWD) {
          a[r] = w[0][0] * a0 + w[1][0] * a1 + w[2][0] * a2;
          a[r + st] = w[0][1] * a0 + w[1][1] * a1 + w[2][1] * a2;
          a[r + 2 * st] = w[0][2] * a0 + w[1][2] * a1 + w[2][2] * a2;
        } else {
          a[r] = w[0][0] * a0 + w[0][1] * a1 + w[0][2] * a2;
          a[r + st] = w[1][0] * a0 + w[1][1] * a1 + w[1][2] * a2;
          a[r + 2 * st] = w[2][0] * a0 + w[2][1] * a1 + w[2][2] * a2;
        }
      }
    }
    return;
  }
  T b[NPE];
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
      for (int k = 0; k < n1; k++)
        acc += (T)(BWD ? Ip<P>::w(cb, id, k) : Ip<P>::w(cb, k, id)) * a[base + k * st];
      b[r] = acc;
    }
#pragma unroll
    for (int i = 0; i < NPE; i++) a[i] = b[i];
  }
}

// Element operators of Q1 (p = 1) elements on the unit cube, applied in the
// Walsh-Hadamard basis.  With the 1D matrices M1 = [[1/3,1/6],[1/6,1/3]] and
// K1 = [[1,-1],[-1,1]], the mass matrix M1 x ... x M1 and the stiffness
// sum_d M1 x .. K1 (axis d) .. x M1 depend on two ranks only through the
// axes on which they differ, so H = [[1,1],[1,-1]]^(x dim) diagonalises them:
// A = H diag(lambda) H / 2^dim, with, for a frequency f of Hamming weight k,
//   mass:      lambda = (1/6)^k (1/2)^(dim-k)
//   stiffness: lambda = 2 k (1/6)^(k-1) (1/2)^(dim-k).
// H u costs dim 2^dim additions, against 4^dim multiply-adds for the dense matrix.
__host__ __device__ constexpr double q1_lambda(int dim, int op, int k) {
  double v = op == 0 ? 2.0 * k : 1.0;
  for (int i = 0; i < dim; i++) {
    const bool diff = i < k;
    if (op == 0 && diff && i == 0) continue;  // the 2 k factor replaces one 1/6 for the axis carrying K1
    v *= diff ? 1.0 / 6.0 : 0.5;
  }
  return v;
}

template <typename T, int DIM>
__device__ __forceinline__ void fwht(T *a) {
  constexpr int NPE = 1 << DIM;
#pragma unroll
  for (int d = 0; d < DIM; d++)
#pragma unroll
    for (int i = 0; i < NPE; i++)
      if (!((i >> d) & 1)) {
        const T x = a[i], y = a[i | (1 << d)];
        a[i] = x + y;
        a[i | (1 << d)] = x - y;
      }
}

// v = scale * A u for the Q1 operator op (0 stiffness, 1 mass): the
// Walsh-Hadamard transform H diagonalises A, A = H diag(lambda) H / 2^dim, with
// lambda depending on the Hamming weight of the mode only.  scale is a power
// of two, so folding it into the eigenvalues (one product per Hamming weight)
// is exact too: both forms give bitwise scale * (H diag(lambda) H u / 2^dim).
// FP32 folds it (measured faster); FP64 scales at the end, which keeps the
// five folded eigenvalues out of registers (the fold cost the 4D FP64 tile
// kernel 9% in time: 1.11 ms against 1.02 ms).
template <typename T, int DIM, int OP>
__device__ __forceinline__ void q1_apply(const T *u, T *v, T scale) {
  constexpr int NPE = 1 << DIM;
  T a[NPE];
#pragma unroll
  for (int i = 0; i < NPE; i++) a[i] = u[i];
  fwht<T, DIM>(a);
  if constexpr (sizeof(T) == 4) {
    T ls[DIM + 1];
#pragma unroll
    for (int k = 0; k <= DIM; k++) ls[k] = (T)(q1_lambda(DIM, OP, k) / NPE) * scale;
#pragma unroll
    for (int f = 0; f < NPE; f++) {
      if (OP == 0 && f == 0)
        a[f] = T(0);
      else
        a[f] *= ls[__builtin_popcount(f)];
    }
    fwht<T, DIM>(a);
#pragma unroll
    for (int i = 0; i < NPE; i++) v[i] = a[i];
  } else {
#pragma unroll
    for (int f = 0; f < NPE; f++) a[f] *= (T)(q1_lambda(DIM, OP, __builtin_popcount(f)) / NPE);
    fwht<T, DIM>(a);
#pragma unroll
    for (int i = 0; i < NPE; i++) v[i] = scale * a[i];
  }
}

// p = 2 element operators by sum factorisation, in place on u:
// f = [B (3 x 3, row-major), lambda (3), D (3^DIM)] with M1 = B^T B,
// K1 = B^T diag(lambda) B and D[r] = sum_d lambda[i_d(r)] (summed in T, axis 0
// first, by q2_table on the host);
// u <- scale (B^T)^(x DIM) D B^(x DIM) u for the stiffness (op 0), without D
// for the mass (op 1); op 2 (identity) scales only.  D comes from the uniform
// table, not summed per work-item: per-lane sums kept the 27 products live in
// vector registers (166 against 113 VGPRs for FP64 in 3D).  scale is a power of
// two, so applying it last gives the bits of applying it with D.
// *p for an address that is the same on every lane, held in scalar registers
// from then on.  The compiler does this itself for a kernel's own pointer
// arguments; through a pointer read from memory (a batch's plan record) it
// cannot prove that the kernel's stores leave the value unchanged, and would
// keep one copy per lane in vector registers.
template <typename T>
__device__ __forceinline__ T uniform_load(const T *p) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t b = __builtin_bit_cast(uint64_t, *p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
    return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
  } else {
    return __builtin_bit_cast(T, (uint32_t)__builtin_amdgcn_readfirstlane(__builtin_bit_cast(uint32_t, *p)));
  }
}

template <typename T, int DIM>
__device__ __forceinline__ void q2_apply(T *u, const T *__restrict__ f, int op, T scale) {
  constexpr int NPE = Pow<3, DIM>::v;
  if (op == 2) {
#pragma unroll
    for (int i = 0; i < NPE; i++) u[i] *= scale;
    return;
  }
  T B[9];
#pragma unroll
  for (int k = 0; k < 9; k++) B[k] = uniform_load(f + k);
  constexpr int strides[4] = {1, 3, 9, 27};
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const int st = strides[d];
#pragma unroll
    for (int r = 0; r < NPE; r++) {
      if ((r / st) % 3 != 0) continue;
      const T a0 = u[r], a1 = u[r + st], a2 = u[r + 2 * st];
      u[r] = B[0] * a0 + B[1] * a1 + B[2] * a2;
      u[r + st] = B[3] * a0 + B[4] * a1 + B[5] * a2;
      u[r + 2 * st] = B[6] * a0 + B[7] * a1 + B[8] * a2;
    }
  }
  if (op == 0)
#pragma unroll
    for (int r = 0; r < NPE; r++) u[r] *= uniform_load(f + 12 + r);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const int st = strides[d];
#pragma unroll
    for (int r = 0; r < NPE; r++) {
      if ((r / st) % 3 != 0) continue;
      const T a0 = u[r], a1 = u[r + st], a2 = u[r + 2 * st];
      u[r] = B[0] * a0 + B[3] * a1 + B[6] * a2;
      u[r + st] = B[1] * a0 + B[4] * a1 + B[7] * a2;
      u[r + 2 * st] = B[2] * a0 + B[5] * a1 + B[8] * a2;
    }
  }
#pragma unroll
  for (int r = 0; r < NPE; r++) u[r] *= scale;
}

// One tile = one workgroup = up to kTileThreads SFC-consecutive leaves.
struct TileHdr {
  int32_t leaf0, nleaf;  // leaves
  int32_t node0, nnode;  // local nodes (tile_nodes / tile_out)
  int32_t grp0, ngrp;    // sibling groups (group_ploc rows)
  int32_t he0, nhe;      // hanging-transpose entries, sorted by target node
  int32_t rd0, nrd;      // stale-buffer readers, grouped by replay block, by hanging index within a block
  int32_t bk0, nbk;      // replay blocks
  int32_t cr0, ncr;      // terms of the parent-buffer state at the tile start (carried slots)
};

// Leaf word: bits 0-4 level, 5-8 child index, 9 hanging, 10-17 hanging index
// within the tile, 18-25 sibling group (tile-local), 26 first leaf of its group.
// A sibling group is a run of consecutive hanging leaves with one parent that
// does not cross a 64-leaf (wavefront) boundary of the tile.
__host__ __device__ inline int aux_level(uint32_t a) { return a & 31u; }
__host__ __device__ inline int aux_child(uint32_t a) { return (a >> 5) & 15u; }
__host__ __device__ inline bool aux_hang(uint32_t a) { return (a >> 9) & 1u; }
__host__ __device__ inline int aux_hidx(uint32_t a) { return (a >> 10) & 255u; }
__host__ __device__ inline int aux_group(uint32_t a) { return (a >> 18) & 255u; }
__host__ __device__ inline bool aux_ghead(uint32_t a) { return (a >> 26) & 1u; }
// Hanging-transpose entry (uint32): sibling group (8) | parent rank (8) << 8 |
// target local node (16) << 16.
// Reader (uint16): hanging index (8) | tile-local leaf (8) << 8.
// Replay block (uint64): start hanging index (16, 0xFFFF = from the carried
// tile-start state) | last hanging index (16) << 16 | first reader (16) << 32 |
// reader count (16) << 48.
// Chain word (in LDS, per hanging index): child (8) | sibling group (16) << 8.

// Per-tile maxima that size the LDS carve.
struct TileDims {
  int max_nodes, max_grp, max_he, max_rd, max_cr, tile_leaves, max_hang;
};

struct TileArgs {
  const TileHdr *tiles;
  const int32_t *tile_nodes, *tile_out;
  const uint16_t *leaf_loc;
  const uint32_t *leaf_aux;
  const uint16_t *group_ploc;
  const uint32_t *hentries;
  const uint16_t *readers;
  const uint64_t *blocks;
  const uint32_t *carry_key;  // per tile: nPe + 1 slot offsets, then buffer rank (8) << 16 | local node (16)
  const uint8_t *dir_in;   // non-null: x reads as zero at boundary nodes
  const uint8_t *dir_out;  // non-null: y is zero at boundary nodes
  const int *cancel;  // non-null: the tile returns at once when *cancel != 0
  int op;  // DKT_OP_STIFFNESS, DKT_OP_MASS or DKT_OP_IDENTITY
  int hp, nq, ld;
  TileDims dims;
};

}  // namespace dkt
