// Conjugate gradients on the device, the caller of the matvec in the
// reference (HeatMat::cgSolve, FEM/examples/src/heatMat.cpp:165-325):
// unpreconditioned CG with one matvec per iteration, alpha = (r.r)/(p.Ap),
// beta = (r1.r1)/(r0.r0), stopping when ||r||_inf / ||b||_inf <= tol
// (normLInfty, include/mathUtils.tcc:120-127; dot, :204-221).
//
// Device design: the scalars never leave the device.  Dot products and max
// norms are two-stage reductions over a fixed partition (cg_blocks(n) blocks
// of kCgThreads work-items, fixed tree order): every block that needs alpha or
// beta re-reduces the same partials in the same order, so all blocks agree
// bit for bit and the solve is reproducible run to run.  Convergence raises a
// device flag; the kernels of iterations queued after it (matvec included)
// return at once, and the host only checks the flag every kCgChunk
// iterations.
//
// Mixed precision, DKT_CG_MIXED_RELIABLE: CG with reliable updates.  The iteration
// runs with FP32 vectors and matvecs (FP32 accumulation inside the matvec) and
// FP64 dot products and scalars, accumulating its update of x in FP32.  When
// the recursive FP32 residual has dropped by kReliableDrop since the last
// update (or reaches tol), the update is added to x in FP64, the true residual
// b - A x is formed with an FP64 matvec and replaces the FP32 one, and the
// iteration continues with the same search direction (beta from the replaced
// residual), so no Krylov information is discarded.  The solve stops on the
// true FP64 residual.  A replacement that does not reduce the true residual
// restarts the search direction.
// DKT_CG_MIXED_RESTART is iterative refinement: each inner solve runs the same
// FP32 CG from zero on A d = r / ||r||_inf to a relative residual of kInnerTol
// (or looser when tol / ||r|| is larger), then x += ||r|| d.
#include <algorithm>
#include <cmath>
#include <deque>

#include "../../include/dkt.h"
#include "capi_common.h"
#include "plan_common.h"
#include "plan_impl.h"

namespace dkt {
namespace {

constexpr int kCgBlocks = 512;
constexpr int kCgThreads = 256;
constexpr int kCgChunk = 8;
constexpr double kInnerTol = 1e-4;
constexpr int kMaxOuter = 64;
constexpr double kReliableDrop = 0.1;
// Scalar slots.
enum { S_RR0 = 0, S_RR1 = 1, S_NORMB = 2, S_RESID = 3, S_SCALE = 4, S_BETA = 5, S_COUNT = 8 };

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *sh) {
  const int t = threadIdx.x;
  __syncthreads();
  sh[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = kCgThreads / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  return sh[0];
}

template <typename T>
__device__ __forceinline__ T block_max(T v, T *sh) {
  const int t = threadIdx.x;
  __syncthreads();
  sh[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = kCgThreads / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  return sh[0];
}

// Sum (or max) of the kCgBlocks partials, identical in every block.
__device__ __forceinline__ double parts_sum(const double *part, double *sh) {
  return block_sum(part[threadIdx.x] + part[threadIdx.x + kCgThreads], sh);
}
__device__ __forceinline__ double parts_max(const double *part, double *sh) {
  return block_max(fmax(part[threadIdx.x], part[threadIdx.x + kCgThreads]), sh);
}

// Blocks of a solve over n entries: at least 4 entries per work-item, at most
// kCgBlocks.  Blocks past it write no partials; the partial arrays keep zeros
// there (zeroed with the workspace), which the fixed-order sums and max norms
// over all kCgBlocks slots ignore.
__host__ __device__ inline int cg_blocks(int n) {
  const int b = (n + 4 * kCgThreads - 1) / (4 * kCgThreads);
  return b < 1 ? 1 : (b > kCgBlocks ? kCgBlocks : b);
}
#define CG_LOOP(i, n) \
  for (int i = blockIdx.x * kCgThreads + threadIdx.x; i < (n); i += cg_blocks(n) * kCgThreads)

// The bodies of the iteration kernels run one solve's block blockIdx.x of
// kCgBlocks; the single-solve kernels and the batched ones (one solve per
// blockIdx.y) share them, so a solve's arithmetic and reduction order do not
// depend on what it is batched with.

// r = b - ax (ax may be null: x = 0), p = r; partials of r.r, max|r|, max|b|.
template <typename T>
__device__ __forceinline__ void init_body(int n, const T *__restrict__ b, const T *__restrict__ ax,
                                          T *__restrict__ r, T *__restrict__ p, double *__restrict__ part) {
  __shared__ double sh[kCgThreads];
  double rr = 0, mr = 0, mb = 0;
  CG_LOOP(i, n) {
    const T ri = ax ? b[i] - ax[i] : b[i];
    r[i] = ri;
    p[i] = ri;
    rr += (double)ri * (double)ri;
    mr = fmax(mr, (double)fabs(ri));
    mb = fmax(mb, (double)fabs(b[i]));
  }
  rr = block_sum(rr, sh);
  mr = block_max(mr, sh);
  mb = block_max(mb, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rr;
    part[kCgBlocks + blockIdx.x] = mr;
    part[2 * kCgBlocks + blockIdx.x] = mb;
  }
}
template <typename T>
__global__ void __launch_bounds__(kCgThreads) k_init(int n, const T *__restrict__ b, const T *__restrict__ ax,
                                                     T *__restrict__ r, T *__restrict__ p, double *__restrict__ part) {
  init_body(n, b, ax, r, p, part);
}

__device__ __forceinline__ void init_fin_body(const double *__restrict__ part, double *__restrict__ scal, double tol,
                                              int *__restrict__ flag) {
  __shared__ double sh[kCgThreads];
  const double rr = parts_sum(part, sh);
  const double mr = parts_max(part + kCgBlocks, sh);
  double mb = parts_max(part + 2 * kCgBlocks, sh);
  if (threadIdx.x == 0) {
    if (mb == 0.0) mb = 1.0;
    scal[S_RR0] = rr;
    scal[S_NORMB] = mb;
    scal[S_RESID] = mr / mb;
    flag[0] = mr / mb <= tol ? 1 : 0;
    flag[1] = 0;
  }
}
__global__ void __launch_bounds__(kCgThreads) k_init_fin(const double *__restrict__ part, double *__restrict__ scal,
                                                         double tol, int *__restrict__ flag) {
  init_fin_body(part, scal, tol, flag);
}

template <typename T>
__device__ __forceinline__ void dot_body(int n, const T *__restrict__ a, const T *__restrict__ b,
                                         double *__restrict__ part, const int *__restrict__ flag) {
  if (flag[0]) return;
  __shared__ double sh[kCgThreads];
  double acc = 0;
  CG_LOOP(i, n) acc += (double)a[i] * (double)b[i];
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}
template <typename T>
__global__ void __launch_bounds__(kCgThreads) k_dot(int n, const T *__restrict__ a, const T *__restrict__ b,
                                                    double *__restrict__ part, const int *__restrict__ flag) {
  dot_body(n, a, b, part, flag);
}

// alpha = rr / (p.Ap); x += alpha p; r -= alpha Ap; partials of r.r and max|r|.
template <typename T>
__device__ __forceinline__ void update_xr_body(int n, T *__restrict__ x, T *__restrict__ r, const T *__restrict__ p,
                                               const T *__restrict__ ap, const double *__restrict__ part_pap,
                                               const double *__restrict__ scal, int par, double *__restrict__ part,
                                               const int *__restrict__ flag) {
  if (flag[0]) return;
  __shared__ double sh[kCgThreads];
  const T alpha = (T)(scal[S_RR0 + par] / parts_sum(part_pap, sh));
  double rr = 0, mr = 0;
  CG_LOOP(i, n) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * ap[i];
    r[i] = ri;
    rr += (double)ri * (double)ri;
    mr = fmax(mr, (double)fabs(ri));
  }
  rr = block_sum(rr, sh);
  mr = block_max(mr, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rr;
    part[kCgBlocks + blockIdx.x] = mr;
  }
}
template <typename T>
__global__ void __launch_bounds__(kCgThreads) k_update_xr(int n, T *__restrict__ x, T *__restrict__ r,
                                                          const T *__restrict__ p, const T *__restrict__ ap,
                                                          const double *__restrict__ part_pap,
                                                          const double *__restrict__ scal, int par,
                                                          double *__restrict__ part, const int *__restrict__ flag) {
  update_xr_body(n, x, r, p, ap, part_pap, scal, par, part, flag);
}

// Convergence test, then beta = r1.r1 / r0.r0 and p = r + beta p.  Block 0
// publishes the new r.r (into the other parity slot), the residual, the
// iteration count and the converged flag.
template <typename T>
__device__ __forceinline__ void update_p_body(int n, T *__restrict__ p, const T *__restrict__ r,
                                              const double *__restrict__ part, double *__restrict__ scal, int par,
                                              double tol, int *__restrict__ flag) {
  if (flag[0]) return;
  __shared__ double sh[kCgThreads];
  const double rr1 = parts_sum(part, sh);
  const double resid = parts_max(part + kCgBlocks, sh) / scal[S_NORMB];
  const bool done = resid <= tol;
  if (!done) {
    const T beta = (T)(rr1 / scal[S_RR0 + par]);
    CG_LOOP(i, n) p[i] = r[i] + beta * p[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[S_RR0 + (par ^ 1)] = rr1;
    scal[S_RESID] = resid;
    flag[1] += 1;
    if (done) flag[0] = 1;
  }
}
template <typename T>
__global__ void __launch_bounds__(kCgThreads) k_update_p(int n, T *__restrict__ p, const T *__restrict__ r,
                                                         const double *__restrict__ part, double *__restrict__ scal,
                                                         int par, double tol, int *__restrict__ flag) {
  update_p_body(n, p, r, part, scal, par, tol, flag);
}

// FP64 residual r = b - ax and its partial max norms (outer refinement step).
__global__ void __launch_bounds__(kCgThreads) k_resid64(int n, const double *__restrict__ b,
                                                        const double *__restrict__ ax, double *__restrict__ r,
                                                        double *__restrict__ part) {
  __shared__ double sh[kCgThreads];
  double mr = 0, mb = 0;
  CG_LOOP(i, n) {
    const double ri = b[i] - ax[i];
    r[i] = ri;
    mr = fmax(mr, fabs(ri));
    mb = fmax(mb, fabs(b[i]));
  }
  mr = block_max(mr, sh);
  mb = block_max(mb, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = mr;
    part[kCgBlocks + blockIdx.x] = mb;
  }
}

// scal[S_RESID] = max|r| / max|b| (max|b| = 1 if b = 0); scal[S_SCALE] = max|r|.
__global__ void __launch_bounds__(kCgThreads) k_resid64_fin(const double *__restrict__ part, double *__restrict__ scal) {
  __shared__ double sh[kCgThreads];
  const double mr = parts_max(part, sh);
  double mb = parts_max(part + kCgBlocks, sh);
  if (threadIdx.x == 0) {
    if (mb == 0.0) mb = 1.0;
    scal[S_NORMB] = mb;
    scal[S_RESID] = mr / mb;
    scal[S_SCALE] = mr;
  }
}

// b32 = r / s (s = max|r|, read on the device).
__global__ void __launch_bounds__(kCgThreads) k_to32(int n, const double *__restrict__ r, const double *__restrict__ scal,
                                                     float *__restrict__ b32, float *__restrict__ x32) {
  const double s = scal[S_SCALE];
  const double inv = s > 0.0 ? 1.0 / s : 0.0;
  CG_LOOP(i, n) {
    b32[i] = (float)(r[i] * inv);
    x32[i] = 0.0f;
  }
}

// x += s d32.
__global__ void __launch_bounds__(kCgThreads) k_add32(int n, double *__restrict__ x, const float *__restrict__ d32,
                                                      const double *__restrict__ scal) {
  const double s = scal[S_SCALE];
  CG_LOOP(i, n) x[i] += s * (double)d32[i];
}

// Reliable-update kernels (FP32 iteration, FP64 x and residual).
// x += x32, x32 = 0.
__global__ void __launch_bounds__(kCgThreads) k_add32u(int n, double *__restrict__ x, float *__restrict__ x32) {
  CG_LOOP(i, n) {
    x[i] += (double)x32[i];
    x32[i] = 0.0f;
  }
}

// r32 = r64 (restart: p32 = r32 and x32 = 0 too); partials of r32.r32.
__global__ void __launch_bounds__(kCgThreads) k_rel32(int n, const double *__restrict__ r64, float *__restrict__ r32,
                                                      float *__restrict__ p32, float *__restrict__ x32, int restart,
                                                      double *__restrict__ part) {
  __shared__ double sh[kCgThreads];
  double rr = 0;
  CG_LOOP(i, n) {
    const float ri = (float)r64[i];
    r32[i] = ri;
    if (restart) {
      p32[i] = ri;
      x32[i] = 0.0f;
    }
    rr += (double)ri * (double)ri;
  }
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

// New r.r into the slot the next iteration reads (par ^ 1), beta = new / old
// r.r (0 on a restart), ||b|| from the FP64 residual, flags cleared.
__global__ void __launch_bounds__(kCgThreads) k_rel32_fin(const double *__restrict__ part, double *__restrict__ scal32,
                                                          const double *__restrict__ scal64, int par, int restart,
                                                          int *__restrict__ flag) {
  __shared__ double sh[kCgThreads];
  const double rr = parts_sum(part, sh);
  if (threadIdx.x == 0) {
    scal32[S_BETA] = restart ? 0.0 : rr / scal32[S_RR0 + par];
    scal32[S_RR0 + (par ^ 1)] = rr;
    scal32[S_NORMB] = scal64[S_NORMB];
    flag[0] = 0;
    if (restart == 2) flag[1] = 0;  // first start: iteration count
  }
}

// p = r + beta p (after a residual replacement).
__global__ void __launch_bounds__(kCgThreads) k_pupd32(int n, float *__restrict__ p32, const float *__restrict__ r32,
                                                       const double *__restrict__ scal32) {
  const float beta = (float)scal32[S_BETA];
  CG_LOOP(i, n) p32[i] = r32[i] + beta * p32[i];
}

template <typename T>
struct CgVectors {
  T *r, *p, *ap;
  double *part, *scal;
};

// Runs up to max_iter CG iterations on A x = b in precision T from the current x.
// Returns the iterations performed; the final relative residual is scal[S_RESID].
template <typename T>
int cg_run(dkt_plan *pl, int dtype, const T *b, T *x, bool x_is_zero, int max_iter, double tol, int masked,
           const CgVectors<T> &v, int *flag, int *host, hipStream_t st, int *matvecs) {
  const int n = (int)pl->N;
  const dim3 grid(cg_blocks(n)), blk(kCgThreads);
  if (!x_is_zero) {
    plan_apply(pl, dtype, x, v.ap, 1, masked, st, nullptr);
    (*matvecs)++;
  }
  hipLaunchKernelGGL((k_init<T>), grid, blk, 0, st, n, b, x_is_zero ? nullptr : v.ap, v.r, v.p, v.part);
  hipLaunchKernelGGL(k_init_fin, dim3(1), blk, 0, st, v.part, v.scal, tol, flag);
  HIP_OK(hipGetLastError());
  int it = 0;
  HIP_OK(hipMemcpyAsync(host, flag, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  if (host[0]) return 0;
  while (it < max_iter) {
    const int chunk = std::min(kCgChunk, max_iter - it);
    for (int c = 0; c < chunk; c++) {
      const int par = (it + c) & 1;
      plan_apply(pl, dtype, v.p, v.ap, 1, masked, st, flag);
      hipLaunchKernelGGL((k_dot<T>), grid, blk, 0, st, n, v.p, v.ap, v.part + 2 * kCgBlocks, flag);
      hipLaunchKernelGGL((k_update_xr<T>), grid, blk, 0, st, n, x, v.r, v.p, v.ap, v.part + 2 * kCgBlocks, v.scal,
                         par, v.part, flag);
      hipLaunchKernelGGL((k_update_p<T>), grid, blk, 0, st, n, v.p, v.r, v.part, v.scal, par, tol, flag);
      HIP_OK(hipGetLastError());
    }
    it += chunk;
    HIP_OK(hipMemcpyAsync(host, flag, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (host[0]) break;
  }
  *matvecs += host[1];
  return host[1];
}

template <typename T>
T read_scalar(const T *dev, int slot, hipStream_t st) {
  T v;
  HIP_OK(hipMemcpyAsync(&v, dev + slot, sizeof(T), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  return v;
}

void ensure_workspace(dkt_plan *pl, bool mixed) {
  CgWorkspace &w = pl->cg;
  const size_t n = pl->N;
  if (w.n != n) {
    w.release();
    w.n = n;
  }
  if (!w.r.p) {
    w.r.alloc(n);
    w.p.alloc(n);
    w.ap.alloc(n);
    w.part64.alloc(4 * kCgBlocks);
    HIP_OK(hipMemset(w.part64.p, 0, 4 * kCgBlocks * sizeof(double)));
    w.scal.alloc(S_COUNT);
    w.flag.alloc(4);
    HIP_OK(hipHostMalloc((void **)&w.host, 4 * sizeof(int), hipHostMallocDefault));
  }
  if (mixed && !w.r32.p) {
    w.d64.alloc(n);
    w.r32.alloc(n);
    w.p32.alloc(n);
    w.ap32.alloc(n);
    w.x32.alloc(n);
    w.b32.alloc(n);
    w.part32.alloc(4 * kCgBlocks);
    HIP_OK(hipMemset(w.part32.p, 0, 4 * kCgBlocks * sizeof(double)));
    w.scal32.alloc(S_COUNT);
  }
}

// Mixed-precision CG with reliable updates (see the file comment).  Returns
// the FP32 iterations; *resid is the final true relative residual.
int cg_reliable(dkt_plan *pl, const double *b, double *x, int max_iter, double tol, int masked, hipStream_t st,
                int *matvecs, int *updates, double *resid) {
  CgWorkspace &w = pl->cg;
  const int n = (int)pl->N;
  const dim3 grid(cg_blocks(n)), blk(kCgThreads);
  CgVectors<float> v{w.r32.p, w.p32.p, w.ap32.p, w.part32.p, w.scal32.p};
  int *flag = w.flag.p, *host = w.host;
  // True FP64 residual of the current x into d64 and scal[S_RESID].
  auto true_residual = [&]() {
    plan_apply(pl, 0, x, w.ap.p, 1, masked, st, nullptr);
    (*matvecs)++;
    hipLaunchKernelGGL(k_resid64, grid, blk, 0, st, n, b, w.ap.p, w.d64.p, w.part64.p);
    hipLaunchKernelGGL(k_resid64_fin, dim3(1), blk, 0, st, w.part64.p, w.scal.p);
    HIP_OK(hipGetLastError());
    return read_scalar(w.scal.p, S_RESID, st);
  };
  *resid = true_residual();
  if (*resid <= tol || max_iter == 0 || !std::isfinite(*resid)) return 0;
  // Start: r32 = p32 = r, x32 = 0 (a restart; the r.r goes to slot 0, flags cleared).
  hipLaunchKernelGGL(k_rel32, grid, blk, 0, st, n, w.d64.p, v.r, v.p, w.x32.p, 2, v.part);
  hipLaunchKernelGGL(k_rel32_fin, dim3(1), blk, 0, st, v.part, v.scal, w.scal.p, 1, 2, flag);
  HIP_OK(hipGetLastError());
  double last = *resid;
  int it = 0, stalls = 0;
  double tol_in = std::max(tol, kReliableDrop * last);
  while (it < max_iter) {
    const int chunk = std::min(kCgChunk, max_iter - it);
    for (int c = 0; c < chunk; c++) {
      const int par = (it + c) & 1;
      plan_apply(pl, 1, v.p, v.ap, 1, masked, st, flag);
      hipLaunchKernelGGL((k_dot<float>), grid, blk, 0, st, n, v.p, v.ap, v.part + 2 * kCgBlocks, flag);
      hipLaunchKernelGGL((k_update_xr<float>), grid, blk, 0, st, n, w.x32.p, v.r, v.p, v.ap, v.part + 2 * kCgBlocks,
                         v.scal, par, v.part, flag);
      hipLaunchKernelGGL((k_update_p<float>), grid, blk, 0, st, n, v.p, v.r, v.part, v.scal, par, tol_in, flag);
      HIP_OK(hipGetLastError());
    }
    HIP_OK(hipMemcpyAsync(host, flag, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    it = host[1];  // iterations completed (the chunk's launches after a stop return at once)
    if (!host[0] && it < max_iter) continue;
    // Reliable update: fold the FP32 update into x, replace the residual.
    hipLaunchKernelGGL(k_add32u, grid, blk, 0, st, n, x, w.x32.p);
    *resid = true_residual();
    (*updates)++;
    if (*resid <= tol || it >= max_iter || !std::isfinite(*resid)) break;
    // No progress since the last replacement: restart the search direction.
    stalls = *resid < 0.9 * last ? 0 : stalls + 1;
    const int restart = stalls > 0 ? 1 : 0;
    if (stalls > 8) break;
    const int par = (it - 1) & 1;  // parity of the iteration that stopped
    hipLaunchKernelGGL(k_rel32, grid, blk, 0, st, n, w.d64.p, v.r, v.p, w.x32.p, restart, v.part);
    hipLaunchKernelGGL(k_rel32_fin, dim3(1), blk, 0, st, v.part, v.scal, w.scal.p, par, restart, flag);
    if (!restart) hipLaunchKernelGGL(k_pupd32, grid, blk, 0, st, n, v.p, v.r, v.scal);
    HIP_OK(hipGetLastError());
    last = std::min(last, *resid);
    tol_in = std::max(tol, kReliableDrop * *resid);
  }
  *matvecs += it;
  return it;
}

// One solve of a batch (dkt_cg_engine, dkt_cg_solve_batch): its vectors,
// workspace and stop rule.
struct CgDesc {
  int n, max_iter;
  double tol;
  const double *b;
  double *x, *r, *p, *ap, *part, *scal;
  int *flag;
};

// Batched FP64 iteration kernels: solve blockIdx.y, block blockIdx.x of its
// kCgBlocks, through the single-solve bodies.
// Blocks at or past cg_blocks(n) of a solve return at once (the grid spans the
// largest solve).  kb_init's block 0 zeroes the partial slots past them: an
// engine slot's workspace is reused by solves of other sizes.
__global__ void __launch_bounds__(kCgThreads) kb_init(const CgDesc *__restrict__ d) {
  const CgDesc e = d[blockIdx.y];
  const int nb = cg_blocks(e.n);
  if (blockIdx.x >= nb) return;
  if (blockIdx.x == 0)
    for (int j = threadIdx.x; j < 4 * kCgBlocks; j += kCgThreads)
      if (j % kCgBlocks >= nb) e.part[j] = 0.0;
  CG_LOOP(i, e.n) e.x[i] = 0.0;
  init_body<double>(e.n, e.b, nullptr, e.r, e.p, e.part);
}
// As k_init_fin; a solve with max_iter = 0 stops at once (flag[0] = 2).
__global__ void __launch_bounds__(kCgThreads) kb_init_fin(const CgDesc *__restrict__ d) {
  const CgDesc e = d[blockIdx.y];
  init_fin_body(e.part, e.scal, e.tol, e.flag);
  if (threadIdx.x == 0 && e.max_iter <= 0 && !e.flag[0]) e.flag[0] = 2;
}
__global__ void __launch_bounds__(kCgThreads) kb_dot(const CgDesc *__restrict__ d) {
  const CgDesc e = d[blockIdx.y];
  if (blockIdx.x >= cg_blocks(e.n)) return;
  dot_body<double>(e.n, e.p, e.ap, e.part + 2 * kCgBlocks, e.flag);
}
__global__ void __launch_bounds__(kCgThreads) kb_update_xr(const CgDesc *__restrict__ d, int par) {
  const CgDesc e = d[blockIdx.y];
  if (blockIdx.x >= cg_blocks(e.n)) return;
  update_xr_body<double>(e.n, e.x, e.r, e.p, e.ap, e.part + 2 * kCgBlocks, e.scal, par, e.part, e.flag);
}
// As k_update_p; after the solve's max_iter-th iteration it stops (flag[0] =
// 2), as the single solve's host loop does.  Blocks that start after block 0
// has raised the flag skip the update of p, which is no longer used.
__global__ void __launch_bounds__(kCgThreads) kb_update_p(const CgDesc *__restrict__ d, int par) {
  const CgDesc e = d[blockIdx.y];
  if (blockIdx.x >= cg_blocks(e.n) || e.flag[0]) return;
  update_p_body<double>(e.n, e.p, e.r, e.part, e.scal, par, e.tol, e.flag);
  if (blockIdx.x == 0 && threadIdx.x == 0 && !e.flag[0] && e.flag[1] >= e.max_iter) e.flag[0] = 2;
}
// out[3k .. 3k + 2] = stop flag (0 running, 1 converged, 2 max_iter),
// iterations, relative residual of solve k.
__global__ void __launch_bounds__(kCgThreads) kb_status(const CgDesc *__restrict__ d, int nsolves,
                                                        double *__restrict__ out) {
  for (int k = threadIdx.x; k < nsolves; k += blockDim.x) {
    out[3 * k] = d[k].flag[0];
    out[3 * k + 1] = d[k].flag[1];
    out[3 * k + 2] = d[k].scal[S_RESID];
  }
}

}  // namespace
}  // namespace dkt

using namespace dkt;

extern "C" int dkt_cg_solve(dkt_plan *pl, int mode, const double *b, double *x, int max_iter, double tol, int masked,
                            double *stats, void *stream) {
  return guarded([&] {
    if (!pl || !b || !x) throw std::invalid_argument("null pointer");
    if (pl->dry) throw std::invalid_argument("a dry-run plan cannot solve");
    if (max_iter < 0 || !(tol >= 0.0)) throw std::invalid_argument("max_iter >= 0 and tol >= 0 required");
    if (mode < DKT_CG_F64 || mode > DKT_CG_MIXED_RELIABLE) throw std::invalid_argument("unknown CG mode");
    // DKT_CG_MIXED: reliable updates on the symmetric (conforming) operator;
    // restarted refinement on the reference traversal's operator, which is not
    // symmetric at hanging nodes and on which restarts converge faster.
    if (mode == DKT_CG_MIXED) mode = pl->conforming ? DKT_CG_MIXED_RELIABLE : DKT_CG_MIXED_RESTART;
    if ((const void *)b == (const void *)x) throw std::invalid_argument("b and x must not alias");
    hipStream_t st = (hipStream_t)stream;
    const bool mixed = mode != DKT_CG_F64;
    ensure_workspace(pl, mixed);
    CgWorkspace &w = pl->cg;
    int matvecs = 0, iters = 0, outer = 0;
    double resid = 0.0;
    if (!mixed) {
      CgVectors<double> v{w.r.p, w.p.p, w.ap.p, w.part64.p, w.scal.p};
      iters = cg_run<double>(pl, 0, b, x, false, max_iter, tol, masked ? 1 : 0, v, w.flag.p, w.host, st, &matvecs);
      resid = read_scalar(w.scal.p, S_RESID, st);
    } else if (mode == DKT_CG_MIXED_RELIABLE) {
      iters = cg_reliable(pl, b, x, max_iter, tol, masked ? 1 : 0, st, &matvecs, &outer, &resid);
    } else {
      const int n = (int)pl->N;
      CgVectors<float> v{w.r32.p, w.p32.p, w.ap32.p, w.part32.p, w.scal32.p};  // FP64 reductions
      while (true) {
        plan_apply(pl, 0, x, w.ap.p, 1, masked ? 1 : 0, st, nullptr);
        matvecs++;
        hipLaunchKernelGGL(k_resid64, dim3(cg_blocks(n)), dim3(kCgThreads), 0, st, n, b, w.ap.p, w.d64.p, w.part64.p);
        hipLaunchKernelGGL(k_resid64_fin, dim3(1), dim3(kCgThreads), 0, st, w.part64.p, w.scal.p);
        HIP_OK(hipGetLastError());
        resid = read_scalar(w.scal.p, S_RESID, st);
        if (resid <= tol || iters >= max_iter || !std::isfinite(resid)) break;
        hipLaunchKernelGGL(k_to32, dim3(cg_blocks(n)), dim3(kCgThreads), 0, st, n, w.d64.p, w.scal.p, w.b32.p, w.x32.p);
        // Inner tolerance: kInnerTol, loosened when the outer target is already close.
        const double inner_tol = std::min(0.5, std::max(kInnerTol, 0.5 * tol / resid));
        const int inner = cg_run<float>(pl, 1, w.b32.p, w.x32.p, true, max_iter - iters, inner_tol, masked ? 1 : 0, v,
                                        w.flag.p, w.host, st, &matvecs);
        hipLaunchKernelGGL(k_add32, dim3(cg_blocks(n)), dim3(kCgThreads), 0, st, n, x, w.x32.p, w.scal.p);
        HIP_OK(hipGetLastError());
        iters += inner;
        outer++;
        if (inner == 0 || outer >= kMaxOuter) break;  // no progress possible: report the last residual
      }
    }
    if (stats) {
      stats[0] = iters;
      stats[1] = resid;
      stats[2] = outer;
      stats[3] = matvecs;
    }
  });
}

// Continuous batching of FP64 CG solves.  Solves join (dkt_cg_engine_add)
// and leave (finished in dkt_cg_engine_step) between chunks of kCgChunk
// iterations; every chunk runs one matvec launch per (dim, p) of the active
// plans and one launch per vector step for all active solves.  A solve starts
// at the first step after it was added, so it always runs an even number of
// iterations per chunk and its own iteration parity: its arithmetic is that of
// its own dkt_cg_solve from x = 0.
struct dkt_cg_engine {
  // A slot owns its CG workspace (kept and grown across solves), so several
  // solves may run on one plan at once.
  struct Slot {
    dkt_plan *pl = nullptr;
    CgDesc d{};
    bool active = false, fresh = false;
    size_t cap = 0;
    DevBuf<double> r, p, ap, part, scal;
    DevBuf<int> flag;
  };
  int masked = 0;
  std::deque<Slot> slots;  // fixed size after creation; Slot is not movable
  bool dirty = true;                       // active set changed: rebuild the launches
  std::vector<int> act;                    // active slots in descriptor order
  std::unique_ptr<MatvecLauncher> mv;
  DevBuf<CgDesc> dact, dnew;
  DevBuf<double> status;
  std::vector<CgDesc> hact, hnew;
  std::vector<double> hs;
};

extern "C" int dkt_cg_engine_create(int max_slots, int masked, dkt_cg_engine **out) {
  return guarded([&] {
    if (!out || max_slots < 1 || max_slots > 65535) throw std::invalid_argument("max_slots must be 1..65535");
    auto e = std::make_unique<dkt_cg_engine>();
    e->masked = masked ? 1 : 0;
    e->slots.resize(max_slots);  // device memory is allocated by add and step
    *out = e.release();
  });
}

extern "C" int dkt_cg_engine_add(dkt_cg_engine *e, dkt_plan *pl, const double *b, double *x, int max_iter, double tol,
                                 int *slot) {
  return guarded([&] {
    if (!e || !pl || !b || !x || !slot) throw std::invalid_argument("null pointer");
    if (pl->dry) throw std::invalid_argument("a dry-run plan cannot solve");
    if (pl->simple || pl->per_level || pl->conforming)
      throw std::invalid_argument("batched solves need tiled device plans");
    if (max_iter < 0 || !(tol >= 0.0)) throw std::invalid_argument("max_iter >= 0 and tol >= 0 required");
    if ((const void *)b == (const void *)x) throw std::invalid_argument("b and x must not alias");
    int free_slot = -1;
    for (int k = 0; k < (int)e->slots.size() && free_slot < 0; k++)
      if (!e->slots[k].active) free_slot = k;
    if (free_slot < 0) throw std::invalid_argument("no free slot");
    auto &s = e->slots[free_slot];
    if (s.cap < pl->N) {
      s.r.alloc(pl->N);
      s.p.alloc(pl->N);
      s.ap.alloc(pl->N);
      s.cap = pl->N;
    }
    if (!s.part.p) {
      s.part.alloc(4 * kCgBlocks);
      s.scal.alloc(S_COUNT);
      s.flag.alloc(4);
    }
    s.pl = pl;
    s.d = CgDesc{(int)pl->N, max_iter, tol, b, x, s.r.p, s.p.p, s.ap.p, s.part.p, s.scal.p, s.flag.p};
    s.active = true;
    s.fresh = true;
    e->dirty = true;
    *slot = free_slot;
  });
}

// Starts the solves added since the last step, runs `chunks` chunks of
// iterations over the active solves and reports those that stopped:
// finished[i] = slot, stats[2i] = iterations, stats[2i + 1] = relative
// residual, *nfinished = their count (arrays of at least max_slots entries).
// Their slots are free again.
extern "C" int dkt_cg_engine_step(dkt_cg_engine *e, int chunks, void *stream, int *finished, double *stats,
                                  int *nfinished) {
  return guarded([&] {
    if (!e || !finished || !nfinished || chunks < 0) throw std::invalid_argument("null pointer or chunks < 0");
    hipStream_t st = (hipStream_t)stream;
    *nfinished = 0;
    if (e->dirty) {
      e->act.clear();
      e->hact.clear();
      e->hnew.clear();
      std::vector<dkt_plan *> pls;
      std::vector<const double *> px;
      std::vector<double *> py;
      std::vector<const int *> cancel;
      for (int k = 0; k < (int)e->slots.size(); k++) {
        auto &s = e->slots[k];
        if (!s.active) continue;
        e->act.push_back(k);
        e->hact.push_back(s.d);
        if (s.fresh) e->hnew.push_back(s.d);
        s.fresh = false;
        pls.push_back(s.pl);
        px.push_back(s.d.p);
        py.push_back(s.d.ap);
        cancel.push_back(s.d.flag);  // flag[0]: stopped
      }
      e->mv = pls.empty() ? nullptr : batch_matvec_f64(pls, px, py, e->masked, cancel);
      if (e->dact.n < e->hact.size()) e->dact.alloc(e->hact.size());
      if (e->dnew.n < e->hnew.size()) e->dnew.alloc(e->hnew.size());
      if (!e->hact.empty())
        HIP_OK(hipMemcpyAsync(e->dact.p, e->hact.data(), e->hact.size() * sizeof(CgDesc), hipMemcpyHostToDevice,
                              st));
      if (!e->hnew.empty()) {
        HIP_OK(hipMemcpyAsync(e->dnew.p, e->hnew.data(), e->hnew.size() * sizeof(CgDesc), hipMemcpyHostToDevice,
                              st));
        const int nn = (int)e->hnew.size();
        int nbn = 1;
        for (const CgDesc &c : e->hnew) nbn = std::max(nbn, cg_blocks(c.n));
        hipLaunchKernelGGL(kb_init, dim3(nbn, nn), dim3(kCgThreads), 0, st, e->dnew.p);
        hipLaunchKernelGGL(kb_init_fin, dim3(1, nn), dim3(kCgThreads), 0, st, e->dnew.p);
        HIP_OK(hipGetLastError());
      }
      e->dirty = false;
    }
    const int K = (int)e->act.size();
    if (K == 0) return;
    if (e->status.n < 3 * e->slots.size()) e->status.alloc(3 * e->slots.size());
    int nb = 1;
    for (const CgDesc &c : e->hact) nb = std::max(nb, cg_blocks(c.n));
    const dim3 grid(nb, K), blk(kCgThreads);
    for (int c = 0; c < chunks * kCgChunk; c++) {
      const int par = c & 1;
      e->mv->launch(st);
      hipLaunchKernelGGL(kb_dot, grid, blk, 0, st, e->dact.p);
      hipLaunchKernelGGL(kb_update_xr, grid, blk, 0, st, e->dact.p, par);
      hipLaunchKernelGGL(kb_update_p, grid, blk, 0, st, e->dact.p, par);
      HIP_OK(hipGetLastError());
    }
    hipLaunchKernelGGL(kb_status, dim3(1), dim3(kCgThreads), 0, st, e->dact.p, K, e->status.p);
    HIP_OK(hipGetLastError());
    e->hs.resize(3 * (size_t)K);
    HIP_OK(hipMemcpyAsync(e->hs.data(), e->status.p, e->hs.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int i = 0; i < K; i++) {
      if (e->hs[3 * i] == 0.0) continue;
      const int k = e->act[i];
      finished[*nfinished] = k;
      if (stats) {
        stats[2 * *nfinished] = e->hs[3 * i + 1];
        stats[2 * *nfinished + 1] = e->hs[3 * i + 2];
      }
      (*nfinished)++;
      e->slots[k].active = false;
      e->dirty = true;
    }
  });
}

extern "C" int dkt_cg_engine_destroy(dkt_cg_engine *e) {
  delete e;
  return 0;
}

extern "C" int dkt_cg_solve_batch(dkt_plan *const *plans, int nplans, const double *const *b, double *const *x,
                                  int max_iter, double tol, int masked, double *stats, void *stream) {
  return guarded([&] {
    if (!plans || !b || !x || nplans < 1) throw std::invalid_argument("null pointer or empty batch");
    dkt_cg_engine *raw = nullptr;
    if (dkt_cg_engine_create(nplans, masked, &raw) != 0) throw std::runtime_error(dkt_last_error());
    std::unique_ptr<dkt_cg_engine> e(raw);
    for (int k = 0; k < nplans; k++) {
      int slot = -1;
      if (dkt_cg_engine_add(e.get(), plans[k], b[k], x[k], max_iter, tol, &slot) != 0)
        throw std::invalid_argument(dkt_last_error());
    }
    std::vector<int> fin(nplans);
    std::vector<double> fst(2 * (size_t)nplans);
    int left = nplans;
    // The first step only starts the solves (they stop at once when b = 0 or max_iter = 0).
    for (int chunks = 0; left > 0; chunks = 1) {
      int nf = 0;
      if (dkt_cg_engine_step(e.get(), chunks, stream, fin.data(), fst.data(), &nf) != 0)
        throw std::runtime_error(dkt_last_error());
      for (int i = 0; i < nf; i++)
        if (stats) {
          stats[2 * fin[i]] = fst[2 * i];
          stats[2 * fin[i] + 1] = fst[2 * i + 1];
        }
      left -= nf;
    }
  });
}
