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
enum { S_RR0 = 0, S_RR1 = 1, S_NORMB = 2, S_RESID = 3, S_SCALE = 4, S_BETA = 5, S_PAR = 6, S_TOLIN = 7, S_COUNT = 8 };
// Stop reasons in flag[0] (0: running), as reported in the solve statistics.
constexpr int kStopConverged = 1, kStopMaxIter = 2, kStopBreakdown = 3, kStopStalled = 4;

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

// The iteration kernels' loops take kCgUnroll of a work-item's entries at a
// time with every load issued before any use (one memory round trip per batch
// where the plain loop waits once per entry), then handle them in index order:
// each work-item's sums keep the order, and the bits, of the plain loop.
constexpr int kCgUnroll = 4;

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
  const int stride = cg_blocks(n) * kCgThreads;
  int i = blockIdx.x * kCgThreads + threadIdx.x;
  for (; i < n - (kCgUnroll - 1) * stride; i += kCgUnroll * stride) {
    T av[kCgUnroll], bv[kCgUnroll];
#pragma unroll
    for (int u = 0; u < kCgUnroll; u++) {
      av[u] = a[i + u * stride];
      bv[u] = b[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < kCgUnroll; u++) acc += (double)av[u] * (double)bv[u];
  }
  for (; i < n; i += stride) acc += (double)a[i] * (double)b[i];
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
                                               int *__restrict__ flag) {
  if (flag[0]) return;
  __shared__ double sh[kCgThreads];
  // Breakdown guard: p.Ap <= 0 (the operator is not positive definite on p)
  // or a non-finite step stops the solve before x or r change.  Every block
  // reduces the same partials in the same order, so all of them agree.
  const double pap = parts_sum(part_pap, sh);
  const double a64 = scal[S_RR0 + par] / pap;
  if (!(pap > 0.0) || !isfinite(a64)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = kStopBreakdown;
    return;
  }
  const T alpha = (T)a64;
  double rr = 0, mr = 0;
  const int stride = cg_blocks(n) * kCgThreads;
  int i = blockIdx.x * kCgThreads + threadIdx.x;
  for (; i < n - (kCgUnroll - 1) * stride; i += kCgUnroll * stride) {
    T xv[kCgUnroll], pv[kCgUnroll], rv[kCgUnroll], av[kCgUnroll];
#pragma unroll
    for (int u = 0; u < kCgUnroll; u++) {
      const int k = i + u * stride;
      xv[u] = x[k];
      pv[u] = p[k];
      rv[u] = r[k];
      av[u] = ap[k];
    }
#pragma unroll
    for (int u = 0; u < kCgUnroll; u++) {
      const int k = i + u * stride;
      x[k] = xv[u] + alpha * pv[u];
      const T ri = rv[u] - alpha * av[u];
      r[k] = ri;
      rr += (double)ri * (double)ri;
      mr = fmax(mr, (double)fabs(ri));
    }
  }
  for (; i < n; i += stride) {
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
                                                          double *__restrict__ part, int *__restrict__ flag) {
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
    const int stride = cg_blocks(n) * kCgThreads;
    int i = blockIdx.x * kCgThreads + threadIdx.x;
    for (; i < n - (kCgUnroll - 1) * stride; i += kCgUnroll * stride) {
      T rv[kCgUnroll], pv[kCgUnroll];
#pragma unroll
      for (int u = 0; u < kCgUnroll; u++) {
        rv[u] = r[i + u * stride];
        pv[u] = p[i + u * stride];
      }
#pragma unroll
      for (int u = 0; u < kCgUnroll; u++) p[i + u * stride] = rv[u] + beta * pv[u];
    }
    for (; i < n; i += stride) p[i] = r[i] + beta * p[i];
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
// Returns the iterations performed; the final relative residual is scal[S_RESID]
// and *stop the stop reason (kStop*).
template <typename T>
int cg_run(dkt_plan *pl, int dtype, const T *b, T *x, bool x_is_zero, int max_iter, double tol, int masked,
           const CgVectors<T> &v, int *flag, int *host, hipStream_t st, int *matvecs, int *stop) {
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
  if (host[0]) {
    *stop = kStopConverged;
    return 0;
  }
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
  *stop = host[0] ? host[0] : kStopMaxIter;
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
  }
  if (!w.flag.p) {
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
                int *matvecs, int *updates, double *resid, int *stop) {
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
  *stop = *resid <= tol ? kStopConverged : (std::isfinite(*resid) ? kStopMaxIter : kStopBreakdown);
  if (*resid <= tol || max_iter == 0 || !std::isfinite(*resid)) return 0;
  // Start: r32 = p32 = r, x32 = 0 (a restart; the r.r goes to slot 0, flags cleared).
  hipLaunchKernelGGL(k_rel32, grid, blk, 0, st, n, w.d64.p, v.r, v.p, w.x32.p, 2, v.part);
  hipLaunchKernelGGL(k_rel32_fin, dim3(1), blk, 0, st, v.part, v.scal, w.scal.p, 1, 2, flag);
  HIP_OK(hipGetLastError());
  double last = *resid;
  int it = 0, stalls = 0, it_restart = 0;
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
    const bool broke = host[0] == kStopBreakdown;
    if (!host[0] && it < max_iter) continue;
    // Reliable update: fold the FP32 update into x, replace the residual.
    hipLaunchKernelGGL(k_add32u, grid, blk, 0, st, n, x, w.x32.p);
    *resid = true_residual();
    (*updates)++;
    if (*resid <= tol) {
      *stop = kStopConverged;
      break;
    }
    if (!std::isfinite(*resid) || (broke && it == it_restart)) {
      *stop = kStopBreakdown;  // non-finite, or p.Ap <= 0 right after a restart
      break;
    }
    if (it >= max_iter) {
      *stop = kStopMaxIter;
      break;
    }
    // No progress since the last replacement, or a breakdown of the FP32
    // iteration: restart the search direction.
    stalls = *resid < 0.9 * last ? 0 : stalls + 1;
    const int restart = (stalls > 0 || broke) ? 1 : 0;
    if (stalls > 8) {
      *stop = kStopStalled;
      break;
    }
    if (restart) it_restart = it;
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

// One solve of a batch (dkt_cg_engine, dkt_cg_solve_batch): its iteration
// vectors (FP64, or FP32 for the iteration of a mixed-precision solve),
// workspace and stop rule.  tolp non-null: the iteration stops at the
// relative residual *tolp (a mixed solve's inner target) instead of tol.
template <typename T>
struct CgDesc {
  int n, max_iter;
  double tol;
  const double *tolp;
  const T *b;
  T *x, *r, *p, *ap;
  double *part, *scal;
  int *flag;
};

// Batched iteration kernels: solve blockIdx.y, block blockIdx.x of its
// kCgBlocks, through the single-solve bodies.
// Blocks at or past cg_blocks(n) of a solve return at once (the grid spans the
// largest solve).  kb_init's block 0 zeroes the partial slots past them: an
// engine slot's workspace is reused by solves of other sizes.
__global__ void __launch_bounds__(kCgThreads) kb_init(const CgDesc<double> *__restrict__ d) {
  const CgDesc<double> e = d[blockIdx.y];
  const int nb = cg_blocks(e.n);
  if (blockIdx.x >= nb) return;
  if (blockIdx.x == 0)
    for (int j = threadIdx.x; j < 4 * kCgBlocks; j += kCgThreads)
      if (j % kCgBlocks >= nb) e.part[j] = 0.0;
  CG_LOOP(i, e.n) e.x[i] = 0.0;
  init_body<double>(e.n, e.b, nullptr, e.r, e.p, e.part);
}
// As k_init_fin; a solve with max_iter = 0 stops at once (flag[0] = 2).
__global__ void __launch_bounds__(kCgThreads) kb_init_fin(const CgDesc<double> *__restrict__ d) {
  const CgDesc<double> e = d[blockIdx.y];
  init_fin_body(e.part, e.scal, e.tol, e.flag);
  if (threadIdx.x == 0 && e.max_iter <= 0 && !e.flag[0]) e.flag[0] = kStopMaxIter;
}
template <typename T>
__global__ void __launch_bounds__(kCgThreads) kb_dot(const CgDesc<T> *__restrict__ d) {
  const CgDesc<T> e = d[blockIdx.y];
  if (blockIdx.x >= cg_blocks(e.n)) return;
  dot_body<T>(e.n, e.p, e.ap, e.part + 2 * kCgBlocks, e.flag);
}
template <typename T>
__global__ void __launch_bounds__(kCgThreads) kb_update_xr(const CgDesc<T> *__restrict__ d, int par) {
  const CgDesc<T> e = d[blockIdx.y];
  if (blockIdx.x >= cg_blocks(e.n)) return;
  update_xr_body<T>(e.n, e.x, e.r, e.p, e.ap, e.part + 2 * kCgBlocks, e.scal, par, e.part, e.flag);
}
// As k_update_p; after the solve's max_iter-th iteration it stops (flag[0] =
// 2), as the single solve's host loop does.  Blocks that start after block 0
// has raised the flag skip the update of p, which is no longer used.  The
// parity of the iteration that stopped goes to scal[S_PAR] (a mixed solve's
// residual replacement continues from it).
template <typename T>
__global__ void __launch_bounds__(kCgThreads) kb_update_p(const CgDesc<T> *__restrict__ d, int par) {
  const CgDesc<T> e = d[blockIdx.y];
  if (blockIdx.x >= cg_blocks(e.n) || e.flag[0]) return;
  update_p_body<T>(e.n, e.p, e.r, e.part, e.scal, par, e.tolp ? *e.tolp : e.tol, e.flag);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (!e.flag[0] && e.flag[1] >= e.max_iter) e.flag[0] = kStopMaxIter;
    if (e.flag[0]) e.scal[S_PAR] = par;
  }
}

// Residual replacement of mixed-precision solves in an engine (the device
// side of cg_reliable's host loop, for every slot at once).  upd: 0 nothing
// to do this step, 1 the FP32 iteration stopped (inner target, max_iter or
// breakdown): fold x32 into x and replace the residual, 2 the solve starts
// (x = 0, r = b).  st: [0] lowest true residual so far, [1] stalls,
// [2] FP32 iterations at the last restart, [3] restart (1) or keep the
// search direction (0).  A stopped solve's flag[0] becomes kFinal + reason.
constexpr int kFinal = 16;
struct RelDesc {
  int n, max_iter;
  double tol;
  const double *b;
  double *x, *r64, *ap64, *part64, *scal64;
  float *x32, *r32, *p32;
  double *part32, *scal32;
  double *st;
  int *flag;
  const int *upd;
};

// x += x32, x32 = 0 (start: x = 0, x32 = 0).
__global__ void __launch_bounds__(kCgThreads) kr_fold(const RelDesc *__restrict__ d) {
  const RelDesc e = d[blockIdx.y];
  const int u = *e.upd;
  if (!u || blockIdx.x >= cg_blocks(e.n)) return;
  CG_LOOP(i, e.n) {
    e.x[i] = u == 2 ? 0.0 : e.x[i] + (double)e.x32[i];
    e.x32[i] = 0.0f;
  }
}
// r64 = b - A x (start: r64 = b); partials of max|r|, max|b|.  At the start,
// block 0 zeroes the partial slots past cg_blocks(n) of both workspaces.
__global__ void __launch_bounds__(kCgThreads) kr_resid(const RelDesc *__restrict__ d) {
  const RelDesc e = d[blockIdx.y];
  const int u = *e.upd;
  const int nb = cg_blocks(e.n);
  if (!u || blockIdx.x >= nb) return;
  if (u == 2 && blockIdx.x == 0)
    for (int j = threadIdx.x; j < 4 * kCgBlocks; j += kCgThreads)
      if (j % kCgBlocks >= nb) {
        e.part64[j] = 0.0;
        e.part32[j] = 0.0;
      }
  __shared__ double sh[kCgThreads];
  double mr = 0, mb = 0;
  CG_LOOP(i, e.n) {
    const double ri = u == 2 ? e.b[i] : e.b[i] - e.ap64[i];
    e.r64[i] = ri;
    mr = fmax(mr, fabs(ri));
    mb = fmax(mb, fabs(e.b[i]));
  }
  mr = block_max(mr, sh);
  mb = block_max(mb, sh);
  if (threadIdx.x == 0) {
    e.part64[blockIdx.x] = mr;
    e.part64[kCgBlocks + blockIdx.x] = mb;
  }
}
// The stop decision of cg_reliable, per solve.
__global__ void __launch_bounds__(kCgThreads) kr_decide(const RelDesc *__restrict__ d) {
  const RelDesc e = d[blockIdx.y];
  const int u = *e.upd;
  if (!u) return;
  __shared__ double sh[kCgThreads];
  const double mr = parts_max(e.part64, sh);
  double mb = parts_max(e.part64 + kCgBlocks, sh);
  if (threadIdx.x != 0) return;
  if (mb == 0.0) mb = 1.0;
  const double resid = mr / mb;
  e.scal64[S_NORMB] = mb;
  e.scal64[S_RESID] = resid;
  double *st = e.st;
  if (u == 2) {  // start
    e.flag[1] = 0;
    st[0] = resid;
    st[1] = 0;
    st[2] = 0;
    st[3] = 1;
    if (resid <= e.tol) e.flag[0] = kFinal + kStopConverged;
    else if (!isfinite(resid)) e.flag[0] = kFinal + kStopBreakdown;
    else if (e.max_iter <= 0) e.flag[0] = kFinal + kStopMaxIter;
    else e.flag[0] = 0;
    e.scal32[S_TOLIN] = fmax(e.tol, kReliableDrop * resid);
    return;
  }
  const int it = e.flag[1];
  const bool broke = e.flag[0] == kStopBreakdown;
  if (resid <= e.tol) {
    e.flag[0] = kFinal + kStopConverged;
  } else if (!isfinite(resid) || (broke && it == (int)st[2])) {
    e.flag[0] = kFinal + kStopBreakdown;
  } else if (it >= e.max_iter) {
    e.flag[0] = kFinal + kStopMaxIter;
  } else {
    const double last = st[0];
    st[1] = resid < 0.9 * last ? 0.0 : st[1] + 1.0;
    const int restart = (st[1] > 0 || broke) ? 1 : 0;
    if (st[1] > 8) {
      e.flag[0] = kFinal + kStopStalled;
    } else {
      if (restart) st[2] = it;
      st[3] = restart;
      st[0] = fmin(last, resid);
      e.scal32[S_TOLIN] = fmax(e.tol, kReliableDrop * resid);
      e.flag[0] = 0;
    }
  }
}
// r32 = r64 (restart: p32 = r32 too); partials of r32.r32.
__global__ void __launch_bounds__(kCgThreads) kr_rel32(const RelDesc *__restrict__ d) {
  const RelDesc e = d[blockIdx.y];
  if (!*e.upd || e.flag[0] || blockIdx.x >= cg_blocks(e.n)) return;
  const bool restart = e.st[3] != 0.0;
  __shared__ double sh[kCgThreads];
  double rr = 0;
  CG_LOOP(i, e.n) {
    const float ri = (float)e.r64[i];
    e.r32[i] = ri;
    if (restart) e.p32[i] = ri;
    rr += (double)ri * (double)ri;
  }
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) e.part32[blockIdx.x] = rr;
}
// The new r.r goes to parity slot 0 (every engine step starts at parity 0);
// beta = new r.r / r.r before the iteration that stopped (0 on a restart).
__global__ void __launch_bounds__(kCgThreads) kr_rel32_fin(const RelDesc *__restrict__ d) {
  const RelDesc e = d[blockIdx.y];
  if (!*e.upd || e.flag[0]) return;
  __shared__ double sh[kCgThreads];
  const double rr = parts_sum(e.part32, sh);
  if (threadIdx.x == 0) {
    const int par = *e.upd == 2 ? 0 : (int)e.scal32[S_PAR];
    e.scal32[S_BETA] = e.st[3] != 0.0 ? 0.0 : rr / e.scal32[S_RR0 + par];
    e.scal32[S_RR0] = rr;
    e.scal32[S_NORMB] = e.scal64[S_NORMB];
  }
}
// p32 = r32 + beta p32 (residual replaced without a restart).
__global__ void __launch_bounds__(kCgThreads) kr_pupd32(const RelDesc *__restrict__ d) {
  const RelDesc e = d[blockIdx.y];
  if (!*e.upd || e.flag[0] || e.st[3] != 0.0 || blockIdx.x >= cg_blocks(e.n)) return;
  const float beta = (float)e.scal32[S_BETA];
  CG_LOOP(i, e.n) e.p32[i] = e.r32[i] + beta * e.p32[i];
}

// out[3k .. 3k + 2] = flag[0], iterations, relative residual of solve k.
struct StatDesc {
  const int *flag;
  const double *resid;
};
__global__ void __launch_bounds__(kCgThreads) kb_status(const StatDesc *__restrict__ d, int nsolves,
                                                        double *__restrict__ out) {
  for (int k = threadIdx.x; k < nsolves; k += blockDim.x) {
    out[3 * k] = d[k].flag[0];
    out[3 * k + 1] = d[k].flag[1];
    out[3 * k + 2] = *d[k].resid;
  }
}

// ---------------------------------------------------------------------------
// BiCGStab (van der Vorst 1992) for nonsymmetric operators (the stabilised
// space-time heat form), FP64, with the same fixed-order reductions as CG:
// every block re-reduces the same partials, so the solve is bitwise
// reproducible.  No counterpart in the reference's library (its heat solves
// go through PETSc's KSP, FEM/examples/src/heatEq.cpp:153-165).
// Scalars: B_RHO r^.r of the current residual, B_RHO0 of the previous one,
// B_ALPHA, B_OMEGA, B_NORMB, B_RESID.
// ---------------------------------------------------------------------------
enum { B_RHO = 0, B_RHO0 = 1, B_ALPHA = 2, B_OMEGA = 3, B_NORMB = 4, B_RESID = 5, B_COUNT = 8 };

struct BiVecs {
  double *r, *rh, *p, *v, *s, *t, *part, *scal;
};

// r = b - ax, r^ = r, p = v = 0; partials of r.r, max|r|, max|b|.
__global__ void __launch_bounds__(kCgThreads) kbi_init(int n, const double *__restrict__ b,
                                                       const double *__restrict__ ax, BiVecs w) {
  __shared__ double sh[kCgThreads];
  double rr = 0, mr = 0, mb = 0;
  CG_LOOP(i, n) {
    const double ri = b[i] - ax[i];
    w.r[i] = ri;
    w.rh[i] = ri;
    w.p[i] = 0.0;
    w.v[i] = 0.0;
    rr += ri * ri;
    mr = fmax(mr, fabs(ri));
    mb = fmax(mb, fabs(b[i]));
  }
  rr = block_sum(rr, sh);
  mr = block_max(mr, sh);
  mb = block_max(mb, sh);
  if (threadIdx.x == 0) {
    w.part[blockIdx.x] = rr;
    w.part[kCgBlocks + blockIdx.x] = mr;
    w.part[2 * kCgBlocks + blockIdx.x] = mb;
  }
}
__global__ void __launch_bounds__(kCgThreads) kbi_init_fin(BiVecs w, double tol, int *__restrict__ flag) {
  __shared__ double sh[kCgThreads];
  const double rr = parts_sum(w.part, sh);
  const double mr = parts_max(w.part + kCgBlocks, sh);
  double mb = parts_max(w.part + 2 * kCgBlocks, sh);
  if (threadIdx.x == 0) {
    if (mb == 0.0) mb = 1.0;
    w.scal[B_RHO] = rr;
    w.scal[B_RHO0] = 1.0;
    w.scal[B_ALPHA] = 1.0;
    w.scal[B_OMEGA] = 1.0;
    w.scal[B_NORMB] = mb;
    w.scal[B_RESID] = mr / mb;
    flag[0] = mr / mb <= tol ? kStopConverged : 0;
    flag[1] = 0;
  }
}
// p = r + beta (p - omega v), beta = (rho / rho0) (alpha / omega).
__global__ void __launch_bounds__(kCgThreads) kbi_p(int n, BiVecs w, const int *__restrict__ flag) {
  if (flag[0]) return;
  const double beta = (w.scal[B_RHO] / w.scal[B_RHO0]) * (w.scal[B_ALPHA] / w.scal[B_OMEGA]);
  const double om = w.scal[B_OMEGA];
  CG_LOOP(i, n) w.p[i] = w.r[i] + beta * (w.p[i] - om * w.v[i]);
}
// partials of a.b into part + slot * kCgBlocks (and c.d into the next slot when c is given).
__global__ void __launch_bounds__(kCgThreads) kbi_dot(int n, const double *__restrict__ a, const double *__restrict__ b,
                                                      const double *__restrict__ c, const double *__restrict__ d,
                                                      double *__restrict__ part, const int *__restrict__ flag) {
  if (flag[0]) return;
  __shared__ double sh[kCgThreads];
  double ab = 0, cd = 0;
  CG_LOOP(i, n) {
    ab += a[i] * b[i];
    if (c) cd += c[i] * d[i];
  }
  ab = block_sum(ab, sh);
  if (c) cd = block_sum(cd, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = ab;
    if (c) part[kCgBlocks + blockIdx.x] = cd;
  }
}
// alpha = rho / (r^.v); s = r - alpha v.  r^.v = 0 or a non-finite alpha: breakdown.
__global__ void __launch_bounds__(kCgThreads) kbi_s(int n, BiVecs w, int *__restrict__ flag) {
  if (flag[0]) return;
  __shared__ double sh[kCgThreads];
  const double rv = parts_sum(w.part + 3 * kCgBlocks, sh);
  const double alpha = w.scal[B_RHO] / rv;
  if (!(rv != 0.0) || !isfinite(alpha)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = kStopBreakdown;
    return;
  }
  CG_LOOP(i, n) w.s[i] = w.r[i] - alpha * w.v[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) w.scal[B_ALPHA] = alpha;
}
// omega = (t.s) / (t.t) (0 when t = 0); x += alpha p + omega s; r = s - omega t;
// partials of r^.r and max|r|.
__global__ void __launch_bounds__(kCgThreads) kbi_xr(int n, double *__restrict__ x, BiVecs w,
                                                     const int *__restrict__ flag) {
  if (flag[0]) return;
  __shared__ double sh[kCgThreads];
  const double ts = parts_sum(w.part + 4 * kCgBlocks, sh);
  const double tt = parts_sum(w.part + 5 * kCgBlocks, sh);
  const double om = tt > 0.0 ? ts / tt : 0.0;
  const double alpha = w.scal[B_ALPHA];
  double rr = 0, mr = 0;
  CG_LOOP(i, n) {
    x[i] += alpha * w.p[i] + om * w.s[i];
    const double ri = w.s[i] - om * w.t[i];
    w.r[i] = ri;
    rr += w.rh[i] * ri;
    mr = fmax(mr, fabs(ri));
  }
  rr = block_sum(rr, sh);
  mr = block_max(mr, sh);
  if (threadIdx.x == 0) {
    w.part[blockIdx.x] = rr;
    w.part[kCgBlocks + blockIdx.x] = mr;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) w.scal[B_OMEGA] = om;
}
// rho0 = rho, rho = r^.r, the residual and the stop test; omega = 0 or rho =
// 0 with the residual above tol: breakdown (the next beta is undefined).
__global__ void __launch_bounds__(kCgThreads) kbi_fin(BiVecs w, double tol, int max_iter, int *__restrict__ flag) {
  if (flag[0]) return;
  __shared__ double sh[kCgThreads];
  const double rho = parts_sum(w.part, sh);
  const double mr = parts_max(w.part + kCgBlocks, sh);
  if (threadIdx.x == 0) {
    const double resid = mr / w.scal[B_NORMB];
    w.scal[B_RHO0] = w.scal[B_RHO];
    w.scal[B_RHO] = rho;
    w.scal[B_RESID] = resid;
    flag[1] += 1;
    if (resid <= tol) flag[0] = kStopConverged;
    else if (!isfinite(resid) || rho == 0.0 || w.scal[B_OMEGA] == 0.0) flag[0] = kStopBreakdown;
    else if (flag[1] >= max_iter) flag[0] = kStopMaxIter;
  }
}

}  // namespace
}  // namespace dkt

using namespace dkt;

extern "C" int dkt_cg_solve(dkt_plan *pl, int mode, const double *b, double *x, int max_iter, double tol, int masked,
                            double *stats, void *stream) {
  int stop = 0;
  const int rc = guarded([&] {
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
      iters = cg_run<double>(pl, 0, b, x, false, max_iter, tol, masked ? 1 : 0, v, w.flag.p, w.host, st, &matvecs,
                             &stop);
      resid = read_scalar(w.scal.p, S_RESID, st);
    } else if (mode == DKT_CG_MIXED_RELIABLE) {
      iters = cg_reliable(pl, b, x, max_iter, tol, masked ? 1 : 0, st, &matvecs, &outer, &resid, &stop);
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
        if (resid <= tol) {
          stop = kStopConverged;
          break;
        }
        if (!std::isfinite(resid)) {
          stop = kStopBreakdown;
          break;
        }
        if (iters >= max_iter) {
          stop = kStopMaxIter;
          break;
        }
        hipLaunchKernelGGL(k_to32, dim3(cg_blocks(n)), dim3(kCgThreads), 0, st, n, w.d64.p, w.scal.p, w.b32.p, w.x32.p);
        // Inner tolerance: kInnerTol, loosened when the outer target is already close.
        const double inner_tol = std::min(0.5, std::max(kInnerTol, 0.5 * tol / resid));
        int inner_stop = 0;
        const int inner = cg_run<float>(pl, 1, w.b32.p, w.x32.p, true, max_iter - iters, inner_tol, masked ? 1 : 0, v,
                                        w.flag.p, w.host, st, &matvecs, &inner_stop);
        hipLaunchKernelGGL(k_add32, dim3(cg_blocks(n)), dim3(kCgThreads), 0, st, n, x, w.x32.p, w.scal.p);
        HIP_OK(hipGetLastError());
        iters += inner;
        outer++;
        if (inner == 0 || outer >= kMaxOuter) {  // no progress possible: report the last residual
          stop = inner == 0 && inner_stop == kStopBreakdown ? kStopBreakdown : kStopStalled;
          break;
        }
      }
    }
    if (stats) {
      stats[0] = iters;
      stats[1] = resid;
      stats[2] = outer;
      stats[3] = matvecs;
      stats[4] = stop;
    }
  });
  if (rc == 0 && stop == kStopBreakdown) {
    g_last_error = "CG breakdown: p.Ap <= 0 or a non-finite step (the operator is not positive definite on the "
                   "search direction); x holds the last iterate";
    return 2;
  }
  return rc;
}

// Continuous batching of CG solves.  Solves join (dkt_cg_engine_add) and
// leave (finished in dkt_cg_engine_step) between chunks of kCgChunk
// iterations; every chunk runs one matvec launch per (kind, dim, p) and
// precision of the active plans and one launch per vector step for all
// active solves of a precision.  A solve starts at the first step after it
// was added, so every step starts at iteration parity 0 and each solve's
// arithmetic is that of its own dkt_cg_solve from x = 0:
//   FP64 solves (DKT_CG_F64) run the FP64 iteration;
//   mixed solves (DKT_CG_MIXED_RELIABLE) run the FP32 iteration and, at the
//   start of the step after their FP32 residual reached its target (or
//   max_iter, or broke down), the FP64 residual replacement of cg_reliable
//   (kr_* kernels) for all such solves at once.
struct dkt_cg_engine {
  // A slot owns its CG workspace (kept and grown across solves), so several
  // solves may run on one plan at once.
  struct Slot {
    dkt_plan *pl = nullptr;
    int mode = DKT_CG_F64;
    CgDesc<double> d{};
    CgDesc<float> d32{};
    RelDesc rel{};
    bool active = false, fresh = false;
    int upd = 0;  // mixed: residual replacement due at the next step (1), start (2)
    size_t cap = 0, cap32 = 0;
    DevBuf<double> r, p, ap, part, scal, st;
    DevBuf<float> r32, p32, ap32, x32;
    DevBuf<double> part32, scal32;
    DevBuf<int> flag;
  };
  int masked = 0;
  std::deque<Slot> slots;  // fixed size after creation; Slot is not movable
  bool dirty = true;       // active set changed: rebuild the launches
  std::vector<int> act64, act32;  // active slots of each precision, in descriptor order
  std::unique_ptr<MatvecLauncher> mv64, mv32, mvu;
  // Tables and partial sums of the batched matvecs, rebuilt at every change of
  // the active set without allocating (each step ends synchronized).
  DevArena arena;
  DevBuf<CgDesc<double>> d64, dnew;
  DevBuf<CgDesc<float>> d32;
  DevBuf<RelDesc> drel;
  DevBuf<StatDesc> dstat;
  DevBuf<int> dupd;  // [slot] upd, [max_slots + slot] 0 when the slot's FP64 matvec runs this step
  DevBuf<double> status;
  int *hupd = nullptr;  // pinned host staging of dupd
  std::vector<double> hs;
  ~dkt_cg_engine() {
    if (hupd) (void)hipHostFree(hupd);
  }
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

extern "C" int dkt_cg_engine_add_mode(dkt_cg_engine *e, dkt_plan *pl, int mode, const double *b, double *x,
                                      int max_iter, double tol, int *slot) {
  return guarded([&] {
    if (!e || !pl || !b || !x || !slot) throw std::invalid_argument("null pointer");
    if (pl->dry) throw std::invalid_argument("a dry-run plan cannot solve");
    if (pl->simple || (pl->per_level && !pl->conforming))
      throw std::invalid_argument("batched solves need tiled device plans");
    if (max_iter < 0 || !(tol >= 0.0)) throw std::invalid_argument("max_iter >= 0 and tol >= 0 required");
    if ((const void *)b == (const void *)x) throw std::invalid_argument("b and x must not alias");
    if (mode == DKT_CG_MIXED) mode = DKT_CG_MIXED_RELIABLE;
    if (mode != DKT_CG_F64 && mode != DKT_CG_MIXED_RELIABLE)
      throw std::invalid_argument("batched solves run DKT_CG_F64 or DKT_CG_MIXED(_RELIABLE)");
    int free_slot = -1;
    for (int k = 0; k < (int)e->slots.size() && free_slot < 0; k++)
      if (!e->slots[k].active) free_slot = k;
    if (free_slot < 0) throw std::invalid_argument("no free slot");
    auto &s = e->slots[free_slot];
    const size_t n = pl->N;
    if (!e->hupd) {
      const size_t S = e->slots.size();
      HIP_OK(hipHostMalloc((void **)&e->hupd, 2 * S * sizeof(int), hipHostMallocDefault));
      e->dupd.alloc(2 * S);
    }
    if (s.cap < n) {
      s.r.alloc(n);
      s.p.alloc(n);
      s.ap.alloc(n);
      s.cap = n;
    }
    if (!s.part.p) {
      s.part.alloc(4 * kCgBlocks);
      s.scal.alloc(S_COUNT);
      s.flag.alloc(4);
      s.st.alloc(4);
    }
    if (mode == DKT_CG_MIXED_RELIABLE) {
      if (s.cap32 < n) {
        s.r32.alloc(n);
        s.p32.alloc(n);
        s.ap32.alloc(n);
        s.x32.alloc(n);
        s.cap32 = n;
      }
      if (!s.part32.p) {
        s.part32.alloc(4 * kCgBlocks);
        s.scal32.alloc(S_COUNT);
      }
    }
    s.pl = pl;
    s.mode = mode;
    const int N = (int)n;
    s.d = CgDesc<double>{N, max_iter, tol, nullptr, b, x, s.r.p, s.p.p, s.ap.p, s.part.p, s.scal.p, s.flag.p};
    if (mode == DKT_CG_MIXED_RELIABLE) {
      // FP64 residual path: r -> r64, ap -> ap64, part/scal -> the FP64 reductions.
      s.d32 = CgDesc<float>{N, max_iter, tol, s.scal32.p + S_TOLIN, nullptr, s.x32.p, s.r32.p, s.p32.p, s.ap32.p,
                            s.part32.p, s.scal32.p, s.flag.p};
      s.rel = RelDesc{N, max_iter, tol, b, x, s.r.p, s.ap.p, s.part.p, s.scal.p, s.x32.p, s.r32.p, s.p32.p,
                      s.part32.p, s.scal32.p, s.st.p, s.flag.p, nullptr};
      s.upd = 2;
    } else {
      s.upd = 0;
    }
    s.active = true;
    s.fresh = true;
    e->dirty = true;
    *slot = free_slot;
  });
}

extern "C" int dkt_cg_engine_add(dkt_cg_engine *e, dkt_plan *pl, const double *b, double *x, int max_iter, double tol,
                                 int *slot) {
  return dkt_cg_engine_add_mode(e, pl, DKT_CG_F64, b, x, max_iter, tol, slot);
}

// Starts the solves added since the last step, runs the due residual
// replacements and `chunks` chunks of iterations over the active solves, and
// reports those that stopped: finished[i] = slot, stats[3i] = iterations,
// stats[3i + 1] = relative residual, stats[3i + 2] = stop reason,
// *nfinished = their count.  Their slots are free again.
extern "C" int dkt_cg_engine_step(dkt_cg_engine *e, int chunks, void *stream, int *finished, double *stats,
                                  int *nfinished) {
  return guarded([&] {
    if (!e || !finished || !nfinished || chunks < 0) throw std::invalid_argument("null pointer or chunks < 0");
    hipStream_t st = (hipStream_t)stream;
    *nfinished = 0;
    const int S = (int)e->slots.size();
    if (e->dirty) {
      e->act64.clear();
      e->act32.clear();
      std::vector<CgDesc<double>> h64, hnew;
      std::vector<CgDesc<float>> h32;
      std::vector<RelDesc> hrel;
      std::vector<StatDesc> hstat;
      std::vector<dkt_plan *> p64, p32;
      std::vector<const void *> x64, x32, xu;
      std::vector<void *> y64, y32, yu;
      std::vector<const int *> c64, c32, cu;
      for (int k = 0; k < S; k++) {
        auto &s = e->slots[k];
        if (!s.active) continue;
        if (s.mode == DKT_CG_F64) {
          e->act64.push_back(k);
          h64.push_back(s.d);
          if (s.fresh) hnew.push_back(s.d);
          p64.push_back(s.pl);
          x64.push_back(s.d.p);
          y64.push_back(s.d.ap);
          c64.push_back(s.d.flag);  // flag[0]: stopped
        } else {
          e->act32.push_back(k);
          h32.push_back(s.d32);
          RelDesc r = s.rel;
          r.upd = e->dupd.p + k;
          hrel.push_back(r);
          p32.push_back(s.pl);
          x32.push_back(s.d32.p);
          y32.push_back(s.d32.ap);
          c32.push_back(s.d32.flag);
          xu.push_back(s.rel.x);
          yu.push_back(s.rel.ap64);
          cu.push_back(e->dupd.p + S + k);
        }
        s.fresh = false;
      }
      for (int k : e->act64) hstat.push_back(StatDesc{e->slots[k].d.flag, e->slots[k].d.scal + S_RESID});
      for (int k : e->act32) hstat.push_back(StatDesc{e->slots[k].d.flag, e->slots[k].rel.scal64 + S_RESID});
      e->mv64.reset();
      e->mv32.reset();
      e->mvu.reset();
      e->arena.reset(st);
      e->mv64 = p64.empty() ? nullptr : batch_matvec(0, p64, x64, y64, e->masked, c64, &e->arena);
      e->mv32 = p32.empty() ? nullptr : batch_matvec(1, p32, x32, y32, e->masked, c32, &e->arena);
      e->mvu = p32.empty() ? nullptr : batch_matvec(0, p32, xu, yu, e->masked, cu, &e->arena);
      auto put = [&](auto &buf, const auto &v) {
        if (v.empty()) return;
        if (buf.n < v.size()) buf.alloc(v.size());
        HIP_OK(hipMemcpyAsync(buf.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice, st));
      };
      put(e->d64, h64);
      put(e->dnew, hnew);
      put(e->d32, h32);
      put(e->drel, hrel);
      put(e->dstat, hstat);
      if (!hnew.empty()) {
        const int nn = (int)hnew.size();
        int nbn = 1;
        for (const auto &c : hnew) nbn = std::max(nbn, cg_blocks(c.n));
        hipLaunchKernelGGL(kb_init, dim3(nbn, nn), dim3(kCgThreads), 0, st, e->dnew.p);
        hipLaunchKernelGGL(kb_init_fin, dim3(1, nn), dim3(kCgThreads), 0, st, e->dnew.p);
        HIP_OK(hipGetLastError());
      }
      e->dirty = false;
    }
    const int K64 = (int)e->act64.size(), K32 = (int)e->act32.size();
    if (K64 + K32 == 0) return;
    int nb64 = 1, nb32 = 1;
    for (int k : e->act64) nb64 = std::max(nb64, cg_blocks(e->slots[k].d.n));
    for (int k : e->act32) nb32 = std::max(nb32, cg_blocks(e->slots[k].d.n));
    // Residual replacements (and starts) of the mixed solves that are due.
    bool any_upd = false;
    for (int k = 0; k < S; k++) {
      const int u = e->slots[k].active && e->slots[k].mode != DKT_CG_F64 ? e->slots[k].upd : 0;
      e->hupd[k] = u;
      e->hupd[S + k] = u == 1 ? 0 : 1;  // the FP64 matvec of x runs for replacements only
      any_upd = any_upd || u != 0;
    }
    if (any_upd) {
      HIP_OK(hipMemcpyAsync(e->dupd.p, e->hupd, 2 * S * sizeof(int), hipMemcpyHostToDevice, st));
      const dim3 g(nb32, K32), b1(1, K32), blk(kCgThreads);
      hipLaunchKernelGGL(kr_fold, g, blk, 0, st, e->drel.p);
      e->mvu->launch(st);
      hipLaunchKernelGGL(kr_resid, g, blk, 0, st, e->drel.p);
      hipLaunchKernelGGL(kr_decide, b1, blk, 0, st, e->drel.p);
      hipLaunchKernelGGL(kr_rel32, g, blk, 0, st, e->drel.p);
      hipLaunchKernelGGL(kr_rel32_fin, b1, blk, 0, st, e->drel.p);
      hipLaunchKernelGGL(kr_pupd32, g, blk, 0, st, e->drel.p);
      HIP_OK(hipGetLastError());
      for (int k : e->act32) e->slots[k].upd = 0;
    }
    const dim3 blk(kCgThreads);
    for (int c = 0; c < chunks * kCgChunk; c++) {
      const int par = c & 1;
      if (K64) {
        const dim3 g(nb64, K64);
        e->mv64->launch(st);
        hipLaunchKernelGGL((kb_dot<double>), g, blk, 0, st, e->d64.p);
        hipLaunchKernelGGL((kb_update_xr<double>), g, blk, 0, st, e->d64.p, par);
        hipLaunchKernelGGL((kb_update_p<double>), g, blk, 0, st, e->d64.p, par);
      }
      if (K32) {
        const dim3 g(nb32, K32);
        e->mv32->launch(st);
        hipLaunchKernelGGL((kb_dot<float>), g, blk, 0, st, e->d32.p);
        hipLaunchKernelGGL((kb_update_xr<float>), g, blk, 0, st, e->d32.p, par);
        hipLaunchKernelGGL((kb_update_p<float>), g, blk, 0, st, e->d32.p, par);
      }
      HIP_OK(hipGetLastError());
    }
    const int K = K64 + K32;
    if (e->status.n < 3 * (size_t)S) e->status.alloc(3 * (size_t)S);
    hipLaunchKernelGGL(kb_status, dim3(1), blk, 0, st, e->dstat.p, K, e->status.p);
    HIP_OK(hipGetLastError());
    e->hs.resize(3 * (size_t)K);
    HIP_OK(hipMemcpyAsync(e->hs.data(), e->status.p, e->hs.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int i = 0; i < K; i++) {
      const bool mixed = i >= K64;
      const int k = mixed ? e->act32[i - K64] : e->act64[i];
      const int f = (int)e->hs[3 * i];
      int reason = 0;
      if (!mixed) {
        reason = f;
      } else if (f >= kFinal) {
        reason = f - kFinal;
      } else if (f != 0) {
        e->slots[k].upd = 1;  // FP32 iteration stopped: residual replacement at the next step
        continue;
      }
      if (!reason) continue;
      finished[*nfinished] = k;
      if (stats) {
        stats[3 * *nfinished] = e->hs[3 * i + 1];
        stats[3 * *nfinished + 1] = e->hs[3 * i + 2];
        stats[3 * *nfinished + 2] = reason;
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
    std::vector<double> fst(3 * (size_t)nplans);
    int left = nplans;
    // The first step only starts the solves (they stop at once when b = 0 or max_iter = 0).
    for (int chunks = 0; left > 0; chunks = 1) {
      int nf = 0;
      if (dkt_cg_engine_step(e.get(), chunks, stream, fin.data(), fst.data(), &nf) != 0)
        throw std::runtime_error(dkt_last_error());
      for (int i = 0; i < nf; i++)
        if (stats) {
          stats[3 * fin[i]] = fst[3 * i];
          stats[3 * fin[i] + 1] = fst[3 * i + 1];
          stats[3 * fin[i] + 2] = fst[3 * i + 2];
        }
      left -= nf;
    }
  });
}

extern "C" int dkt_bicgstab_solve(dkt_plan *pl, const double *b, double *x, int max_iter, double tol, int masked,
                                  double *stats, void *stream) {
  int stop = 0;
  const int rc = guarded([&] {
    if (!pl || !b || !x) throw std::invalid_argument("null pointer");
    if (pl->dry) throw std::invalid_argument("a dry-run plan cannot solve");
    if (max_iter < 0 || !(tol >= 0.0)) throw std::invalid_argument("max_iter >= 0 and tol >= 0 required");
    if ((const void *)b == (const void *)x) throw std::invalid_argument("b and x must not alias");
    hipStream_t st = (hipStream_t)stream;
    const int n = (int)pl->N;
    CgWorkspace &w = pl->cg;
    if (w.n != pl->N) {
      w.release();
      w.n = pl->N;
    }
    if (!w.bi.p) {
      w.bi.alloc(6 * pl->N);
      w.bipart.alloc(6 * kCgBlocks);
      HIP_OK(hipMemset(w.bipart.p, 0, 6 * kCgBlocks * sizeof(double)));
      w.biscal.alloc(B_COUNT);
    }
    if (!w.flag.p) {
      w.flag.alloc(4);
      HIP_OK(hipHostMalloc((void **)&w.host, 4 * sizeof(int), hipHostMallocDefault));
    }
    double *base = w.bi.p;
    BiVecs v{base, base + pl->N, base + 2 * pl->N, base + 3 * pl->N, base + 4 * pl->N, base + 5 * pl->N,
             w.bipart.p, w.biscal.p};
    int *flag = w.flag.p, *host = w.host;
    const int mask = masked ? 1 : 0;
    const dim3 grid(cg_blocks(n)), blk(kCgThreads);
    plan_apply(pl, 0, x, v.v, 1, mask, st, nullptr);
    hipLaunchKernelGGL(kbi_init, grid, blk, 0, st, n, b, v.v, v);
    hipLaunchKernelGGL(kbi_init_fin, dim3(1), blk, 0, st, v, tol, flag);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(host, flag, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    int it = 0;
    stop = host[0] ? kStopConverged : (max_iter == 0 ? kStopMaxIter : 0);
    while (!stop) {
      for (int c = 0; c < kCgChunk; c++) {
        hipLaunchKernelGGL(kbi_p, grid, blk, 0, st, n, v, flag);
        plan_apply(pl, 0, v.p, v.v, 1, mask, st, flag);
        hipLaunchKernelGGL(kbi_dot, grid, blk, 0, st, n, v.rh, v.v, (const double *)nullptr,
                           (const double *)nullptr, v.part + 3 * kCgBlocks, flag);
        hipLaunchKernelGGL(kbi_s, grid, blk, 0, st, n, v, flag);
        plan_apply(pl, 0, v.s, v.t, 1, mask, st, flag);
        hipLaunchKernelGGL(kbi_dot, grid, blk, 0, st, n, v.t, v.s, v.t, v.t, v.part + 4 * kCgBlocks, flag);
        hipLaunchKernelGGL(kbi_xr, grid, blk, 0, st, n, x, v, flag);
        hipLaunchKernelGGL(kbi_fin, dim3(1), blk, 0, st, v, tol, max_iter, flag);
        HIP_OK(hipGetLastError());
      }
      HIP_OK(hipMemcpyAsync(host, flag, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      it = host[1];
      stop = host[0];
    }
    const double resid = read_scalar(w.biscal.p, B_RESID, st);
    if (stats) {
      stats[0] = it;
      stats[1] = resid;
      stats[2] = 0;
      stats[3] = 1 + 2 * it;
      stats[4] = stop;
    }
  });
  if (rc == 0 && stop == kStopBreakdown) {
    g_last_error = "BiCGStab breakdown: r^.v = 0, rho = 0 or omega = 0; x holds the last iterate";
    return DKT_ERR_BREAKDOWN;
  }
  return rc;
}
