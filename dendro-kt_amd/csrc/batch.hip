// Batched matvec over many independent meshes (no counterpart in the
// reference, which runs one fem::matvec per call, FEM/include/matvec.h:231).
//
// Plans of the same (dim, p) run as one launch: workgroup b finds its record
// by a binary search over the records' tile offsets and runs that tile with
// the record's maps and vectors (a device table of BatchPlan records, one per
// plan and chunk of right-hand sides that fits the tiles' LDS, each chunk with
// partial slots of its own), followed by one batched fix-up launch.  With
// DKT_BATCH_GRAPH the two launches of every (dim, p) group are captured once
// into a hipGraph and replayed.  Each plan's
// result is bitwise the same as its own dkt_plan_matvec: the tile work and
// the fix-up order are the plan's.
#include <algorithm>
#include <functional>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "../../include/dkt.h"
#include "capi_common.h"
#include "plan_common.h"
#include "plan_impl.h"
#include "tile_kernel.h"
#include "conf_kernel.h"

using namespace dkt;

namespace {

struct GroupBase {
  virtual ~GroupBase() = default;
  virtual void launch(hipStream_t st) = 0;
};

// Device copies of a group's tables: in the caller's arena when there is one,
// else in buffers the group owns.
struct GroupMem {
  DevArena *arena = nullptr;
  std::vector<std::unique_ptr<DevBuf<unsigned char>>> own;
  template <typename T>
  T *upload(const std::vector<T> &v) {
    if (arena) return arena->upload(v);
    own.emplace_back(new DevBuf<unsigned char>());
    own.back()->alloc(std::max<size_t>(1, v.size() * sizeof(T)));
    if (!v.empty()) HIP_OK(hipMemcpy(own.back()->p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return (T *)own.back()->p;
  }
  template <typename T>
  T *alloc(size_t count) {
    if (arena) return (T *)arena->take(count * sizeof(T));
    own.emplace_back(new DevBuf<unsigned char>());
    own.back()->alloc(std::max<size_t>(1, count * sizeof(T)));
    return (T *)own.back()->p;
  }
};

// Shared nodes of one plan of a batch (the fix-up of either kind of tile).
template <typename T>
struct FixRec {
  const int32_t *snode, *soff;
  const T *partial;
  T *y;
  const uint8_t *dir_out;
  const int *cancel;
  int nq, ld, nshared;
};

// Shared nodes of a batch: workgroup b sums block b - block_start[k] of plan k.
template <typename T>
__global__ void __launch_bounds__(256) fixup_kernel_rec(const FixRec<T> *__restrict__ recs,
                                                        const int *__restrict__ block_start, int nplans) {
  const int k = find_range(block_start, nplans, blockIdx.x);
  const FixRec<T> &fr = recs[k];
  const int i = (blockIdx.x - block_start[k]) * 256 + threadIdx.x;
  if (i >= fr.nshared || (fr.cancel && *fr.cancel)) return;
  const int n = fr.snode[i];
  const int a = fr.soff[i], b = fr.soff[i + 1];
  const bool z = fr.dir_out && fr.dir_out[n];
  for (int q = 0; q < fr.nq; q++) fr.y[(size_t)n * fr.ld + q] = z ? T(0) : sum_partials(fr.partial, a, b, fr.nq, q);
}

// The fix-up records of a group (one per tile record, host copies): a batch
// sums the shared nodes of all its groups in one launch after their tiles.
template <typename T>
struct FixSource {
  virtual ~FixSource() = default;
  std::vector<FixRec<T>> fix;
};

struct FixLaunchBase {
  virtual ~FixLaunchBase() = default;
  virtual void launch(hipStream_t st) = 0;
};

template <typename T>
struct FixAll : FixLaunchBase {
  GroupMem mem;
  FixRec<T> *table = nullptr;
  int *block_start = nullptr;
  int nrec = 0, nblocks = 0;
  FixAll(const std::vector<FixRec<T>> &recs, DevArena *arena) {
    mem.arena = arena;
    std::vector<int> bs{0};
    for (const FixRec<T> &r : recs) bs.push_back(bs.back() + (r.nshared + 255) / 256);
    nrec = (int)recs.size();
    nblocks = bs.back();
    table = mem.upload(recs);
    block_start = mem.upload(bs);
  }
  void launch(hipStream_t st) override {
    if (!nblocks) return;
    hipLaunchKernelGGL((fixup_kernel_rec<T>), dim3(nblocks), dim3(256), 0, st, table, block_start, nrec);
    HIP_OK(hipGetLastError());
  }
};

template <typename T, int DIM, int P>
struct Group : GroupBase, FixSource<T> {
  GroupMem mem;
  // One record per (column chunk, plan): columns [q0, q0 + nq) of the plan,
  // every chunk in the same launch with partial slots of its own.
  BatchPlan<T> *table = nullptr;
  int *tile_start = nullptr;
  T *partial = nullptr;
  int nrec = 0, ntiles = 0, threads = 0;
  size_t lds = 0;
  bool one = false;  // every record has one column

  Group(const std::vector<dkt_plan *> &pls, const std::vector<const void *> &xs, const std::vector<void *> &ys,
        int nrhs, int masked, const std::vector<const int *> &cancel, DevArena *arena = nullptr, int max_cols = 0,
        bool colmajor = false) {
    mem.arena = arena;
    constexpr int NPE = Pow<P + 1, DIM>::v;
    const bool f64 = sizeof(T) == 8;
    auto dims_of = [](const dkt_plan *pl) {
      return TileDims{pl->max_nodes, pl->max_grp, pl->max_he, pl->max_rd, pl->max_cr, pl->tile_leaves, pl->max_hang};
    };
    // Columns per launch: the most every plan's tiles fit in LDS.
    // Column-major vectors ([nrhs][N]): one contiguous column per launch.
    int chunk = colmajor ? 1 : (max_cols > 0 ? std::min(nrhs, max_cols) : nrhs);
    for (dkt_plan *pl : pls) {
      if (tile_lds_bytes<T, NPE>(dims_of(pl), 1) > kLdsLimit) throw std::invalid_argument("tile does not fit in LDS");
      while (chunk > 1 && tile_lds_bytes<T, NPE>(dims_of(pl), chunk) > kLdsLimit) chunk--;
    }
    const int nchunks = (nrhs + chunk - 1) / chunk;
    std::vector<int> ts{0};
    std::vector<size_t> poff;
    size_t ptot = 0;
    for (int c = 0; c < nchunks; c++)
      for (dkt_plan *pl : pls) {
        threads = std::max(threads, pl->tile_leaves);
        poff.push_back(ptot);
        ptot += pl->npartial * (size_t)chunk;
        ts.push_back(ts.back() + (int)pl->ntiles);
      }
    partial = mem.alloc<T>(std::max<size_t>(1, ptot));
    one = chunk == 1;
    std::vector<BatchPlan<T>> hp(pls.size() * nchunks);
    for (int c = 0; c < nchunks; c++) {
      const int q0 = c * chunk, nq = std::min(chunk, nrhs - q0);
      for (size_t k = 0; k < pls.size(); k++) {
        dkt_plan *pl = pls[k];
        BatchPlan<T> &r = hp[c * pls.size() + k];
        lds = std::max(lds, tile_lds_bytes<T, NPE>(dims_of(pl), nq));
        TileArgs &A = r.A;
        A.tiles = pl->tiles.p;
        A.tile_nodes = pl->tile_nodes.p;
        A.tile_out = pl->tile_out.p;
        A.leaf_loc = pl->leaf_loc.p;
        A.leaf_aux = pl->leaf_aux.p;
        A.group_ploc = pl->group_ploc.p;
        A.hentries = pl->hentries.p;
        A.readers = pl->readers.p;
        A.blocks = pl->blocks.p;
        A.carry_key = pl->carry_key.p;
        A.dir_in = masked == 1 ? pl->dirichlet.p : nullptr;
        A.dir_out = masked ? pl->dirichlet.p : nullptr;
        A.cancel = cancel.empty() ? nullptr : cancel[k];
        A.op = pl->op;
        A.hp = pl->hp;
        A.nq = nq;
        A.ld = colmajor ? 1 : nrhs;
        A.dims = dims_of(pl);
        r.kref = P == 2 ? (f64 ? (const T *)pl->kfac64.p : (const T *)pl->kfac32.p)
                        : (f64 ? (const T *)pl->kmat64.p : (const T *)pl->kmat32.p);
        r.carry_w = f64 ? (const T *)pl->carry_w64.p : (const T *)pl->carry_w32.p;
        const size_t coff = colmajor ? (size_t)q0 * pl->N : (size_t)q0;
        r.x = (const T *)xs[k] + coff;
        r.y = (T *)ys[k] + coff;
        r.partial = partial + poff[c * pls.size() + k];
        r.snode = pl->snode.p;
        r.soff = pl->soff.p;
        r.nshared = (int)pl->nshared;
        this->fix.push_back(FixRec<T>{pl->snode.p, pl->soff.p, r.partial, r.y, A.dir_out, A.cancel, nq, A.ld,
                                      (int)pl->nshared});
      }
    }
    table = mem.upload(hp);
    nrec = (int)hp.size();
    ntiles = ts.back();
    tile_start = mem.upload(ts);
  }

  void launch(hipStream_t st) override {
    if (ntiles && one)
      hipLaunchKernelGGL((tile_kernel_batch<T, DIM, P, true>), dim3(ntiles), dim3(threads), lds, st, table,
                         tile_start, nrec);
    else if (ntiles)
      hipLaunchKernelGGL((tile_kernel_batch<T, DIM, P, false>), dim3(ntiles), dim3(threads), lds, st, table,
                         tile_start, nrec);
    HIP_OK(hipGetLastError());
  }
};

// A conforming plan of a batch.
template <typename T>
struct ConfBatchPlan {
  ConfArgs A;
  const T *kref, *term_w, *x;
  T *y, *partial;
};

// All tiles of a batch of same-(dim, p) conforming plans in one launch.
template <typename T, int DIM, int P, bool CAP>
__global__ void __launch_bounds__(kTileThreads) conf_tile_kernel_batch(const ConfBatchPlan<T> *__restrict__ plans,
                                                                       const int *__restrict__ tile_start,
                                                                       int nplans) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int k = find_range(tile_start, nplans, t);
  const ConfBatchPlan<T> bp = plans[k];  // a copy: its fields live in scalar registers
  if (bp.A.cancel && *bp.A.cancel) return;
  conf_tile_body<T, DIM, P, CAP>(bp.A, t - tile_start[k], bp.kref, bp.term_w, bp.x, bp.y, bp.partial);
}

// Conforming plans of one (dim, p) (CAP: with per-level captured matrices):
// one tile launch and one fix-up launch, with one record per (column chunk
// that fits the tiles' LDS, plan).  Each result is bitwise the plan's own
// matvec (same tile work, same fix-up order).
template <typename T, int DIM, int P, bool CAP>
struct ConfGroup : GroupBase, FixSource<T> {
  GroupMem mem;
  ConfBatchPlan<T> *table = nullptr;
  int *tile_start = nullptr;
  T *partial = nullptr;
  int nrec = 0, ntiles = 0, threads = 0;
  size_t lds = 0;

  ConfGroup(const std::vector<dkt_plan *> &pls, const std::vector<const void *> &xs, const std::vector<void *> &ys,
            int nrhs, int masked, const std::vector<const int *> &cancel, DevArena *arena = nullptr,
            int max_cols = 0, bool colmajor = false) {
    mem.arena = arena;
    constexpr int NPE = Pow<P + 1, DIM>::v;
    const bool f64 = sizeof(T) == 8;
    auto lds_of = [](const dkt_plan *pl, int nq) { return conf_lds_bytes<T, NPE>(pl->max_nodes, nq, pl->nw); };
    int chunk = colmajor ? 1 : (max_cols > 0 ? std::min(nrhs, max_cols) : nrhs);
    for (dkt_plan *pl : pls) {
      if (lds_of(pl, 1) > kLdsLimit) throw std::invalid_argument("tile does not fit in LDS");
      while (chunk > 1 && lds_of(pl, chunk) > kLdsLimit) chunk--;
    }
    const int nchunks = (nrhs + chunk - 1) / chunk;
    std::vector<int> ts{0};
    std::vector<size_t> poff;
    size_t ptot = 0;
    for (int c = 0; c < nchunks; c++)
      for (dkt_plan *pl : pls) {
        threads = std::max(threads, pl->tile_leaves);
        poff.push_back(ptot);
        ptot += pl->npartial * (size_t)chunk;
        ts.push_back(ts.back() + (int)pl->ntiles);
      }
    partial = mem.alloc<T>(std::max<size_t>(1, ptot));
    std::vector<ConfBatchPlan<T>> hp(pls.size() * nchunks);
    std::vector<FixRec<T>> &hf = this->fix;
    hf.resize(pls.size() * nchunks);
    for (int c = 0; c < nchunks; c++) {
      const int q0 = c * chunk, nq = std::min(chunk, nrhs - q0);
      for (size_t k = 0; k < pls.size(); k++) {
        dkt_plan *pl = pls[k];
        const size_t ri = c * pls.size() + k;
        lds = std::max(lds, lds_of(pl, nq));
        ConfArgs &A = hp[ri].A;
        A.tiles = pl->ctiles.p;
        A.tile_nodes = pl->tile_nodes.p;
        A.tile_out = pl->tile_out.p;
        A.leaf_loc = pl->leaf_loc.p;
        A.leaf_lev = pl->leaf_lev.p;
        A.virt_off = pl->virt_off.p;
        A.term_pk = pl->term_pk.p;
        A.term_spk = pl->term_spk.p;
        A.seg = pl->seg.p;
        A.nw = pl->nw;
        A.dir_in = masked == 1 ? pl->dirichlet.p : nullptr;
        A.dir_out = masked ? pl->dirichlet.p : nullptr;
        A.cancel = cancel.empty() ? nullptr : cancel[k];
        A.op = pl->op;
        A.hp = pl->hp;
        A.nq = nq;
        A.ld = colmajor ? 1 : nrhs;
        A.max_points = pl->max_nodes;
        A.tile_leaves = pl->tile_leaves;
        hp[ri].kref = P == 2 && !pl->per_level ? (f64 ? (const T *)pl->kfac64.p : (const T *)pl->kfac32.p)
                                               : (f64 ? (const T *)pl->kmat64.p : (const T *)pl->kmat32.p);
        hp[ri].term_w = f64 ? (const T *)pl->term_w64.p : (const T *)pl->term_w32.p;
        const size_t coff = colmajor ? (size_t)q0 * pl->N : (size_t)q0;
        hp[ri].x = (const T *)xs[k] + coff;
        hp[ri].y = (T *)ys[k] + coff;
        hp[ri].partial = partial + poff[ri];
        hf[ri] = FixRec<T>{pl->snode.p, pl->soff.p, partial + poff[ri], (T *)ys[k] + coff, A.dir_out, A.cancel,
                           nq, A.ld, (int)pl->nshared};
      }
    }
    table = mem.upload(hp);
    nrec = (int)hp.size();
    ntiles = ts.back();
    tile_start = mem.upload(ts);
  }

  void launch(hipStream_t st) override {
    if (ntiles)
      hipLaunchKernelGGL((conf_tile_kernel_batch<T, DIM, P, CAP>), dim3(ntiles), dim3(threads), lds, st, table,
                         tile_start, nrec);
    HIP_OK(hipGetLastError());
  }
};

// ---------------------------------------------------------------------------
// One persistent launch for a whole batch (DKT_BATCH_PERSISTENT): the tiles of
// every plan and column chunk, any mix of (dim, p), then the fix-up of their
// shared nodes, in a single kernel.
//
// Work is claimed with global counters: tile work comes from eight queues, one
// contiguous run of tiles per queue, and workgroup b takes from queue b % 8
// first (workgroups are dispatched to the eight XCDs round-robin, so each XCD's
// L2 sees one SFC-contiguous run, as xcd_tile arranges for the plain
// launches), then from the other queues in turn.  A workgroup moves on to the
// fix-up blocks only when every queue is empty, so every tile is then claimed
// by a running workgroup; a fix-up block of a (plan, chunk) record waits until
// every tile of that record has finished (a per-record completion counter,
// released by each tile after its write-out), a wait that ends once the tiles
// in flight finish.  The wait polls the counter with relaxed atomic loads and
// acquires once when it is complete: an acquiring load per poll would
// invalidate the L2 on every poll and starve the very tiles it waits for.  A
// bounded wait records an error flag instead of hanging.  The last workgroup
// to leave resets every counter, so the next launch (a graph replay) starts
// from zero with no memset node.
// Each plan's result is bitwise its own dkt_plan_matvec (same tile work, same
// fix-up order).
// ---------------------------------------------------------------------------
constexpr int kQueues = 8;
constexpr long long kSpinLimit = 1ll << 20;  // polls (about a microsecond each) before a wait gives up

struct PersistCtl {
  unsigned int next[kQueues];
  unsigned int fix_next;
  unsigned int exited;
  unsigned int err;
};

template <typename T>
struct PersistRec {
  BatchPlan<T> bp;
  int kind;  // 0: (2,1) 1: (3,1) 2: (4,1) 3: (2,2) 4: (3,2)
  int ntiles;
};

// The tile bodies a persistent kernel carries: bit k of KINDS for kind k.  A
// kernel's registers are those of its largest body, so a batch runs the
// smallest instantiation that covers its kinds.
constexpr int kKindsAll = 31, kKinds3D = (1 << 1) | (1 << 4);

template <typename T, bool ONE, int KINDS>
__device__ __forceinline__ void persist_tile(const PersistRec<T> &rec, int t) {
  const BatchPlan<T> &bp = rec.bp;
  if ((KINDS & 1) && rec.kind == 0) tile_body<T, 2, 1, ONE>(bp.A, t, bp.kref, bp.carry_w, bp.x, bp.y, bp.partial);
  if ((KINDS & 2) && rec.kind == 1) tile_body<T, 3, 1, ONE>(bp.A, t, bp.kref, bp.carry_w, bp.x, bp.y, bp.partial);
  if ((KINDS & 4) && rec.kind == 2) tile_body<T, 4, 1, ONE>(bp.A, t, bp.kref, bp.carry_w, bp.x, bp.y, bp.partial);
  if ((KINDS & 8) && rec.kind == 3) tile_body<T, 2, 2, ONE>(bp.A, t, bp.kref, bp.carry_w, bp.x, bp.y, bp.partial);
  if ((KINDS & 16) && rec.kind == 4) tile_body<T, 3, 2, ONE>(bp.A, t, bp.kref, bp.carry_w, bp.x, bp.y, bp.partial);
}

template <typename T, bool ONE, int KINDS>
__global__ void __launch_bounds__(kTileThreads) tile_kernel_persistent(
    const PersistRec<T> *__restrict__ recs, const int *__restrict__ tile_start, const int *__restrict__ fix_start,
    int nrec, int ntiles, int nfix, unsigned int *__restrict__ done, PersistCtl *__restrict__ ctl) {
  __shared__ int s_item;
  __shared__ int s_ok;
  for (int qi = 0; qi < kQueues; qi++) {
    const int q = (blockIdx.x + qi) % kQueues;  // own queue first, then the others
    const int lo = (int)((long long)ntiles * q / kQueues), hi = (int)((long long)ntiles * (q + 1) / kQueues);
  for (;;) {
    if (threadIdx.x == 0) s_item = lo + (int)atomicAdd(&ctl->next[q], 1u);
    __syncthreads();
    const int t = s_item;
    __syncthreads();
    if (t >= hi) break;
    const int r = find_range(tile_start, nrec, t);
    const PersistRec<T> &rec = recs[r];
    if (!(rec.bp.A.cancel && *rec.bp.A.cancel)) persist_tile<T, ONE, KINDS>(rec, t - tile_start[r]);
    __syncthreads();  // every work-item's write-out is issued
    if (threadIdx.x == 0) {
      __threadfence();
      atomicAdd(&done[r], 1u);
    }
  }
  }
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(&ctl->fix_next, 1u);
    __syncthreads();
    const int f = s_item;
    __syncthreads();
    if (f >= nfix) break;
    const int r = find_range(fix_start, nrec, f);
    const PersistRec<T> &rec = recs[r];
    if (threadIdx.x == 0) {
      long long spins = 0;
      int ok = 1;
      const unsigned int need = (unsigned int)rec.ntiles;
      while (__hip_atomic_load(&done[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        if (++spins > kSpinLimit) {
          atomicOr(&ctl->err, 1u);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(8);
      }
      __atomic_thread_fence(__ATOMIC_ACQUIRE);  // the record's partials are visible from here on
      s_ok = ok;
    }
    __syncthreads();
    const BatchPlan<T> &bp = rec.bp;
    if (!s_ok || (bp.A.cancel && *bp.A.cancel)) continue;
    const int nq = bp.A.nq, ld = bp.A.ld;
    for (int k = threadIdx.x; k < 256; k += blockDim.x) {
      const int i = (f - fix_start[r]) * 256 + k;
      if (i >= bp.nshared) break;
      const int n = bp.snode[i];
      const int a = bp.soff[i], b = bp.soff[i + 1];
      const bool z = bp.A.dir_out && bp.A.dir_out[n];
      for (int qq = 0; qq < nq; qq++) bp.y[(size_t)n * ld + qq] = z ? T(0) : sum_partials(bp.partial, a, b, nq, qq);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&ctl->exited, 1u) == gridDim.x - 1) {  // the last workgroup: reset for the next launch
      for (int i = 0; i < kQueues; i++) ctl->next[i] = 0;
      ctl->fix_next = 0;
      for (int i = 0; i < nrec; i++) done[i] = 0;
      __threadfence();
      ctl->exited = 0;
    }
  }
}

int plan_kind(const dkt_plan *pl) {
  if (pl->order == 1) return pl->dim - 2;
  if (pl->dim == 2) return 3;
  if (pl->dim == 3) return 4;
  throw std::invalid_argument("unsupported (dim, order)");
}

template <typename T>
struct PersistGroup : GroupBase {
  DevBuf<PersistRec<T>> recs;
  DevBuf<int> tile_start, fix_start;
  DevBuf<unsigned int> done;
  DevBuf<PersistCtl> ctl;
  DevBuf<T> partial;
  int nrec = 0, ntiles = 0, nfix = 0, threads = 0, grid = 0, kinds = 0;
  bool one = false;
  size_t lds = 0;
  using KernelFn = void (*)(const PersistRec<T> *, const int *, const int *, int, int, int, unsigned int *,
                            PersistCtl *);
  KernelFn fn = nullptr;

  template <bool ONE>
  KernelFn pick() {
    switch (kinds) {
      case 1: return tile_kernel_persistent<T, ONE, 1>;
      case 2: return tile_kernel_persistent<T, ONE, 2>;
      case 4: return tile_kernel_persistent<T, ONE, 4>;
      case 8: return tile_kernel_persistent<T, ONE, 8>;
      case 16: return tile_kernel_persistent<T, ONE, 16>;
      default:
        if ((kinds & ~kKinds3D) == 0) return tile_kernel_persistent<T, ONE, kKinds3D>;
        return tile_kernel_persistent<T, ONE, kKindsAll>;
    }
  }

  PersistGroup(const std::vector<dkt_plan *> &pls, const std::vector<const void *> &xs, const std::vector<void *> &ys,
               int nrhs, int masked) {
    const bool f64 = sizeof(T) == 8;
    auto dims_of = [](const dkt_plan *pl) {
      return TileDims{pl->max_nodes, pl->max_grp, pl->max_he, pl->max_rd, pl->max_cr, pl->tile_leaves, pl->max_hang};
    };
    auto lds_of = [&](const dkt_plan *pl, int nq) {
      const TileDims D = dims_of(pl);
      switch (plan_kind(pl)) {
        case 0: return tile_lds_bytes<T, 4>(D, nq);
        case 1: return tile_lds_bytes<T, 8>(D, nq);
        case 2: return tile_lds_bytes<T, 16>(D, nq);
        case 3: return tile_lds_bytes<T, 9>(D, nq);
        default: return tile_lds_bytes<T, 27>(D, nq);
      }
    };
    // Columns per launch: the most that every plan's tiles fit in LDS; one
    // column runs the single-column tile bodies.
    int chunk = nrhs;
    for (dkt_plan *pl : pls) {
      if (lds_of(pl, 1) > kLdsLimit) throw std::invalid_argument("tile does not fit in LDS");
      while (chunk > 1 && lds_of(pl, chunk) > kLdsLimit) chunk--;
    }
    one = chunk == 1;
    std::vector<PersistRec<T>> hr;
    std::vector<int> ts{0}, fs{0};
    std::vector<size_t> poff;
    size_t ptot = 0;
    for (int q0 = 0; q0 < nrhs; q0 += chunk) {
      const int nq = std::min(chunk, nrhs - q0);
      if (nq != chunk && !one) {
        // a short last chunk runs with the multi-column body too (nq is a runtime argument there)
      }
      for (size_t k = 0; k < pls.size(); k++) {
        dkt_plan *pl = pls[k];
        threads = std::max(threads, pl->tile_leaves);
        lds = std::max(lds, lds_of(pl, nq));
        PersistRec<T> r{};
        TileArgs &A = r.bp.A;
        A.tiles = pl->tiles.p;
        A.tile_nodes = pl->tile_nodes.p;
        A.tile_out = pl->tile_out.p;
        A.leaf_loc = pl->leaf_loc.p;
        A.leaf_aux = pl->leaf_aux.p;
        A.group_ploc = pl->group_ploc.p;
        A.hentries = pl->hentries.p;
        A.readers = pl->readers.p;
        A.blocks = pl->blocks.p;
        A.carry_key = pl->carry_key.p;
        A.dir_in = masked == 1 ? pl->dirichlet.p : nullptr;
        A.dir_out = masked ? pl->dirichlet.p : nullptr;
        A.cancel = nullptr;
        A.op = pl->op;
        A.hp = pl->hp;
        A.nq = nq;
        A.ld = nrhs;
        A.dims = dims_of(pl);
        r.bp.kref = pl->order == 2 ? (f64 ? (const T *)pl->kfac64.p : (const T *)pl->kfac32.p)
                                   : (f64 ? (const T *)pl->kmat64.p : (const T *)pl->kmat32.p);
        r.bp.carry_w = f64 ? (const T *)pl->carry_w64.p : (const T *)pl->carry_w32.p;
        r.bp.x = (const T *)xs[k] + q0;
        r.bp.y = (T *)ys[k] + q0;
        poff.push_back(ptot);
        ptot += pl->npartial * (size_t)nq;
        r.bp.snode = pl->snode.p;
        r.bp.soff = pl->soff.p;
        r.bp.nshared = (int)pl->nshared;
        r.kind = plan_kind(pl);
        kinds |= 1 << r.kind;
        r.ntiles = (int)pl->ntiles;
        hr.push_back(r);
        ts.push_back(ts.back() + (int)pl->ntiles);
        fs.push_back(fs.back() + (int)((pl->nshared + 255) / 256));
      }
    }
    partial.alloc(std::max<size_t>(1, ptot));
    for (size_t i = 0; i < hr.size(); i++) hr[i].bp.partial = partial.p + poff[i];
    nrec = (int)hr.size();
    ntiles = ts.back();
    nfix = fs.back();
    recs.upload(hr);
    tile_start.upload(ts);
    fix_start.upload(fs);
    done.upload(std::vector<unsigned int>(nrec, 0u));
    ctl.upload(std::vector<PersistCtl>(1, PersistCtl{}));
    // Persistent grid: as many workgroups as fit on the device at once (at
    // least one per queue), never more than there is work.
    int dev = 0, cus = 0, per_cu = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (lds > kLdsLimit) throw std::invalid_argument("tile does not fit in LDS");
    fn = one ? pick<true>() : pick<false>();
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds));
    grid = std::max(kQueues, std::min(std::max(1, cus * std::max(1, per_cu)), std::max(ntiles + nfix, kQueues)));
  }

  void launch(hipStream_t st) override {
    hipLaunchKernelGGL(fn, dim3(grid), dim3(threads), lds, st, (const PersistRec<T> *)recs.p,
                       (const int *)tile_start.p, (const int *)fix_start.p, nrec, ntiles, nfix, done.p, ctl.p);
    HIP_OK(hipGetLastError());
  }

  unsigned int error_flag() {
    PersistCtl h{};
    HIP_OK(hipMemcpy(&h, ctl.p, sizeof(h), hipMemcpyDeviceToHost));
    return h.err;
  }
};

template <typename T>
// kind: 0 reference-semantics tiles, 1 conforming tiles, 2 conforming tiles
// with per-level captured element matrices.
GroupBase *make_group(int kind, int dim, int p, const std::vector<dkt_plan *> &pls,
                      const std::vector<const void *> &xs, const std::vector<void *> &ys, int nrhs, int masked,
                      const std::vector<const int *> &cancel = {}, DevArena *arena = nullptr, int max_cols = 0,
                      bool colmajor = false) {
#define DKT_GROUP(D, PP)                                                                                     \
  if (dim == D && p == PP) {                                                                                 \
    if (kind == 2)                                                                                           \
      return new ConfGroup<T, D, PP, true>(pls, xs, ys, nrhs, masked, cancel, arena, max_cols, colmajor);    \
    if (kind == 1)                                                                                           \
      return new ConfGroup<T, D, PP, false>(pls, xs, ys, nrhs, masked, cancel, arena, max_cols, colmajor);   \
    return new Group<T, D, PP>(pls, xs, ys, nrhs, masked, cancel, arena, max_cols, colmajor);                \
  }
  DKT_GROUP(2, 1)
  DKT_GROUP(3, 1)
  DKT_GROUP(4, 1)
  DKT_GROUP(2, 2)
  DKT_GROUP(3, 2)
#undef DKT_GROUP
  throw std::invalid_argument("unsupported (dim, order)");
}

// Plans grouped by (kind as make_group's, dim, p): indices into the caller's list.
std::map<std::tuple<int, int, int>, std::vector<int>> by_kind_of(dkt_plan *const *plans, int nplans) {
  std::map<std::tuple<int, int, int>, std::vector<int>> by_kind;
  for (int k = 0; k < nplans; k++) {
    dkt_plan *pl = plans[k];
    if (pl->dry || pl->simple || (pl->per_level && !pl->conforming))
      throw std::invalid_argument("batch plans must be tiled device plans (reference or conforming)");
    by_kind[{pl->conforming ? (pl->per_level ? 2 : 1) : 0, pl->dim, pl->order}].push_back(k);
  }
  return by_kind;
}

// One fix-up launch for the shared nodes of every group (the groups' tiles
// have all run before it; their partial slots and outputs are disjoint).
template <typename T>
std::unique_ptr<FixLaunchBase> fix_all(const std::vector<std::unique_ptr<GroupBase>> &groups, DevArena *arena) {
  std::vector<FixRec<T>> recs;
  for (const auto &g : groups)
    if (auto *fs = dynamic_cast<FixSource<T> *>(g.get())) recs.insert(recs.end(), fs->fix.begin(), fs->fix.end());
  return std::make_unique<FixAll<T>>(recs, arena);
}

struct GroupsLauncher : MatvecLauncher {
  std::vector<std::unique_ptr<GroupBase>> groups;
  std::unique_ptr<FixLaunchBase> fixes;
  void launch(hipStream_t st) override {
    for (auto &g : groups) g->launch(st);
    if (fixes) fixes->launch(st);
  }
};

}  // namespace

std::unique_ptr<MatvecLauncher> dkt::batch_matvec(int dtype, const std::vector<dkt_plan *> &plans,
                                                  const std::vector<const void *> &x, const std::vector<void *> &y,
                                                  int mask, const std::vector<const int *> &cancel, DevArena *arena) {
  auto out = std::make_unique<GroupsLauncher>();
  for (const auto &kv : by_kind_of(plans.data(), (int)plans.size())) {
    std::vector<dkt_plan *> pls;
    std::vector<const void *> xs;
    std::vector<void *> ys;
    std::vector<const int *> cs;
    for (int k : kv.second) {
      pls.push_back(plans[k]);
      xs.push_back(x[k]);
      ys.push_back(y[k]);
      cs.push_back(cancel[k]);
    }
    const auto [conf, dim, p] = kv.first;
    out->groups.emplace_back(dtype == 0 ? make_group<double>(conf, dim, p, pls, xs, ys, 1, mask, cs, arena)
                                        : make_group<float>(conf, dim, p, pls, xs, ys, 1, mask, cs, arena));
  }
  out->fixes = dtype == 0 ? fix_all<double>(out->groups, arena) : fix_all<float>(out->groups, arena);
  return out;
}

struct dkt_batch {
  std::vector<std::unique_ptr<GroupBase>> groups;
  std::unique_ptr<FixLaunchBase> fixes;  // the shared nodes of every group, after their tiles
  void launch(hipStream_t st) {
    for (auto &g : groups) g->launch(st);
    if (fixes) fixes->launch(st);
  }
  // Graph capture: the groups' tile launches are independent (their records,
  // partial slots and outputs are disjoint), so each goes on a branch of its
  // own, forked from and joined back to the caller's stream before the
  // fix-up; the replayed graph may then run them side by side.  The branch
  // streams have the default priority.
  std::vector<hipStream_t> branch;
  std::vector<hipEvent_t> done;
  hipEvent_t fork = nullptr;
  // Streams and events of the branches, made before the capture starts.
  void prepare_branches() {
    if (fork || groups.size() < 2) return;
    HIP_OK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    for (size_t g = 1; g < groups.size(); g++) {
      hipStream_t s2 = nullptr;
      HIP_OK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
      branch.push_back(s2);
      hipEvent_t e = nullptr;
      HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      done.push_back(e);
    }
  }
  void capture(hipStream_t st) {
    if (groups.size() < 2 || !fork) {
      launch(st);
      return;
    }
    HIP_OK(hipEventRecord(fork, st));
    for (size_t g = 1; g < groups.size(); g++) {
      HIP_OK(hipStreamWaitEvent(branch[g - 1], fork, 0));
      groups[g]->launch(branch[g - 1]);
      HIP_OK(hipEventRecord(done[g - 1], branch[g - 1]));
    }
    groups[0]->launch(st);
    for (size_t g = 1; g < groups.size(); g++) HIP_OK(hipStreamWaitEvent(st, done[g - 1], 0));
    if (fixes) fixes->launch(st);
  }
  std::function<unsigned int()> error_flag;  // persistent launches: the spin-limit flag
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~dkt_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : done) (void)hipEventDestroy(e);
    for (hipStream_t s2 : branch) (void)hipStreamDestroy(s2);
    if (fork) (void)hipEventDestroy(fork);
  }
};

extern "C" int dkt_batch_create(dkt_plan *const *plans, int nplans, int dtype, int nrhs, const void *const *x,
                                void *const *y, int masked, int flags, dkt_batch **out) {
  return guarded([&] {
    if (!plans || !x || !y || !out || nplans < 1) throw std::invalid_argument("null pointer or empty batch");
    if (dtype != 0 && dtype != 1) throw std::invalid_argument("dtype must be 0 (f64) or 1 (f32)");
    if (nrhs < 1 || nrhs > 64) throw std::invalid_argument("nrhs must be 1..64");
    if (masked < 0 || masked > 2) throw std::invalid_argument("masked must be 0, 1 or 2");
    for (int k = 0; k < nplans; k++) {
      if (!plans[k] || !x[k] || !y[k]) throw std::invalid_argument("null plan or vector");
      if (x[k] == (const void *)y[k]) throw std::invalid_argument("x and y must not alias");
    }
    const int cbits = (flags >> DKT_BATCH_COLUMNS_SHIFT) & 15;
    if (cbits > 7) throw std::invalid_argument("columns per launch must be 1, 2, 4, ..., 64");
    const int max_cols = cbits ? 1 << (cbits - 1) : 0;
    auto b = std::make_unique<dkt_batch>();
    if (flags & DKT_BATCH_PERSISTENT)
      throw std::invalid_argument(
          "DKT_BATCH_PERSISTENT is switched off: its launch did not finish on the device and the cause is not "
          "found; use the per-(dim, p) launches (with DKT_BATCH_GRAPH for a replayed graph)");
    if (flags & DKT_BATCH_PERSISTENT) {
      std::vector<dkt_plan *> pls(plans, plans + nplans);
      for (dkt_plan *pl : pls)
        if (pl->dry || pl->simple || pl->per_level || pl->conforming)
          throw std::invalid_argument("a persistent batch runs tiled reference-semantics plans");
      std::vector<const void *> xs(x, x + nplans);
      std::vector<void *> ys(y, y + nplans);
      if (dtype == 0) {
        auto *g = new PersistGroup<double>(pls, xs, ys, nrhs, masked);
        b->error_flag = [g] { return g->error_flag(); };
        b->groups.emplace_back(g);
      } else {
        auto *g = new PersistGroup<float>(pls, xs, ys, nrhs, masked);
        b->error_flag = [g] { return g->error_flag(); };
        b->groups.emplace_back(g);
      }
      b->use_graph = (flags & DKT_BATCH_GRAPH) != 0;
      *out = b.release();
      return;
    }
    for (const auto &kv : by_kind_of(plans, nplans)) {
      std::vector<dkt_plan *> pls;
      std::vector<const void *> xs;
      std::vector<void *> ys;
      for (int k : kv.second) {
        pls.push_back(plans[k]);
        xs.push_back(x[k]);
        ys.push_back(y[k]);
      }
      const auto [conf, dim, p] = kv.first;
      const bool cm = (flags & DKT_BATCH_COLUMN_MAJOR) != 0;
      GroupBase *g = dtype == 0
                         ? make_group<double>(conf, dim, p, pls, xs, ys, nrhs, masked, {}, nullptr, max_cols, cm)
                         : make_group<float>(conf, dim, p, pls, xs, ys, nrhs, masked, {}, nullptr, max_cols, cm);
      b->groups.emplace_back(g);
    }
    b->fixes = dtype == 0 ? fix_all<double>(b->groups, nullptr) : fix_all<float>(b->groups, nullptr);
    b->use_graph = (flags & DKT_BATCH_GRAPH) != 0;
    *out = b.release();
  });
}

extern "C" int dkt_batch_run(dkt_batch *b, void *stream) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("null pointer");
    hipStream_t st = (hipStream_t)stream;
    if (!b->use_graph) {
      b->launch(st);
      return;
    }
    if (!st) throw std::invalid_argument("graph replay needs a non-default stream");
    if (!b->exec) {
      b->prepare_branches();
      HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        b->capture(st);
      } catch (...) {
        hipGraph_t dropped = nullptr;
        (void)hipStreamEndCapture(st, &dropped);
        if (dropped) (void)hipGraphDestroy(dropped);
        throw;
      }
      HIP_OK(hipStreamEndCapture(st, &b->graph));
      HIP_OK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIP_OK(hipGraphLaunch(b->exec, st));
  });
}

extern "C" int dkt_batch_destroy(dkt_batch *b) {
  delete b;
  return 0;
}

extern "C" int dkt_batch_status(dkt_batch *b, int *err) {
  return guarded([&] {
    if (!b || !err) throw std::invalid_argument("null pointer");
    *err = b->error_flag ? (int)b->error_flag() : 0;
  });
}
