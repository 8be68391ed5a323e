// Batched matvec over many independent meshes (no counterpart in the
// reference, which runs one fem::matvec per call, FEM/include/matvec.h:231).
//
// Plans of the same (dim, p) run as one launch: workgroup b finds its plan by
// a binary search over the plans' tile offsets and runs that plan's tile with
// the plan's maps and vectors (a device table of BatchPlan records), followed
// by one batched fix-up launch (per chunk of right-hand sides that fits the
// tiles' LDS).  With DKT_BATCH_GRAPH the launches of every
// (dim, p) group are captured once into a hipGraph and replayed.  Each plan's
// result is bitwise the same as its own dkt_plan_matvec: the tile work and
// the fix-up order are the plan's.
#include <algorithm>
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

template <typename T, int DIM, int P>
struct Group : GroupBase {
  // One table of plan records per column chunk (columns [q0, q0 + nq)).
  std::vector<std::unique_ptr<DevBuf<BatchPlan<T>>>> tables;
  DevBuf<int> tile_start, block_start;
  DevBuf<T> partial;
  int nplans = 0, ntiles = 0, nblocks = 0, threads = 0;
  std::vector<size_t> lds;
  std::vector<int> nqs;  // columns per chunk

  Group(const std::vector<dkt_plan *> &pls, const std::vector<const void *> &xs, const std::vector<void *> &ys,
        int nrhs, int masked, const std::vector<const int *> &cancel) {
    constexpr int NPE = Pow<P + 1, DIM>::v;
    const bool f64 = sizeof(T) == 8;
    auto dims_of = [](const dkt_plan *pl) {
      return TileDims{pl->max_nodes, pl->max_grp, pl->max_he, pl->max_rd, pl->max_cr, pl->tile_leaves, pl->max_hang};
    };
    // Columns per launch: the most every plan's tiles fit in LDS.
    int chunk = nrhs;
    for (dkt_plan *pl : pls) {
      if (tile_lds_bytes<T, NPE>(dims_of(pl), 1) > kLdsLimit) throw std::invalid_argument("tile does not fit in LDS");
      while (chunk > 1 && tile_lds_bytes<T, NPE>(dims_of(pl), chunk) > kLdsLimit) chunk--;
    }
    std::vector<int> ts{0}, bs{0};
    std::vector<size_t> poff;
    size_t ptot = 0;
    for (dkt_plan *pl : pls) {
      threads = std::max(threads, pl->tile_leaves);
      poff.push_back(ptot);
      ptot += pl->npartial * (size_t)chunk;
      ts.push_back(ts.back() + (int)pl->ntiles);
      bs.push_back(bs.back() + (int)((pl->nshared + 255) / 256));
    }
    partial.alloc(std::max<size_t>(1, ptot));
    for (int q0 = 0; q0 < nrhs; q0 += chunk) {
      const int nq = std::min(chunk, nrhs - q0);
      std::vector<BatchPlan<T>> hp(pls.size());
      size_t need = 0;
      for (size_t k = 0; k < pls.size(); k++) {
        dkt_plan *pl = pls[k];
        need = std::max(need, tile_lds_bytes<T, NPE>(dims_of(pl), nq));
        TileArgs &A = hp[k].A;
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
        A.ld = nrhs;
        A.dims = dims_of(pl);
        hp[k].kref = P == 2 ? (f64 ? (const T *)pl->kfac64.p : (const T *)pl->kfac32.p)
                            : (f64 ? (const T *)pl->kmat64.p : (const T *)pl->kmat32.p);
        hp[k].carry_w = f64 ? (const T *)pl->carry_w64.p : (const T *)pl->carry_w32.p;
        hp[k].x = (const T *)xs[k] + q0;
        hp[k].y = (T *)ys[k] + q0;
        hp[k].partial = partial.p + poff[k];
        hp[k].snode = pl->snode.p;
        hp[k].soff = pl->soff.p;
        hp[k].nshared = (int)pl->nshared;
      }
      tables.emplace_back(new DevBuf<BatchPlan<T>>());
      tables.back()->upload(hp);
      lds.push_back(need);
      nqs.push_back(nq);
    }
    nplans = (int)pls.size();
    ntiles = ts.back();
    nblocks = bs.back();
    tile_start.upload(ts);
    block_start.upload(bs);
  }

  void launch(hipStream_t st) override {
    for (size_t c = 0; c < tables.size(); c++) {
      if (ntiles && nqs[c] == 1)
        hipLaunchKernelGGL((tile_kernel_batch<T, DIM, P, true>), dim3(ntiles), dim3(threads), lds[c], st,
                           tables[c]->p, tile_start.p, nplans);
      else if (ntiles)
        hipLaunchKernelGGL((tile_kernel_batch<T, DIM, P, false>), dim3(ntiles), dim3(threads), lds[c], st,
                           tables[c]->p, tile_start.p, nplans);
      if (nblocks)
        hipLaunchKernelGGL((fixup_kernel_batch<T>), dim3(nblocks), dim3(256), 0, st, tables[c]->p, block_start.p,
                           nplans);
      HIP_OK(hipGetLastError());
    }
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

// A conforming plan of a batch.
template <typename T>
struct ConfBatchPlan {
  ConfArgs A;
  const T *kref, *term_w, *x;
  T *y, *partial;
};

// All tiles of a batch of same-(dim, p) conforming plans in one launch.
template <typename T, int DIM, int P>
__global__ void __launch_bounds__(kTileThreads) conf_tile_kernel_batch(const ConfBatchPlan<T> *__restrict__ plans,
                                                                       const int *__restrict__ tile_start,
                                                                       int nplans) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int k = find_range(tile_start, nplans, t);
  const ConfBatchPlan<T> &bp = plans[k];
  if (bp.A.cancel && *bp.A.cancel) return;
  conf_tile_body<T, DIM, P>(bp.A, t - tile_start[k], bp.kref, bp.term_w, bp.x, bp.y, bp.partial);
}

// Conforming plans of one (dim, p): one tile launch and one fix-up launch per
// column chunk that fits the tiles' LDS.  Each result is bitwise the plan's
// own matvec (same tile work, same fix-up order).
template <typename T, int DIM, int P>
struct ConfGroup : GroupBase {
  std::vector<std::unique_ptr<DevBuf<ConfBatchPlan<T>>>> tables;
  std::vector<std::unique_ptr<DevBuf<FixRec<T>>>> fixes;
  DevBuf<int> tile_start, block_start;
  DevBuf<T> partial;
  int nplans = 0, ntiles = 0, nblocks = 0, threads = 0;
  std::vector<size_t> lds;

  ConfGroup(const std::vector<dkt_plan *> &pls, const std::vector<const void *> &xs, const std::vector<void *> &ys,
            int nrhs, int masked, const std::vector<const int *> &cancel) {
    constexpr int NPE = Pow<P + 1, DIM>::v;
    const bool f64 = sizeof(T) == 8;
    auto lds_of = [](const dkt_plan *pl, int nq) {
      return conf_lds_bytes<T, NPE>(pl->max_nodes, pl->max_virt, pl->max_terms, nq);
    };
    int chunk = nrhs;
    for (dkt_plan *pl : pls) {
      if (lds_of(pl, 1) > kLdsLimit) throw std::invalid_argument("tile does not fit in LDS");
      while (chunk > 1 && lds_of(pl, chunk) > kLdsLimit) chunk--;
    }
    std::vector<int> ts{0}, bs{0};
    std::vector<size_t> poff;
    size_t ptot = 0;
    for (dkt_plan *pl : pls) {
      threads = std::max(threads, pl->tile_leaves);
      poff.push_back(ptot);
      ptot += pl->npartial * (size_t)chunk;
      ts.push_back(ts.back() + (int)pl->ntiles);
      bs.push_back(bs.back() + (int)((pl->nshared + 255) / 256));
    }
    partial.alloc(std::max<size_t>(1, ptot));
    for (int q0 = 0; q0 < nrhs; q0 += chunk) {
      const int nq = std::min(chunk, nrhs - q0);
      std::vector<ConfBatchPlan<T>> hp(pls.size());
      std::vector<FixRec<T>> hf(pls.size());
      size_t need = 0;
      for (size_t k = 0; k < pls.size(); k++) {
        dkt_plan *pl = pls[k];
        need = std::max(need, lds_of(pl, nq));
        ConfArgs &A = hp[k].A;
        A.tiles = pl->ctiles.p;
        A.tile_nodes = pl->tile_nodes.p;
        A.tile_out = pl->tile_out.p;
        A.leaf_loc = pl->leaf_loc.p;
        A.leaf_lev = pl->leaf_lev.p;
        A.virt_off = pl->virt_off.p;
        A.term_node = pl->term_node.p;
        A.term_virt = pl->term_virt.p;
        A.term_perm = pl->term_perm.p;
        A.dir_in = masked == 1 ? pl->dirichlet.p : nullptr;
        A.dir_out = masked ? pl->dirichlet.p : nullptr;
        A.cancel = cancel.empty() ? nullptr : cancel[k];
        A.op = pl->op;
        A.hp = pl->hp;
        A.nq = nq;
        A.ld = nrhs;
        A.max_points = pl->max_nodes;
        A.max_virt = pl->max_virt;
        A.max_terms = pl->max_terms;
        A.tile_leaves = pl->tile_leaves;
        hp[k].kref = P == 2 ? (f64 ? (const T *)pl->kfac64.p : (const T *)pl->kfac32.p)
                            : (f64 ? (const T *)pl->kmat64.p : (const T *)pl->kmat32.p);
        hp[k].term_w = f64 ? (const T *)pl->term_w64.p : (const T *)pl->term_w32.p;
        hp[k].x = (const T *)xs[k] + q0;
        hp[k].y = (T *)ys[k] + q0;
        hp[k].partial = partial.p + poff[k];
        hf[k] = FixRec<T>{pl->snode.p, pl->soff.p, partial.p + poff[k], (T *)ys[k] + q0, A.dir_out, A.cancel,
                          nq, nrhs, (int)pl->nshared};
      }
      tables.emplace_back(new DevBuf<ConfBatchPlan<T>>());
      tables.back()->upload(hp);
      fixes.emplace_back(new DevBuf<FixRec<T>>());
      fixes.back()->upload(hf);
      lds.push_back(need);
    }
    nplans = (int)pls.size();
    ntiles = ts.back();
    nblocks = bs.back();
    tile_start.upload(ts);
    block_start.upload(bs);
  }

  void launch(hipStream_t st) override {
    for (size_t c = 0; c < tables.size(); c++) {
      if (ntiles)
        hipLaunchKernelGGL((conf_tile_kernel_batch<T, DIM, P>), dim3(ntiles), dim3(threads), lds[c], st,
                           tables[c]->p, tile_start.p, nplans);
      if (nblocks)
        hipLaunchKernelGGL((fixup_kernel_rec<T>), dim3(nblocks), dim3(256), 0, st, fixes[c]->p, block_start.p,
                           nplans);
      HIP_OK(hipGetLastError());
    }
  }
};

template <typename T>
GroupBase *make_group(bool conforming, int dim, int p, const std::vector<dkt_plan *> &pls,
                      const std::vector<const void *> &xs, const std::vector<void *> &ys, int nrhs, int masked,
                      const std::vector<const int *> &cancel = {}) {
#define DKT_GROUP(D, PP)                                                                       \
  if (dim == D && p == PP) {                                                                   \
    if (conforming) return new ConfGroup<T, D, PP>(pls, xs, ys, nrhs, masked, cancel);         \
    return new Group<T, D, PP>(pls, xs, ys, nrhs, masked, cancel);                             \
  }
  DKT_GROUP(2, 1)
  DKT_GROUP(3, 1)
  DKT_GROUP(4, 1)
  DKT_GROUP(2, 2)
  DKT_GROUP(3, 2)
#undef DKT_GROUP
  throw std::invalid_argument("unsupported (dim, order)");
}

// Plans grouped by (conforming, dim, p): indices into the caller's list.
std::map<std::tuple<int, int, int>, std::vector<int>> by_kind_of(dkt_plan *const *plans, int nplans) {
  std::map<std::tuple<int, int, int>, std::vector<int>> by_kind;
  for (int k = 0; k < nplans; k++) {
    dkt_plan *pl = plans[k];
    if (pl->dry || pl->simple || pl->per_level)
      throw std::invalid_argument("batch plans must be tiled device plans (reference or conforming)");
    by_kind[{pl->conforming ? 1 : 0, pl->dim, pl->order}].push_back(k);
  }
  return by_kind;
}

struct GroupsLauncher : MatvecLauncher {
  std::vector<std::unique_ptr<GroupBase>> groups;
  void launch(hipStream_t st) override {
    for (auto &g : groups) g->launch(st);
  }
};

}  // namespace

std::unique_ptr<MatvecLauncher> dkt::batch_matvec(int dtype, const std::vector<dkt_plan *> &plans,
                                                  const std::vector<const void *> &x, const std::vector<void *> &y,
                                                  int mask, const std::vector<const int *> &cancel) {
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
    out->groups.emplace_back(dtype == 0 ? make_group<double>(conf != 0, dim, p, pls, xs, ys, 1, mask, cs)
                                        : make_group<float>(conf != 0, dim, p, pls, xs, ys, 1, mask, cs));
  }
  return out;
}

struct dkt_batch {
  std::vector<std::unique_ptr<GroupBase>> groups;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~dkt_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
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
    auto b = std::make_unique<dkt_batch>();
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
      GroupBase *g = dtype == 0 ? make_group<double>(conf != 0, dim, p, pls, xs, ys, nrhs, masked)
                                : make_group<float>(conf != 0, dim, p, pls, xs, ys, nrhs, masked);
      b->groups.emplace_back(g);
    }
    b->use_graph = (flags & DKT_BATCH_GRAPH) != 0;
    *out = b.release();
  });
}

extern "C" int dkt_batch_run(dkt_batch *b, void *stream) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("null pointer");
    hipStream_t st = (hipStream_t)stream;
    if (!b->use_graph) {
      for (auto &g : b->groups) g->launch(st);
      return;
    }
    if (!st) throw std::invalid_argument("graph replay needs a non-default stream");
    if (!b->exec) {
      HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        for (auto &g : b->groups) g->launch(st);
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
