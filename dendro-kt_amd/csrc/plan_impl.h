// The device plan object shared by the matvec (plan.hip) and the solvers (cg.hip).
#pragma once
#include <algorithm>
#include <memory>
#include <vector>

#include "../../include/dkt.h"
#include "plan_common.h"
#include "conf_kernel.h"

// Solver workspace: CG vectors and reduction partials, allocated on first use.
struct CgWorkspace {
  size_t n = 0;
  dkt::DevBuf<double> r, p, ap, d64, part64, scal, part32, scal32;  // part32/scal32: reductions of the FP32 solve
  dkt::DevBuf<float> r32, p32, ap32, x32, b32;
  dkt::DevBuf<int> flag;  // [0] cancel (converged), [1] iterations done
  dkt::DevBuf<double> bi, bipart, biscal;  // BiCGStab: r, r^, p, v, s, t; partials; scalars
  int *host = nullptr;    // pinned copy of flag
  void release() {
    for (auto *b : {&r, &p, &ap, &d64, &part64, &scal, &part32, &scal32, &bi, &bipart, &biscal}) b->release();
    for (auto *b : {&r32, &p32, &ap32, &x32, &b32}) b->release();
    flag.release();
    if (host) (void)hipHostFree(host);
    host = nullptr;
    n = 0;
  }
  CgWorkspace() = default;
  CgWorkspace(const CgWorkspace &) = delete;
  CgWorkspace &operator=(const CgWorkspace &) = delete;
  ~CgWorkspace() { release(); }
};

struct dkt_plan {
  int dim = 0, order = 0, npe = 0, op = 0, hp = 0, maxd = 0;
  size_t N = 0, L = 0, H = 0;
  bool per_level = false, simple = false, dry = false, conforming = false;
  dkt::DevBuf<uint8_t> dirichlet;
  dkt::DevBuf<double> kmat64;
  dkt::DevBuf<float> kmat32;
  // p = 2: B (3 x 3) and lambda of the sum-factorised element operators (tile kernels)
  dkt::DevBuf<double> kfac64;
  dkt::DevBuf<float> kfac32;
  // simple path
  dkt::DevBuf<int32_t> slot, hang_of, pslot, hc_off, hc_node, csr_src;
  dkt::DevBuf<uint8_t> level, child, hc_rank;
  dkt::DevBuf<double> hc_w64;
  dkt::DevBuf<float> hc_w32;
  dkt::DevBuf<int64_t> csr_off;
  size_t nbuf = 0;
  // tiled path
  int tile_leaves = dkt::kTileThreads;
  size_t ntiles = 0, nshared = 0, npartial = 0;
  int nw = 0;  // conforming: entries of the term weight table
  int max_nodes = 0, max_grp = 0, max_he = 0, max_rd = 0, max_cr = 0, max_virt = 0, max_terms = 0, max_hang = 0;
  size_t lds_bytes = 0;  // LDS per workgroup of the tiled kernel, FP64, one column
  dkt::DevBuf<dkt::TileHdr> tiles;
  dkt::DevBuf<int32_t> tile_nodes, tile_out, snode, soff;
  dkt::DevBuf<uint16_t> leaf_loc, group_ploc, readers;
  dkt::DevBuf<uint32_t> leaf_aux, carry_key;
  dkt::DevBuf<uint32_t> hentries;
  dkt::DevBuf<uint64_t> blocks;
  // conforming tiles
  dkt::DevBuf<dkt::ConfTileHdr> ctiles;
  dkt::DevBuf<uint16_t> virt_off;
  dkt::DevBuf<uint32_t> seg;  // node-sorted term segments: node | first term << 16
  dkt::DevBuf<uint32_t> term_pk, term_spk;  // point | weight index << 16, row order and sorted by node
  dkt::DevBuf<uint8_t> leaf_lev;
  dkt::DevBuf<double> term_w64;  // the distinct term weights (nw)
  dkt::DevBuf<float> term_w32;
  dkt::DevBuf<double> carry_w64;
  dkt::DevBuf<float> carry_w32;
  // scratch
  dkt::DevBuf<double> scratch64;
  dkt::DevBuf<float> scratch32;
  double map_bytes = 0;
  // Tile statistics: mean readers, max readers, mean carry terms, max carry terms,
  // mean groups, mean nodes, hanging-transpose entries, mean cone length,
  // mean and max longest replay of tiles with readers.
  std::vector<double> stats = std::vector<double>(10, 0.0);
  // Optional event timing (dkt_plan_profile): (start, after tile kernel, end) per launch.
  bool profiling = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  hipEvent_t next_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      HIP_OK(hipEventCreate(&e));
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }
  CgWorkspace cg;
  ~dkt_plan() {
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
  }
};

namespace dkt {
// Device scratch supplied by the caller instead of the plan's own buffer.
struct ExtScratch {
  void *p;
  size_t bytes;
};
// y = A x on `st` (dtype 0 f64, 1 f32; mask as dkt_plan_matvec's `masked`).  When `cancel` is non-null and *cancel
// is nonzero at launch time, the kernels return without touching y.  With `ext`, the matvec's partial sums (or
// the simple path's contributions) go to ext->p, and a matvec that needs more than ext->bytes is refused; no
// device memory is allocated then.
void plan_apply(dkt_plan *pl, int dtype, const void *x, void *y, int nrhs, int mask, hipStream_t st,
                const int *cancel, const ExtScratch *ext = nullptr);

// Grow-only device memory for launch tables that are rebuilt often (the CG
// engine rebuilds its batched matvecs whenever a solve joins or leaves):
// reset() starts a rebuild, take() hands out 256-byte aligned pieces, upload()
// copies a host vector into one with hipMemcpyAsync on the rebuild's stream.
// A rebuild that outgrows the arena adds a block; the next reset() replaces
// the blocks by one of their total size, so after the first few rebuilds no
// device memory is allocated or freed (hipFree waits for the whole device,
// other streams included).  The host copies stay alive until the next reset();
// the caller must have synchronized the stream before it resets.
struct DevArena {
  struct Block {
    DevBuf<unsigned char> buf;
    size_t used = 0;
  };
  std::vector<std::unique_ptr<Block>> blocks;
  std::vector<std::vector<unsigned char>> staging;
  hipStream_t stream = nullptr;

  void reset(hipStream_t st) {
    stream = st;
    staging.clear();
    if (blocks.size() > 1) {
      size_t total = 0;
      for (auto &b : blocks) total += b->buf.n;
      blocks.clear();
      add_block(total);
    }
    for (auto &b : blocks) b->used = 0;
  }
  void *take(size_t bytes) {
    bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
    Block *b = blocks.empty() ? nullptr : blocks.back().get();
    if (!b || b->used + bytes > b->buf.n) b = add_block(std::max(bytes, blocks.empty() ? (size_t)1 << 20 : 2 * b->buf.n));
    void *p = b->buf.p + b->used;
    b->used += bytes;
    return p;
  }
  template <typename T>
  T *upload(const std::vector<T> &v) {
    T *p = (T *)take(v.size() * sizeof(T));
    if (v.empty()) return p;
    staging.emplace_back((const unsigned char *)v.data(), (const unsigned char *)(v.data() + v.size()));
    HIP_OK(hipMemcpyAsync(p, staging.back().data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream));
    return p;
  }

 private:
  Block *add_block(size_t bytes) {
    blocks.emplace_back(new Block());
    blocks.back()->buf.alloc(bytes);
    return blocks.back().get();
  }
};

// Matvecs y_k = A_k x_k (dtype 0 f64, 1 f32, one column) of many tiled plans,
// reference or conforming, one launch per (kind, dim, p) (batch.hip); plan
// k's kernels return at once while *cancel[k] is nonzero.  Each result is
// bitwise the plan's own matvec.  With an arena, the launch tables and
// partial sums live in it (nothing is allocated once it is large enough) and
// the launcher is valid until the arena's next reset().
struct MatvecLauncher {
  virtual ~MatvecLauncher() = default;
  virtual void launch(hipStream_t st) = 0;
};
std::unique_ptr<MatvecLauncher> batch_matvec(int dtype, const std::vector<dkt_plan *> &plans,
                                             const std::vector<const void *> &x, const std::vector<void *> &y,
                                             int mask, const std::vector<const int *> &cancel,
                                             DevArena *arena = nullptr);
}  // namespace dkt
