// Host mesh layer (see mesh.h).  Citations are to the reference Dendro-KT
// sources whose behaviour each function reproduces.
#include "mesh.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <functional>
#include <parallel/algorithm>
#include <unordered_map>
#include <stdexcept>
#include <cstdio>

namespace dkt {

// ---------------------------------------------------------------------------
// Space-filling curve tables.
//
// Hilbert: the K-dimensional harmonious Hilbert curve of Haverkort (2012,
// "Harmonious Hilbert curves and other extradimensional space-filling curves",
// p.23), the curve the reference tabulates (gen/RotationTableHilbert.cpp,
// src/KDhcurvedata_DATA.cpp).  An orientation is a permutation `a` of the K
// axes plus a reflection bit vector `m`; bit strings are written with
// Haverkort's index 0 as the most significant of K bits.  Orientations are
// numbered in first-visit order of a depth-first refinement from the identity,
// which yields the same row numbering as the reference tables.
// ---------------------------------------------------------------------------
namespace {

struct Orient {
  std::array<int, 4> a{};
  unsigned m = 0;
  int K = 0;
  unsigned permute(unsigned bits) const {
    unsigned out = 0;
    for (int i = 0; i < K; i++) out = (out << 1) | ((bits >> (K - 1 - a[i])) & 1u);
    return out;
  }
  unsigned apply(unsigned loc) const { return permute(loc) ^ m; }
  // (this) o (q): permutations compose as r.a[i] = q.a[a[i]], reflections as m ^ permute(q.m).
  Orient compose(const Orient &q) const {
    Orient r;
    r.K = K;
    for (int i = 0; i < K; i++) r.a[i] = q.a[a[i]];
    r.m = m ^ permute(q.m);
    return r;
  }
  unsigned code() const {
    unsigned c = m;
    for (int i = 0; i < K; i++) c = c * 8 + (unsigned)a[i];
    return c;
  }
};

// Haverkort's refinement operator: location (Gray code) and orientation of the
// child at curve rank r.
void refine(int K, int r, unsigned &loc, Orient &o) {
  const unsigned c = (unsigned)(r ^ (r >> 1));
  loc = c;
  o.K = K;
  if (r == 0) {
    o.m = 0;
  } else {
    const unsigned cm1 = (unsigned)((r - 1) ^ ((r - 1) >> 1));
    o.m = (cm1 & ~1u) | ((~c) & 1u);
  }
  const int endr = r & 1;
  int front = 0;
  for (int i = 0; i < K; i++) front += (((r >> (K - 1 - i)) & 1) != endr);
  int L = front - 1, R = K - 1;
  for (int i = 0; i < K; i++) {
    if (((r >> (K - 1 - i)) & 1) != endr)
      o.a[i] = L--;
    else
      o.a[i] = R--;
  }
}

SfcTables build_hilbert(int K) {
  const int nch = 1 << K;
  std::vector<unsigned> locs(nch);
  std::vector<Orient> kids(nch);
  for (int r = 0; r < nch; r++) refine(K, r, locs[r], kids[r]);

  Orient id;
  id.K = K;
  for (int i = 0; i < K; i++) id.a[i] = i;
  id.m = 0;

  std::map<unsigned, int> rank;
  std::vector<Orient> byRank;
  // Depth-first, children in curve order; numbering on first visit.
  std::vector<std::pair<Orient, int>> stack;
  rank[id.code()] = 0;
  byRank.push_back(id);
  stack.push_back({id, 0});
  while (!stack.empty()) {
    auto &top = stack.back();
    if (top.second >= nch) {
      stack.pop_back();
      continue;
    }
    const Orient c = top.first.compose(kids[top.second]);
    top.second++;
    if (rank.find(c.code()) == rank.end()) {
      rank[c.code()] = (int)byRank.size();
      byRank.push_back(c);
      stack.push_back({c, 0});
    }
  }
  SfcTables t;
  t.dim = K;
  t.nch = nch;
  t.nrot = (int)byRank.size();
  t.perm.assign((size_t)t.nrot * nch, 0);
  t.inv.assign((size_t)t.nrot * nch, 0);
  t.child.assign((size_t)t.nrot * nch, 0);
  for (int p = 0; p < t.nrot; p++) {
    const Orient &po = byRank[p];
    for (int cs = 0; cs < nch; cs++) {
      const unsigned mort = po.apply(locs[cs]);
      const Orient co = po.compose(kids[cs]);
      t.perm[p * nch + cs] = (uint8_t)mort;
      t.inv[p * nch + mort] = (uint8_t)cs;
      t.child[p * nch + mort] = rank.at(co.code());
    }
  }
  return t;
}

SfcTables build_morton(int K) {
  SfcTables t;
  t.dim = K;
  t.nch = 1 << K;
  t.nrot = 1;
  t.perm.resize(t.nch);
  t.inv.resize(t.nch);
  t.child.assign(t.nch, 0);
  for (int c = 0; c < t.nch; c++) t.perm[c] = t.inv[c] = (uint8_t)c;
  return t;
}

}  // namespace

const SfcTables &sfc_tables(int dim, int curve) {
  static std::mutex mu;
  static std::map<int, SfcTables> cache;
  if (dim < 1 || dim > 4) throw std::invalid_argument("dim must be 1..4");
  std::lock_guard<std::mutex> g(mu);
  const int key = dim * 2 + (curve == CURVE_HILBERT ? 1 : 0);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  SfcTables t = (curve == CURVE_HILBERT && dim >= 2) ? build_hilbert(dim) : build_morton(dim);
  return cache.emplace(key, std::move(t)).first->second;
}

// ---------------------------------------------------------------------------
// Keys and sorting.
// ---------------------------------------------------------------------------
static inline unsigned morton_digit(const Oct &o, int dim, int maxd, int lev) {
  unsigned m = 0;
  const int sh = maxd - lev;
  for (int d = 0; d < dim; d++) m |= ((o.x[d] >> sh) & 1u) << d;
  return m;
}

u128 sfc_key(const MeshParams &mp, const Oct &o) {
  const SfcTables &t = sfc_tables(mp.dim, mp.curve);
  int rot = 0;
  u128 key = 0;
  const int lev = std::min((int)o.lev, mp.maxd);
  for (int l = 1; l <= lev; l++) {
    const unsigned m = morton_digit(o, mp.dim, mp.maxd, l);
    key = (key << mp.dim) | t.inv[rot * t.nch + m];
    rot = t.child[rot * t.nch + m];
  }
  // the levels below the octant's own contribute zero digits
  if (lev < mp.maxd) key <<= (unsigned)(mp.dim * (mp.maxd - lev));
  return key;
}

namespace {
struct KeyedIdx {
  u128 key;
  u32 lev;
  u32 idx;
};
inline bool keyed_less(const KeyedIdx &a, const KeyedIdx &b) {
  if (a.key != b.key) return a.key < b.key;
  if (a.lev != b.lev) return a.lev < b.lev;
  return a.idx < b.idx;
}
inline bool is_ancestor(const Oct &a, const Oct &b, int dim, int maxd) {
  if (a.lev >= b.lev) return false;
  const int sh = maxd - (int)a.lev;
  for (int d = 0; d < dim; d++)
    if (sh < 32 ? ((a.x[d] >> sh) != (b.x[d] >> sh)) : false) return false;
  return true;
}
inline bool same_oct(const Oct &a, const Oct &b, int dim) {
  if (a.lev != b.lev) return false;
  for (int d = 0; d < dim; d++)
    if (a.x[d] != b.x[d]) return false;
  return true;
}

std::vector<KeyedIdx> keyed(const MeshParams &mp, const std::vector<Oct> &octs) {
  std::vector<KeyedIdx> k(octs.size());
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < octs.size(); i++) k[i] = {sfc_key(mp, octs[i]), octs[i].lev, (u32)i};
  __gnu_parallel::sort(k.begin(), k.end(), keyed_less);
  return k;
}
}  // namespace

void sort_octants(const MeshParams &mp, std::vector<Oct> &octs, bool remove_ancestors) {
  auto k = keyed(mp, octs);
  std::vector<Oct> out;
  out.reserve(octs.size());
  for (size_t i = 0; i < k.size(); i++) {
    const Oct &cur = octs[k[i].idx];
    if (i + 1 < k.size()) {
      const Oct &nx = octs[k[i + 1].idx];
      if (same_oct(cur, nx, mp.dim)) continue;
      if (remove_ancestors && is_ancestor(cur, nx, mp.dim, mp.maxd)) continue;
    }
    out.push_back(cur);
  }
  octs.swap(out);
}

// ---------------------------------------------------------------------------
// Tree construction (locTreeConstruction) and balancing (distTreeBalancing).
// ---------------------------------------------------------------------------
namespace {
struct Builder {
  const MeshParams &mp;
  const SfcTables &t;
  const std::vector<KeyedIdx> &k;
  size_t max_pts;
  std::vector<Oct> &out;

  void rec(size_t b, size_t e, const Oct &cell, int rot) {
    const int l = (int)cell.lev + 1;  // level of the children
    while (b < e && (int)k[b].lev < l) b++;  // the cell itself / ancestors precede in pre-order
    const int sh = mp.dim * (mp.maxd - l);
    const u32 len = 1u << (mp.maxd - l);
    size_t s = b;
    for (int cs = 0; cs < t.nch; cs++) {
      size_t f = s;
      while (f < e && (unsigned)((k[f].key >> sh) & (u128)(t.nch - 1)) == (unsigned)cs) f++;
      const unsigned m = t.perm[rot * t.nch + cs];
      Oct ch;
      ch.lev = (u32)l;
      for (int d = 0; d < 4; d++) ch.x[d] = d < mp.dim ? cell.x[d] + (((m >> d) & 1u) ? len : 0u) : 0u;
      if (l < mp.maxd && f - s > max_pts)
        rec(s, f, ch, t.child[rot * t.nch + m]);
      else
        out.push_back(ch);
      s = f;
    }
  }
};
}  // namespace

std::vector<Oct> construct_tree(const MeshParams &mp, std::vector<Oct> &pts, size_t max_pts) {
  const SfcTables &t = sfc_tables(mp.dim, mp.curve);
  auto k = keyed(mp, pts);
  // The first levels of the recursion are expanded here, in visiting order,
  // into independent subtrees (a leaf is a subtree of one octant); the
  // subtrees are built in parallel and concatenated in that order, which is
  // the order of the sequential recursion.
  struct Task {
    size_t b, e;
    Oct cell;
    int rot;
    bool leaf;
  };
  std::vector<Task> tasks;
  const int split_levels = mp.dim >= 4 ? 2 : 3;
  std::function<void(size_t, size_t, const Oct &, int)> expand = [&](size_t b, size_t e, const Oct &cell, int rot) {
    const int l = (int)cell.lev + 1;
    while (b < e && (int)k[b].lev < l) b++;
    const int sh = mp.dim * (mp.maxd - l);
    const u32 len = 1u << (mp.maxd - l);
    size_t s = b;
    for (int cs = 0; cs < t.nch; cs++) {
      size_t f = s;
      while (f < e && (unsigned)((k[f].key >> sh) & (u128)(t.nch - 1)) == (unsigned)cs) f++;
      const unsigned m = t.perm[rot * t.nch + cs];
      Oct ch;
      ch.lev = (u32)l;
      for (int d = 0; d < 4; d++) ch.x[d] = d < mp.dim ? cell.x[d] + (((m >> d) & 1u) ? len : 0u) : 0u;
      const int crot = t.child[rot * t.nch + m];
      if (l < mp.maxd && f - s > max_pts) {
        if (l < split_levels)
          expand(s, f, ch, crot);
        else
          tasks.push_back({s, f, ch, crot, false});
      } else {
        tasks.push_back({s, f, ch, crot, true});
      }
      s = f;
    }
  };
  Oct root{};
  root.lev = 0;
  expand(0, k.size(), root, 0);
  std::vector<std::vector<Oct>> part(tasks.size());
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t i = 0; i < tasks.size(); i++) {
    const Task &tk = tasks[i];
    if (tk.leaf) {
      part[i].push_back(tk.cell);
    } else {
      Builder bl{mp, t, k, max_pts, part[i]};
      bl.rec(tk.b, tk.e, tk.cell, tk.rot);
    }
  }
  size_t total = 0;
  for (const auto &v : part) total += v.size();
  std::vector<Oct> out;
  out.reserve(total);
  for (auto &v : part) {
    out.insert(out.end(), v.begin(), v.end());
    std::vector<Oct>().swap(v);
  }
  return out;
}

namespace {
// Distinct octants of v, all of level lev (any order), through sorted keys of
// their coordinates (maxd + 1 bits each).  Entries with lev == kSkip are
// dropped.
constexpr u32 kSkip = 0xFFFFFFFFu;
template <typename K>
void dedupe_octs_t(std::vector<Oct> &v, int dim, int maxd, u32 lev) {
  const int cb = maxd + 1;
  std::vector<K> k(v.size());
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < v.size(); i++) {
    if (v[i].lev == kSkip) {
      k[i] = ~(K)0;  // above every key: coordinates never have all cb bits set
      continue;
    }
    K key = 0;
    for (int d = 0; d < dim; d++) key = (key << cb) | (K)v[i].x[d];
    k[i] = key;
  }
  __gnu_parallel::sort(k.begin(), k.end());
  k.erase(std::unique(k.begin(), k.end()), k.end());
  while (!k.empty() && k.back() == ~(K)0) k.pop_back();
  const K cmask = ((K)1 << cb) - 1;
  v.resize(k.size());
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < k.size(); i++) {
    K key = k[i];
    Oct o{};
    o.lev = lev;
    for (int d = dim - 1; d >= 0; d--) {
      o.x[d] = (u32)(key & cmask);
      key >>= cb;
    }
    v[i] = o;
  }
}
void dedupe_octs(std::vector<Oct> &v, int dim, int maxd, u32 lev) {
  if (dim * (maxd + 1) <= 64)
    dedupe_octs_t<u64>(v, dim, maxd, lev);
  else
    dedupe_octs_t<u128>(v, dim, maxd, lev);
}
}  // namespace

std::vector<Oct> balance_tree(const MeshParams &mp, std::vector<Oct> &pts, size_t max_pts) {
  std::vector<Oct> t0 = construct_tree(mp, pts, max_pts);
  // propagateNeighbours (src/tsort.cpp:777-820): bottom-up, every octant adds its
  // parent and the parent's in-domain neighbours (all 3^dim-1) at the parent level.
  std::vector<std::vector<Oct>> lv(mp.maxd + 1);
  for (const Oct &o : t0) lv[o.lev].push_back(o);
  const u32 top = (u32)1 << mp.maxd;
  int nbr = 1;
  for (int d = 0; d < mp.dim; d++) nbr *= 3;
  for (int l = mp.maxd; l > 0; l--) {
    const u32 plen = 1u << (mp.maxd - (l - 1));
    std::vector<Oct> par(lv[l].size());
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < lv[l].size(); i++) {
      Oct p = lv[l][i];
      p.lev = (u32)(l - 1);
      for (int d = 0; d < mp.dim; d++) p.x[d] &= ~(plen - 1);
      par[i] = p;
    }
    dedupe_octs(par, mp.dim, mp.maxd, (u32)(l - 1));
    std::vector<Oct> &dst = lv[l - 1];
    const size_t base = dst.size();
    dst.resize(base + par.size() * (size_t)nbr);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < par.size(); i++) {
      const Oct &p = par[i];
      for (int code = 0; code < nbr; code++) {
        Oct q = p;
        int c = code;
        bool ok = true;
        for (int d = 0; d < mp.dim; d++) {
          const int off = c % 3 - 1;
          c /= 3;
          if (off < 0) {
            if (p.x[d] < plen) ok = false;
            else q.x[d] = p.x[d] - plen;
          } else if (off > 0) {
            if (p.x[d] >= top - plen) ok = false;
            else q.x[d] = p.x[d] + plen;
          }
        }
        if (!ok) q.lev = kSkip;
        dst[base + i * nbr + code] = q;
      }
    }
    dedupe_octs(dst, mp.dim, mp.maxd, (u32)(l - 1));
  }
  std::vector<Oct> seeds;
  size_t tot = 0;
  for (auto &v : lv) tot += v.size();
  seeds.reserve(tot);
  for (auto &v : lv) seeds.insert(seeds.end(), v.begin(), v.end());
  lv.clear();
  return construct_tree(mp, seeds, 1);
}

// ---------------------------------------------------------------------------
// CG node enumeration (SFC_NodeSort::countCGNodes, include/nsort.tcc:879-1313,
// single-rank compaction :519-533; DA::construct src/oda.cpp:90-107).
// ---------------------------------------------------------------------------
namespace {
struct Inst {
  u128 pos;   // packed coordinates
  u32 lev;
  u32 elem;
};

inline bool on_boundary(const Oct &o, int dim, int maxd) {
  const u32 mask = (1u << maxd) - 1u;
  for (int d = 0; d < dim; d++)
    if ((o.x[d] & mask) == 0) return true;
  return false;
}

// Number of axes along which the point is interior to its cell at level lev.
inline int cell_dim(const Oct &o, int dim, int maxd, int lev) {
  const u32 mask = (1u << (maxd - lev)) - 1u;
  int c = 0;
  for (int d = 0; d < dim; d++) c += (o.x[d] & mask) != 0;
  return c;
}

// Full point key: child_sfc digits of the point's coordinate bits for levels
// 1..maxd (rotation chain from the root), as locTreeSortAsPoints orders points.
inline u128 point_key(const SfcTables &t, const Oct &o, int dim, int maxd) {
  int rot = 0;
  u128 key = 0;
  for (int l = 1; l <= maxd; l++) {
    const unsigned m = morton_digit(o, dim, maxd, l);
    key = (key << dim) | t.inv[rot * t.nch + m];
    rot = t.child[rot * t.nch + m];
  }
  return key;
}

struct InteriorKey {
  u128 key;
  int lc;   // level of the finest open container (getFinestOpenContainer)
  int hp;   // first incident hyperplane at level lc+1 (get_firstIncidentHyperplane)
  u32 idx;
};
}  // namespace

// A 256-bit key for dims whose coordinates and level do not fit 128 bits
// (4D with maxdepth 30): the operations resolve_instances uses.
struct WideKey {
  u128 hi = 0, lo = 0;
  WideKey() = default;
  WideKey(u128 v) : hi(0), lo(v) {}
  friend WideKey operator<<(WideKey a, int s) {
    if (s == 0) return a;
    WideKey r;
    if (s >= 128) {
      r.hi = a.lo << (s - 128);
    } else {
      r.hi = (a.hi << s) | (a.lo >> (128 - s));
      r.lo = a.lo << s;
    }
    return r;
  }
  friend WideKey operator>>(WideKey a, int s) {
    if (s == 0) return a;
    WideKey r;
    if (s >= 128) {
      r.lo = a.hi >> (s - 128);
    } else {
      r.lo = (a.lo >> s) | (a.hi << (128 - s));
      r.hi = a.hi >> s;
    }
    return r;
  }
  friend WideKey operator|(WideKey a, WideKey b) { a.hi |= b.hi; a.lo |= b.lo; return a; }
  friend WideKey operator&(WideKey a, WideKey b) { a.hi &= b.hi; a.lo &= b.lo; return a; }
  friend WideKey operator-(WideKey a, WideKey b) {
    WideKey r;
    r.lo = a.lo - b.lo;
    r.hi = a.hi - b.hi - (a.lo < b.lo ? 1 : 0);
    return r;
  }
  friend bool operator==(const WideKey &a, const WideKey &b) { return a.hi == b.hi && a.lo == b.lo; }
  friend bool operator!=(const WideKey &a, const WideKey &b) { return !(a == b); }
  friend bool operator<(const WideKey &a, const WideKey &b) { return a.hi != b.hi ? a.hi < b.hi : a.lo < b.lo; }
  WideKey &operator>>=(int s) { return *this = *this >> s; }
  explicit operator u32() const { return (u32)lo; }
};

// Exterior node instances of every element, sorted by location and level;
// per location the coarsest instance is kept when the location is on the
// domain boundary or carries several levels, else when all 2^(dim - cell dim)
// incident cells produced it (no hanging node).  K holds the key
// (coordinates with maxd + 1 bits each, then 5 bits of level).
template <typename K>
std::vector<Oct> resolve_instances(const MeshParams &mp, const std::vector<Oct> &elems, const std::vector<int> &extIdx) {
  const int dim = mp.dim, maxd = mp.maxd, p = mp.order, cb = maxd + 1;
  const int next = (int)extIdx.size();
  const size_t total = elems.size() * (size_t)next;
  std::unique_ptr<K[]> keys(new K[total]);  // filled in parallel (first touch)
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < elems.size(); e++) {
    const Oct &el = elems[e];
    const u64 len = (u64)1 << (maxd - el.lev);
    for (int j = 0; j < next; j++) {
      int rr = extIdx[j];
      K k = 0;
      for (int d = 0; d < dim; d++) {
        const int i = rr % (p + 1);
        rr /= (p + 1);
        k = (k << cb) | (K)(u32)(len * (u64)i / (u64)p + el.x[d]);
      }
      keys[e * next + j] = (k << 5) | (K)el.lev;
    }
  }
  __gnu_parallel::sort(keys.get(), keys.get() + total);
  const K cmask = ((K)1 << cb) - 1;
  auto decode = [&](K k) {
    Oct o{};
    o.lev = (u32)(k & 31);
    k >>= 5;
    for (int d = dim - 1; d >= 0; d--) {
      o.x[d] = (u32)(k & cmask);
      k >>= cb;
    }
    return o;
  };
  std::vector<Oct> sel;
  sel.reserve(total / 4);
  for (size_t b = 0; b < total;) {
    size_t e = b + 1;
    const K pos = keys[b] >> 5;
    while (e < total && (keys[e] >> 5) == pos) e++;
    const Oct first = decode(keys[b]);  // coarsest instance (sorted by level)
    const bool mixed = (u32)(keys[e - 1] & 31) != first.lev;
    bool keep;
    if (on_boundary(first, dim, maxd) || mixed) {
      keep = true;
    } else {
      const int cd = cell_dim(first, dim, maxd, first.lev);
      keep = (e - b) == ((size_t)1 << (dim - cd));
    }
    if (keep) sel.push_back(first);
    b = e;
  }
  return sel;
}

NodeSet enumerate_nodes(const MeshParams &mp, const std::vector<Oct> &elems) {
  const int dim = mp.dim, maxd = mp.maxd, p = mp.order;
  if (p < 1) throw std::invalid_argument("order must be >= 1");
  int npe = 1, nint = 1;
  for (int d = 0; d < dim; d++) npe *= (p + 1), nint *= (p - 1);
  const int next = npe - nint;

  // 1. Exterior node instances (Element::appendExteriorNodes, nsort.tcc:360-389),
  //    as sortable keys: coordinates (maxd + 1 bits each) then the level.
  std::vector<int> extIdx;
  for (int r = 0; r < npe; r++) {
    bool ext = false;
    int rr = r;
    for (int d = 0; d < dim; d++) {
      const int i = rr % (p + 1);
      rr /= (p + 1);
      if (i == 0 || i == p) ext = true;
    }
    if (ext) extIdx.push_back(r);
  }
  if (p > 2) throw std::invalid_argument("orders above 2 are not supported");
  const int key_bits = dim * (maxd + 1) + 5;
  // 2. Resolve duplicates and hanging nodes per location
  //    (resolveInterface_lowOrder nsort.tcc:1266-1313; boundary: coarsest wins, :913-926).
  const std::vector<Oct> inst = key_bits <= 64    ? resolve_instances<u64>(mp, elems, extIdx)
                                : key_bits <= 128 ? resolve_instances<u128>(mp, elems, extIdx)
                                                  : resolve_instances<WideKey>(mp, elems, extIdx);
  // 3. Order: interior nodes by (container post-order, hyperplane, point key);
  //    boundary nodes after them, by (root bucket, point key).
  const SfcTables &t = sfc_tables(dim, mp.curve);
  const size_t nsel = inst.size();
  std::vector<uint8_t> is_bdy(nsel);
  std::vector<InteriorKey> ik_all(nsel);
  std::vector<std::pair<std::pair<int, u128>, u32>> bk_all(nsel);
#pragma omp parallel for schedule(static)
  for (size_t si = 0; si < nsel; si++) {
    const u32 s = (u32)si;
    const Oct &o = inst[s];
    is_bdy[s] = on_boundary(o, dim, maxd);
    if (is_bdy[s]) {
      unsigned m0 = 0;
      for (int d = 0; d < dim; d++) m0 |= ((o.x[d] >> maxd) & 1u) << d;
      bk_all[s] = {{(int)t.inv[m0], point_key(t, o, dim, maxd)}, s};
    } else {
      int maxlow = 0;
      for (int d = 0; d < dim; d++) {
        const int low = __builtin_ctz(o.x[d]);
        maxlow = std::max(maxlow, low);
      }
      const int lc = maxd - maxlow - 1;
      const u32 hmask = (1u << (maxd - (lc + 1))) - 1u;
      int hp = dim;
      for (int d = 0; d < dim; d++)
        if ((o.x[d] & hmask) == 0) {
          hp = d;
          break;
        }
      ik_all[s] = {point_key(t, o, dim, maxd), lc, hp, s};
    }
  }
  std::vector<InteriorKey> ik;
  std::vector<std::pair<std::pair<int, u128>, u32>> bk;
  for (size_t s = 0; s < nsel; s++) {
    if (is_bdy[s])
      bk.push_back(bk_all[s]);
    else
      ik.push_back(ik_all[s]);
  }
  std::vector<InteriorKey>().swap(ik_all);
  std::vector<std::pair<std::pair<int, u128>, u32>>().swap(bk_all);
  const int D = dim, MD = maxd;
  __gnu_parallel::sort(ik.begin(), ik.end(), [D, MD](const InteriorKey &a, const InteriorKey &b) {
    const int l = std::min(a.lc, b.lc);
    const int sh = D * (MD - l);
    const u128 pa = sh >= 128 ? 0 : (a.key >> sh), pb = sh >= 128 ? 0 : (b.key >> sh);
    if (pa != pb) return pa < pb;
    if (a.lc != b.lc) return a.lc > b.lc;
    if (a.hp != b.hp) return a.hp < b.hp;
    return a.key < b.key;
  });
  __gnu_parallel::sort(bk.begin(), bk.end());

  NodeSet ns;
  ns.nodes.reserve(nsel + elems.size() * (size_t)nint);
  for (const auto &k : ik) ns.nodes.push_back(inst[k.idx]);
  for (const auto &k : bk) ns.nodes.push_back(inst[k.second]);
  ns.n_exterior = ns.nodes.size();

  // 4. Element-interior nodes, appended in element order (oda.cpp:100-101).
  if (nint > 0) {
    for (const Oct &el : elems) {
      const u64 len = (u64)1 << (maxd - el.lev);
      for (int j = 0; j < nint; j++) {
        int rr = j;
        Oct o{};
        o.lev = el.lev;
        for (int d = 0; d < dim; d++) {
          const int i = rr % (p - 1);
          rr /= (p - 1);
          o.x[d] = (u32)(len * (u64)(i + 1) / (u64)p + el.x[d]);
        }
        ns.nodes.push_back(o);
      }
    }
  }
  for (size_t i = 0; i < ns.nodes.size(); i++)
    if (on_boundary(ns.nodes[i], dim, maxd)) ns.bdy.push_back((u32)i);
  return ns;
}

// ---------------------------------------------------------------------------
// Traversal flattening (fem::matvec_rec, FEM/include/matvec.h:243-548).
// ---------------------------------------------------------------------------
void interp_1d(int p, int c, double *ip) {
  // Parent nodes at t_k = k/p on [0,1]; child c covers [c/2, (c+1)/2].
  const int n = p + 1;
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) {
      const double x = 0.5 * ((double)c + (double)i / (double)p);
      double l = 1.0;
      for (int j = 0; j < n; j++)
        if (j != k) l *= (x - (double)j / p) / ((double)k / p - (double)j / p);
      ip[k * n + i] = l;
    }
}

namespace {
struct Flattener {
  const MeshParams &mp;
  const SfcTables &t;
  const std::vector<Oct> &nodes;
  Traversal &tr;
  int npe;
  std::vector<std::vector<int32_t>> buf;  // per level: duplicated child buckets
  std::vector<int32_t> scratch;
  double ip[2][9];
  size_t stale_needed = 0;
  std::vector<std::vector<std::vector<std::pair<int, double>>>> wnz;  // [child][r] -> (s, weight != 0)

  void init_weights() {
    const int n1 = mp.order + 1;
    const int nch = 1 << mp.dim;
    wnz.assign(nch, std::vector<std::vector<std::pair<int, double>>>(npe));
    for (int c = 0; c < nch; c++)
      for (int r = 0; r < npe; r++)
        for (int s = 0; s < npe; s++) {
          double w = 1.0;
          int rr = r, ss = s;
          for (int d = 0; d < mp.dim; d++) {
            w *= ip[(c >> d) & 1][(ss % n1) * n1 + (rr % n1)];
            rr /= n1;
            ss /= n1;
          }
          if (w != 0.0) wnz[c][r].push_back({s, w});
        }
  }

  // Lexicographic rank of node x on cell (TNPoint::get_lexNodeRank, nsort.tcc:283-303).
  int lex_rank(const Oct &x, const Oct &cell) const {
    const u64 len = (u64)1 << (mp.maxd - cell.lev);
    int rank = 0, stride = 1;
    for (int d = 0; d < mp.dim; d++) {
      const u64 off = (u64)x.x[d] - (u64)cell.x[d];
      if ((u64)x.x[d] < (u64)cell.x[d] || off > len || (off * mp.order) % len != 0)
        throw std::runtime_error("node is not on the lattice of its traversal cell");
      const int i = (int)(off * mp.order / len);
      rank += i * stride;
      stride *= (mp.order + 1);
    }
    return rank;
  }

  void leaf(const Oct &cell, const int32_t *ids, size_t n, const int32_t *pids, size_t pn, const Oct &parent) {
    const size_t L = tr.leaves.size();
    tr.leaves.push_back(cell);
    const unsigned cm = morton_digit(cell, mp.dim, mp.maxd, cell.lev);
    tr.child.push_back((uint8_t)cm);
    tr.slot.resize((L + 1) * npe, -1);
    int32_t *sl = &tr.slot[L * npe];
    for (size_t i = 0; i < n; i++) sl[lex_rank(nodes[ids[i]], cell)] = ids[i];
    bool all = true;
    for (int r = 0; r < npe; r++) all = all && sl[r] >= 0;
    if (all) {
      tr.hang_of.push_back(-1);
      return;
    }
    if (pids == nullptr || pn == 0) throw std::runtime_error("hanging leaf without parent nodes");
    const size_t H = tr.pslot.size() / npe;
    tr.hang_of.push_back((int32_t)H);
    tr.pslot.resize((H + 1) * npe, -1);
    int32_t *ps = &tr.pslot[H * npe];
    for (size_t i = 0; i < pn; i++) {
      const Oct &x = nodes[pids[i]];
      if ((int)x.lev != (int)cell.lev - 1) continue;
      ps[lex_rank(x, parent)] = pids[i];
    }
  }

  void rec(const Oct &cell, int rot, const int32_t *ids, size_t n, const int32_t *pids, size_t pn,
           const Oct &parent) {
    if (n == 0) return;
    const int lev = (int)cell.lev;
    bool isLeaf = true;
    for (size_t i = 0; i < n && isLeaf; i++) isLeaf = (int)nodes[ids[i]].lev <= lev;
    if (isLeaf) {
      leaf(cell, ids, n, pids, pn, parent);
      return;
    }
    // Bucket with duplication onto closed children (top_down, matvec.h:74-199).
    const int nch = t.nch;
    const u32 h = 1u << (mp.maxd - lev - 1);
    std::vector<int32_t> &b = buf[lev];
    int cnt[16] = {0}, off[17];
    scratch.resize(n);
    // First pass: counts. Children of node i are encoded as (lo mask, dup mask).
    std::vector<std::pair<unsigned, unsigned>> cc(n);
    for (size_t i = 0; i < n; i++) {
      const Oct &x = nodes[ids[i]];
      unsigned lo = 0, dup = 0;
      for (int d = 0; d < mp.dim; d++) {
        const u32 mid = cell.x[d] + h;
        if (x.x[d] == mid)
          dup |= 1u << d;
        else if (x.x[d] > mid)
          lo |= 1u << d;
      }
      cc[i] = {lo, dup};
      for (unsigned sub = dup;; sub = (sub - 1) & dup) {
        cnt[t.inv[rot * nch + (lo | sub)]]++;
        if (sub == 0) break;
      }
    }
    off[0] = 0;
    for (int c = 0; c < nch; c++) off[c + 1] = off[c] + cnt[c];
    b.resize(off[nch]);
    int fill[16];
    for (int c = 0; c < nch; c++) fill[c] = off[c];
    for (size_t i = 0; i < n; i++) {
      const unsigned lo = cc[i].first, dup = cc[i].second;
      // Ascending child order within the duplicate set, as the reference enumerates it.
      for (unsigned sub = 0;; sub = (sub - dup) & dup) {
        b[fill[t.inv[rot * nch + (lo | sub)]]++] = ids[i];
        if (sub == dup) break;
      }
    }
    // Keep this level's bucket alive while recursing (deeper levels use other buffers).
    std::vector<int32_t> mine;
    mine.swap(b);
    for (int cs = 0; cs < nch; cs++) {
      const unsigned m = t.perm[rot * nch + cs];
      Oct ch = cell;
      ch.lev = (u32)(lev + 1);
      for (int d = 0; d < mp.dim; d++) ch.x[d] = cell.x[d] + (((m >> d) & 1u) ? h : 0u);
      rec(ch, t.child[rot * nch + m], mine.data() + off[cs], (size_t)cnt[cs], ids, n, cell);
    }
    mine.swap(b);
  }
};
}  // namespace

// The reference interpolates missing nodes from its static parent buffer
// (matvec.h:283,411-456): present parent-level nodes overwrite their ranks, the
// buffer is interpolated in place (IKD_Parent2Child with in == out,
// refel.h:214-240), and missing leaf nodes read the result.  A rank whose
// parent-level node is absent keeps the value the previous hanging leaf's
// interpolation left there (zero before the first hanging leaf of a call).
// StaleBuffer evaluates that buffer lazily, as sparse node combinations.
namespace {
void add_scaled(Combo &dst, const Combo &src, double w) {
  for (const auto &e : src) dst.push_back({e.first, e.second * w});
}
void compact(Combo &c) {
  std::sort(c.begin(), c.end(), [](const std::pair<int32_t, double> &a, const std::pair<int32_t, double> &b) {
    return a.first < b.first;
  });
  size_t w = 0;
  for (size_t i = 0; i < c.size(); i++) {
    if (w > 0 && c[w - 1].first == c[i].first)
      c[w - 1].second += c[i].second;
    else
      c[w++] = c[i];
  }
  c.resize(w);
}
}  // namespace

StaleBuffer::StaleBuffer(const MeshParams &mp, const Traversal &tr) : tr_(tr), npe_(tr.nPe) {
  const size_t H = tr.pslot.size() / npe_;
  leaf_of_.resize(H);
  for (size_t e = 0; e < tr.hang_of.size(); e++)
    if (tr.hang_of[e] >= 0) leaf_of_[tr.hang_of[e]] = e;
  double ip[2][9];
  interp_1d(mp.order, 0, ip[0]);
  interp_1d(mp.order, 1, ip[1]);
  const int n1 = mp.order + 1, nch = 1 << mp.dim;
  wnz_.assign(nch, std::vector<std::vector<std::pair<int, double>>>(npe_));
  for (int c = 0; c < nch; c++)
    for (int r = 0; r < npe_; r++)
      for (int s = 0; s < npe_; s++) {
        double w = 1.0;
        int rr = r, ss = s;
        for (int d = 0; d < mp.dim; d++) {
          w *= ip[(c >> d) & 1][(ss % n1) * n1 + (rr % n1)];
          rr /= n1;
          ss /= n1;
        }
        if (w != 0.0) wnz_[c][r].push_back({s, w});
      }
}

const Combo &StaleBuffer::after(size_t k0, int s0, std::unordered_map<uint64_t, Combo> &memo_) const {
  auto key = [this](size_t k, int s) { return (uint64_t)k * (uint64_t)npe_ + (uint64_t)s; };
  auto it = memo_.find(key(k0, s0));
  if (it != memo_.end()) return it->second;
  std::vector<std::pair<size_t, int>> stack{{k0, s0}};
  while (!stack.empty()) {
    const size_t k = stack.back().first;
    const int s = stack.back().second;
    if (memo_.count(key(k, s))) {
      stack.pop_back();
      continue;
    }
    const int c = tr_.child[leaf_of_[k]];
    const int32_t *ps = &tr_.pslot[k * npe_];
    bool ready = true;
    for (const auto &tw : wnz_[c][s])
      if (ps[tw.first] < 0 && k > 0 && !memo_.count(key(k - 1, tw.first))) {
        stack.push_back({k - 1, tw.first});
        ready = false;
      }
    if (!ready) continue;
    Combo v;
    for (const auto &tw : wnz_[c][s]) {
      if (ps[tw.first] >= 0)
        v.push_back({ps[tw.first], tw.second});
      else if (k > 0)
        add_scaled(v, memo_[key(k - 1, tw.first)], tw.second);
    }
    compact(v);
    memo_.emplace(key(k, s), std::move(v));
    stack.pop_back();
  }
  return memo_[key(k0, s0)];
}

void compute_stale_terms(const MeshParams &mp, Traversal &tr) {
  if (tr.hc_ready) return;
  const int npe = tr.nPe;
  const size_t H = tr.pslot.size() / npe;
  StaleBuffer sb(mp, tr);
  tr.hc_off.assign(1, 0);
  tr.hc_rank.clear();
  tr.hc_node.clear();
  tr.hc_w.clear();
  tr.stale_needed = 0;
  for (size_t h = 0; h < H; h++) {
    const size_t e = sb.leaf_of(h);
    const int c = tr.child[e];
    const int32_t *sl = &tr.slot[e * npe];
    const int32_t *ps = &tr.pslot[h * npe];
    bool stale = false;
    for (int r = 0; r < npe; r++) {
      if (sl[r] >= 0) continue;
      Combo corr;
      for (const auto &sw : sb.weights(c, r))
        if (ps[sw.first] < 0 && h > 0) add_scaled(corr, sb.after(h - 1, sw.first), sw.second);
      compact(corr);
      for (const auto &e2 : corr) {
        if (e2.second == 0.0) continue;
        tr.hc_rank.push_back((uint8_t)r);
        tr.hc_node.push_back(e2.first);
        tr.hc_w.push_back(e2.second);
        stale = true;
      }
    }
    if (stale) tr.stale_needed++;
    tr.hc_off.push_back((int64_t)tr.hc_node.size());
  }
  tr.hc_ready = true;
}

Traversal build_conforming(const MeshParams &mp, const std::vector<Oct> &elems, const NodeSet &ns) {
  const int dim = mp.dim, p = mp.order, n1 = p + 1;
  int npe = 1;
  for (int d = 0; d < dim; d++) npe *= n1;
  Traversal tr;
  tr.nPe = npe;
  using Key = std::array<u32, 5>;  // coordinates, level (level 0xFFFFFFFF: node lookup)
  struct KeyHash {
    size_t operator()(const Key &k) const {
      u64 h = 0x9E3779B97F4A7C15ull;
      for (u32 v : k) h = (h ^ v) * 0xBF58476D1CE4E5B9ull;
      return (size_t)(h ^ (h >> 31));
    }
  };
  // Node ids by coordinates: an open-addressing table, filled without a
  // per-node allocation and probed by every element rank.
  struct NodeIndex {
    struct Slot {
      u32 x[4];
      int32_t id;
    };
    std::vector<Slot> slot;
    size_t mask = 0;
    static size_t hash(const u32 *x) {
      u64 h = 0x9E3779B97F4A7C15ull;
      for (int d = 0; d < 4; d++) h = (h ^ x[d]) * 0xBF58476D1CE4E5B9ull;
      return (size_t)(h ^ (h >> 31));
    }
    int32_t find(const u32 *x) const {
      for (size_t i = hash(x) & mask;; i = (i + 1) & mask) {
        const Slot &s = slot[i];
        if (s.id < 0) return -1;
        if (s.x[0] == x[0] && s.x[1] == x[1] && s.x[2] == x[2] && s.x[3] == x[3]) return s.id;
      }
    }
  } at;
  {
    size_t cap = 16;
    while (cap < ns.nodes.size() * 2) cap <<= 1;
    at.slot.assign(cap, NodeIndex::Slot{{0, 0, 0, 0}, -1});
    at.mask = cap - 1;
    // Filled in parallel: a node claims the first free slot of its probe
    // sequence with a compare-and-swap on the id, then writes its key (node
    // positions are distinct; lookups start after the fill).
#pragma omp parallel for schedule(static)
    for (size_t n = 0; n < ns.nodes.size(); n++) {
      u32 x[4] = {0, 0, 0, 0};
      for (int d = 0; d < dim; d++) x[d] = ns.nodes[n].x[d];
      for (size_t i = NodeIndex::hash(x) & at.mask;; i = (i + 1) & at.mask) {
        int32_t expected = -1;
        if (__atomic_compare_exchange_n(&at.slot[i].id, &expected, (int32_t)n, false, __ATOMIC_RELAXED,
                                        __ATOMIC_RELAXED)) {
          for (int d = 0; d < 4; d++) at.slot[i].x[d] = x[d];
          break;
        }
      }
    }
  }
  double ip[2][9];
  interp_1d(p, 0, ip[0]);
  interp_1d(p, 1, ip[1]);
  using Memo = std::unordered_map<Key, Combo, KeyHash>;
  // Value at point x, a node position of the cell (anchor a, level lev).
  std::function<const Combo &(const u32 *, const u32 *, int, Memo &)> resolve = [&](const u32 *x, const u32 *a, int lev,
                                                                                    Memo &memo) -> const Combo & {
    Key key{0, 0, 0, 0, (u32)lev};
    for (int d = 0; d < dim; d++) key[d] = x[d];
    auto m = memo.find(key);
    if (m != memo.end()) return m->second;
    Combo v;
    const u32 xp[4] = {key[0], key[1], key[2], key[3]};
    if (const int32_t f = at.find(xp); f >= 0) {
      v.push_back({f, 1.0});
    } else {
      if (lev == 0) throw std::runtime_error("conforming constraint: a root node is missing");
      const u32 len = 1u << (mp.maxd - lev), step = len / (u32)p;
      const u32 plen = len << 1, pstep = plen / (u32)p;
      u32 pa[4] = {0, 0, 0, 0};
      int c[4] = {0, 0, 0, 0}, t[4] = {0, 0, 0, 0};
      for (int d = 0; d < dim; d++) {
        pa[d] = a[d] & ~(plen - 1);
        c[d] = (int)((a[d] - pa[d]) / len);
        t[d] = (int)((x[d] - a[d]) / step);
      }
      int pn = 1;
      for (int d = 0; d < dim; d++) pn *= n1;
      for (int u = 0; u < pn; u++) {
        double w = 1.0;
        u32 y[4] = {0, 0, 0, 0};
        int uu = u;
        for (int d = 0; d < dim; d++) {
          const int ud = uu % n1;
          uu /= n1;
          w *= ip[c[d]][ud * n1 + t[d]];
          y[d] = pa[d] + (u32)ud * pstep;
        }
        if (w == 0.0) continue;
        add_scaled(v, resolve(y, pa, lev - 1, memo), w);
      }
      compact(v);
    }
    return memo.emplace(key, std::move(v)).first->second;
  };
  // One record per mesh element (SFC order): the node at each rank where the
  // point is a node, else the constraint terms.
  tr.leaves = elems;
  const size_t E = elems.size();
  tr.child.assign(E, 0);
  tr.slot.assign(E * npe, -1);
  tr.hang_of.assign(E, -1);
  tr.cf_off.assign(1, 0);
  // Elements in chunks on worker threads (each with its own memo of resolved
  // points), terms concatenated in element order.
  constexpr size_t kChunk = 2048;
  const size_t nchunks = (E + kChunk - 1) / kChunk;
  struct ChunkTerms {
    std::vector<uint8_t> rank;
    std::vector<int32_t> node;
    std::vector<double> w;
    std::vector<uint32_t> count;  // per element of the chunk: terms (0: no constrained rank)
  };
  std::vector<ChunkTerms> chunks(nchunks);
  std::vector<std::string> errs(nchunks);
#pragma omp parallel
  {
    Memo memo;
#pragma omp for schedule(dynamic, 1)
    for (size_t ci = 0; ci < nchunks; ci++) {
      try {
        if (memo.size() > (1u << 20)) memo.clear();
        ChunkTerms &ct = chunks[ci];
        const size_t e0 = ci * kChunk, e1 = std::min(E, e0 + kChunk);
        ct.count.assign(e1 - e0, 0);
        for (size_t e = e0; e < e1; e++) {
          const Oct &cell = elems[e];
          const u32 len = 1u << (mp.maxd - cell.lev), step = len / (u32)p;
          int c = 0;
          for (int d = 0; d < dim; d++) c |= (int)((cell.x[d] / len) & 1u) << d;
          tr.child[e] = (uint8_t)c;
          const size_t before = ct.node.size();
          for (int r = 0; r < npe; r++) {
            Key k{0, 0, 0, 0, 0xFFFFFFFFu};
            int rr = r;
            for (int d = 0; d < dim; d++) {
              k[d] = cell.x[d] + (u32)(rr % n1) * step;
              rr /= n1;
            }
            if (const int32_t f = at.find(k.data()); f >= 0) {
              tr.slot[e * npe + r] = f;
              continue;
            }
            for (const auto &tw : resolve(k.data(), cell.x, (int)cell.lev, memo)) {
              ct.rank.push_back((uint8_t)r);
              ct.node.push_back(tw.first);
              ct.w.push_back(tw.second);
            }
          }
          ct.count[e - e0] = (uint32_t)(ct.node.size() - before);
        }
      } catch (const std::exception &ex) {
        errs[ci] = ex.what();
      }
    }
  }
  for (const auto &er : errs)
    if (!er.empty()) throw std::runtime_error(er);
  // Concatenation in element order: chunk offsets first, then each chunk
  // copied to its place in parallel.
  std::vector<size_t> rec0(nchunks + 1, 0), term0(nchunks + 1, 0);
  for (size_t ci = 0; ci < nchunks; ci++) {
    size_t nrec = 0;
    for (uint32_t c : chunks[ci].count) nrec += c ? 1 : 0;
    rec0[ci + 1] = rec0[ci] + nrec;
    term0[ci + 1] = term0[ci] + chunks[ci].node.size();
  }
  tr.cf_off.assign(rec0[nchunks] + 1, 0);
  tr.cf_rank.resize(term0[nchunks]);
  tr.cf_node.resize(term0[nchunks]);
  tr.cf_w.resize(term0[nchunks]);
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t ci = 0; ci < nchunks; ci++) {
    ChunkTerms &ct = chunks[ci];
    const size_t e0 = ci * kChunk, t0 = term0[ci];
    std::copy(ct.rank.begin(), ct.rank.end(), tr.cf_rank.begin() + t0);
    std::copy(ct.node.begin(), ct.node.end(), tr.cf_node.begin() + t0);
    std::copy(ct.w.begin(), ct.w.end(), tr.cf_w.begin() + t0);
    size_t h = rec0[ci], k = t0;
    for (size_t i = 0; i < ct.count.size(); i++) {
      if (!ct.count[i]) continue;
      tr.hang_of[e0 + i] = (int32_t)h;
      k += ct.count[i];
      tr.cf_off[++h] = (int64_t)k;
    }
    ct = ChunkTerms();
  }
  // (pslot stays empty: conforming records have no parent-level slots.)
  tr.cf_ready = true;
  return tr;
}

Traversal flatten_traversal(const MeshParams &mp, const std::vector<Oct> &nodes) {
  int npe = 1;
  for (int d = 0; d < mp.dim; d++) npe *= (mp.order + 1);
  const SfcTables &t = sfc_tables(mp.dim, mp.curve);
  // The first kSplit levels of the recursion are expanded here, in visiting
  // order, into independent subtree tasks (a subtree's records depend only on
  // its cell, its node bucket and its parent's bucket); the tasks run
  // concurrently into separate record lists, concatenated in order.
  constexpr int kSplit = 2;
  struct Ref {
    int vec;
    size_t off, n;
  };
  struct Task {
    Oct cell, parent;
    int rot;
    Ref ids, pids;
  };
  std::vector<std::vector<int32_t>> store(1);
  store[0].resize(nodes.size());
  std::iota(store[0].begin(), store[0].end(), 0);
  std::vector<Task> tasks;
  std::function<void(const Oct &, int, Ref, Ref, const Oct &)> expand = [&](const Oct &cell, int rot, Ref ids, Ref pids,
                                                                             const Oct &parent) {
    if (ids.n == 0) return;
    const int lev = (int)cell.lev;
    const int32_t *id = store[ids.vec].data() + ids.off;
    bool isLeaf = true;
    for (size_t i = 0; i < ids.n && isLeaf; i++) isLeaf = (int)nodes[id[i]].lev <= lev;
    if (isLeaf || lev >= kSplit) {
      tasks.push_back({cell, parent, rot, ids, pids});
      return;
    }
    // Same bucketing as Flattener::rec.
    const int nch = t.nch;
    const u32 h = 1u << (mp.maxd - lev - 1);
    size_t cnt[16] = {0}, off[17];
    std::vector<std::pair<unsigned, unsigned>> cc(ids.n);
    for (size_t i = 0; i < ids.n; i++) {
      const Oct &x = nodes[id[i]];
      unsigned lo = 0, dup = 0;
      for (int d = 0; d < mp.dim; d++) {
        const u32 mid = cell.x[d] + h;
        if (x.x[d] == mid)
          dup |= 1u << d;
        else if (x.x[d] > mid)
          lo |= 1u << d;
      }
      cc[i] = {lo, dup};
      for (unsigned sub = dup;; sub = (sub - 1) & dup) {
        cnt[t.inv[rot * nch + (lo | sub)]]++;
        if (sub == 0) break;
      }
    }
    off[0] = 0;
    for (int c = 0; c < nch; c++) off[c + 1] = off[c] + cnt[c];
    const int vec = (int)store.size();
    store.emplace_back(off[nch]);
    id = store[ids.vec].data() + ids.off;  // store may have moved
    std::vector<int32_t> &b = store[vec];
    size_t fill[16];
    for (int c = 0; c < nch; c++) fill[c] = off[c];
    for (size_t i = 0; i < ids.n; i++) {
      const unsigned lo = cc[i].first, dup = cc[i].second;
      for (unsigned sub = 0;; sub = (sub - dup) & dup) {
        b[fill[t.inv[rot * nch + (lo | sub)]]++] = id[i];
        if (sub == dup) break;
      }
    }
    for (int cs = 0; cs < nch; cs++) {
      const unsigned m = t.perm[rot * nch + cs];
      Oct ch = cell;
      ch.lev = (u32)(lev + 1);
      for (int d = 0; d < mp.dim; d++) ch.x[d] = cell.x[d] + (((m >> d) & 1u) ? h : 0u);
      expand(ch, t.child[rot * nch + m], Ref{vec, off[cs], cnt[cs]}, ids, cell);
    }
  };
  Oct root{};
  root.lev = 0;
  expand(root, 0, Ref{0, 0, nodes.size()}, Ref{0, 0, 0}, root);
  std::vector<Traversal> parts(tasks.size());
  std::vector<std::string> errs(tasks.size());
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t k = 0; k < tasks.size(); k++) {
    try {
      Traversal &pt = parts[k];
      pt.nPe = npe;
      Flattener f{mp, t, nodes, pt, npe, {}, {}, {}};
      f.buf.resize(mp.maxd + 2);
      interp_1d(mp.order, 0, f.ip[0]);
      interp_1d(mp.order, 1, f.ip[1]);
      const Task &tk = tasks[k];
      const int32_t *pid = tk.pids.n ? store[tk.pids.vec].data() + tk.pids.off : nullptr;
      f.rec(tk.cell, tk.rot, store[tk.ids.vec].data() + tk.ids.off, tk.ids.n, pid, tk.pids.n, tk.parent);
    } catch (const std::exception &e) {
      errs[k] = e.what();
    }
  }
  for (const auto &e : errs)
    if (!e.empty()) throw std::runtime_error(e);
  Traversal tr;
  tr.nPe = npe;
  size_t L = 0, H = 0;
  for (const auto &pt : parts) {
    L += pt.leaves.size();
    H += pt.pslot.size() / npe;
  }
  tr.leaves.reserve(L);
  tr.child.reserve(L);
  tr.slot.reserve(L * npe);
  tr.hang_of.reserve(L);
  tr.pslot.reserve(H * npe);
  for (auto &pt : parts) {
    const int32_t h0 = (int32_t)(tr.pslot.size() / npe);
    tr.leaves.insert(tr.leaves.end(), pt.leaves.begin(), pt.leaves.end());
    tr.child.insert(tr.child.end(), pt.child.begin(), pt.child.end());
    tr.slot.insert(tr.slot.end(), pt.slot.begin(), pt.slot.end());
    for (int32_t h : pt.hang_of) tr.hang_of.push_back(h < 0 ? h : h + h0);
    tr.pslot.insert(tr.pslot.end(), pt.pslot.begin(), pt.pslot.end());
    pt = Traversal();
  }
  return tr;
}

// ---------------------------------------------------------------------------
// Synthetic inputs (identical to tests/golden/gen_inputs.py).
// ---------------------------------------------------------------------------
static inline u64 mix64(u64 z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static const u64 kGolden = 0x9E3779B97F4A7C15ull;

double keyed_uniform(u64 seed, const u32 *c, int dim) {
  u64 s = mix64(seed + kGolden);
  for (int d = 0; d < dim; d++) s = mix64(s ^ ((u64)c[d] * kGolden + (u64)d));
  return 2.0 * ((double)(s >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

static bool sphere_split(const MeshParams &mp, const Oct &c, int base, int maxl, const double *ctr, double r) {
  if ((int)c.lev < base) return true;
  if ((int)c.lev >= maxl) return false;
  const double scale = 1.0 / (double)((u64)1 << mp.maxd);
  const double h = (double)((u64)1 << (mp.maxd - c.lev)) * scale;
  double mind2 = 0.0, maxd2 = 0.0;
  for (int d = 0; d < mp.dim; d++) {
    const double a = (double)c.x[d] * scale;
    const double lo = a - ctr[d];
    const double hi = (a + h) - ctr[d];
    const double gap = std::max(0.0, std::max(lo, -hi));
    const double g2 = gap * gap;
    mind2 = mind2 + g2;
    const double l2 = lo * lo, h2 = hi * hi;
    maxd2 = maxd2 + std::max(l2, h2);
  }
  const double r2 = r * r;
  return mind2 <= r2 && r2 <= maxd2;
}

std::vector<Oct> sphere_tree(const MeshParams &mp, int base_level, int max_level, const double *center,
                             double radius) {
  std::vector<Oct> out;
  const int nch = 1 << mp.dim;
  std::vector<Oct> stack;
  Oct root{};
  root.lev = 0;
  // Morton-order depth-first emission (the order is irrelevant: the tree is sorted later).
  struct F {
    static void rec(const MeshParams &mp, const Oct &c, int base, int maxl, const double *ctr, double r,
                    std::vector<Oct> &out, int nch) {
      if (c.lev > 0 && !sphere_split(mp, c, base, maxl, ctr, r)) {
        out.push_back(c);
        return;
      }
      const u32 h = 1u << (mp.maxd - c.lev - 1);
      for (int ch = 0; ch < nch; ch++) {
        Oct k = c;
        k.lev = c.lev + 1;
        for (int d = 0; d < mp.dim; d++) k.x[d] = c.x[d] + (((ch >> d) & 1) ? h : 0u);
        rec(mp, k, base, maxl, ctr, r, out, nch);
      }
    }
  };
  F::rec(mp, root, base_level, max_level, center, radius, out, nch);
  return out;
}

std::vector<Oct> gauss_points(const MeshParams &mp, size_t n, int ncomp, double sigma, u64 seed) {
  const int dim = mp.dim;
  const int per = 1 + 12 * dim;
  auto u32at = [seed](u64 i) -> u64 {
    return mix64((seed * kGolden) ^ (i * kGolden + 0x632BE59BD9B4E019ull)) >> 32;
  };
  std::vector<double> ctr((size_t)ncomp * dim);
  for (int i = 0; i < ncomp * dim; i++) ctr[i] = (double)u32at((u64)i) * (1.0 / 4294967296.0);
  std::vector<Oct> out(n);
  const double top = (double)((u64)1 << mp.maxd);
  const u64 base = (u64)ncomp * dim;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) {
    const u64 b = base + (u64)i * per;
    const int comp = (int)(u32at(b) % (u64)ncomp);
    Oct o{};
    o.lev = (u32)mp.maxd;
    for (int d = 0; d < dim; d++) {
      int64_t s = 0;
      for (int k = 0; k < 12; k++) s += (int64_t)u32at(b + 1 + (u64)d * 12 + k);
      const double z = (double)(s - (int64_t)6 * 4294967296LL) * (1.0 / 4294967296.0);
      const double sz = sigma * z;
      double x = ctr[comp * dim + d] + sz;
      x = std::min(std::max(x, 0.0), 1.0);
      double xi = std::floor(x * top);
      xi = std::min(xi, top - 1.0);
      o.x[d] = (u32)xi;
    }
    out[i] = o;
  }
  return out;
}

}  // namespace dkt
