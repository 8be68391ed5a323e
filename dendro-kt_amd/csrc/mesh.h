// Host mesh layer: SFC tables, tree construction and 2:1 balancing, CG node
// enumeration in Dendro-KT's node order, and the flattening of the reference's
// recursive matvec traversal into per-leaf gather maps.
//
// Everything here runs once per mesh on the host.  The results are plain
// arrays handed to the device (dkt_plan in plan.h).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace dkt {

using u32 = uint32_t;
using u64 = uint64_t;
using u128 = unsigned __int128;

enum Curve { CURVE_MORTON = 0, CURVE_HILBERT = 1 };

// Child-order tables of a space-filling curve (the reference's rotations[] and
// HILBERT_TABLE[], src/KDhcurvedata.cpp:25-60).  Row r is an orientation.
struct SfcTables {
  int dim = 0, nch = 0, nrot = 0;
  std::vector<uint8_t> perm;   // [nrot][nch]: child_sfc -> Morton child
  std::vector<uint8_t> inv;    // [nrot][nch]: Morton child -> child_sfc
  std::vector<int32_t> child;  // [nrot][nch]: Morton child -> child orientation
};
const SfcTables &sfc_tables(int dim, int curve);

// A cell or point: anchor coordinates (integer grid 2^maxd) and level.
struct Oct {
  u32 x[4];
  u32 lev;
};

struct MeshParams {
  int dim = 3;
  int maxd = 30;   // the reference's m_uiMaxDepth
  int order = 1;
  int curve = CURVE_HILBERT;
};

// SFC key of an octant: child_sfc digits for levels 1..lev (dim bits each, level
// 1 most significant), zero below the octant's level.
u128 sfc_key(const MeshParams &mp, const Oct &o);

// Sorts octants in SFC order with ancestors before descendants (the order of
// SFC_Tree::locTreeSort, include/tsort.tcc:14-80) and removes duplicates and
// ancestors of later octants (SFC_Tree::locRemoveDuplicates, src/tsort.cpp:718-741).
void sort_octants(const MeshParams &mp, std::vector<Oct> &octs, bool remove_ancestors);

// Complete tree whose leaves hold at most max_pts of the given octants/points
// (SFC_Tree::locTreeConstruction semantics, src/tsort.cpp:569-645; leaf level
// capped at maxd).  `pts` is sorted in place.
std::vector<Oct> construct_tree(const MeshParams &mp, std::vector<Oct> &pts, size_t max_pts);

// 2:1 balanced complete tree refining the tree built from `pts`
// (SFC_Tree::distTreeBalancing on one rank, src/tsort.cpp:860-878).
std::vector<Oct> balance_tree(const MeshParams &mp, std::vector<Oct> &pts, size_t max_pts);

// CG nodes of a balanced tree, in the reference's single-rank DA order
// (ot::DA::construct, src/oda.cpp:53-151).
struct NodeSet {
  std::vector<Oct> nodes;   // exterior nodes in countCGNodes order, then element-interior nodes
  size_t n_exterior = 0;
  std::vector<u32> bdy;     // indices of domain-boundary nodes (DA::m_uiBdyNodeIds)
};
NodeSet enumerate_nodes(const MeshParams &mp, const std::vector<Oct> &elems);

// Flattened traversal (fem::matvec_rec, FEM/include/matvec.h:243-548): one
// record per traversal leaf, in traversal (SFC) order.
struct Traversal {
  int nPe = 0;
  std::vector<Oct> leaves;        // leaf cells
  std::vector<uint8_t> child;     // Morton index of the leaf within its parent
  std::vector<int32_t> slot;      // [L][nPe]: node index, or -1 if the leaf lacks that node
  std::vector<int32_t> hang_of;   // [L]: index into pslot rows, -1 if the leaf has all nodes
  std::vector<int32_t> pslot;     // [H][nPe]: parent-level node at parent lexicographic rank, or -1
  // Stale parent-buffer terms of hanging leaf h: for k in [hc_off[h], hc_off[h+1]),
  // missing rank hc_rank[k] adds hc_w[k] * x[hc_node[k]] to its interpolated value.
  std::vector<int64_t> hc_off;
  std::vector<uint8_t> hc_rank;
  std::vector<int32_t> hc_node;
  std::vector<double> hc_w;
  size_t stale_needed = 0;        // hanging leaves that read a stale parent-buffer entry
  bool hc_ready = false;          // hc_* computed (compute_stale_terms)
  // Conforming constraints (build_conforming): for record h and k in
  // [cf_off[h], cf_off[h+1]), missing rank cf_rank[k] takes cf_w[k] * x[cf_node[k]].
  std::vector<int64_t> cf_off;
  std::vector<uint8_t> cf_rank;
  std::vector<int32_t> cf_node;
  std::vector<double> cf_w;
  bool cf_ready = false;
};
Traversal flatten_traversal(const MeshParams &mp, const std::vector<Oct> &nodes);

using Combo = std::vector<std::pair<int32_t, double>>;

// Lazy evaluation of the reference's static parent buffer: after(k, s) is the
// value of buffer rank s after the k-th hanging leaf (traversal order) has
// interpolated in place, as a sparse combination of x entries.
class StaleBuffer {
 public:
  StaleBuffer(const MeshParams &mp, const Traversal &tr);
  const Combo &after(size_t k, int s) { return after(k, s, memo_); }
  // Same, with a caller-owned memo (concurrent callers each pass their own).
  const Combo &after(size_t k, int s, std::unordered_map<uint64_t, Combo> &memo) const;
  size_t leaf_of(size_t h) const { return leaf_of_[h]; }
  // Nonzero interpolation weights (parent rank s, w) of child rank r, child c.
  const std::vector<std::pair<int, double>> &weights(int c, int r) const { return wnz_[c][r]; }

 private:
  const Traversal &tr_;
  int npe_;
  std::vector<size_t> leaf_of_;
  std::vector<std::vector<std::vector<std::pair<int, double>>>> wnz_;
  std::unordered_map<uint64_t, Combo> memo_;
};

// Per-leaf stale terms hc_* (all hanging leaves); computed on demand.
void compute_stale_terms(const MeshParams &mp, Traversal &tr);

// The conforming element list: one record per mesh element (not per reference
// traversal leaf: the reference traversal leaves part of the domain
// unvisited), with the node at each rank where the point is a node and, for
// the other ranks, constraint terms cf_*: the value at a point that is not a
// node is the parent cell's interpolant there, taken recursively through
// coarser cells until every point is a node (the continuous piecewise-
// polynomial space on the 2:1 balanced tree).  The operator
// sum_e C_e^T K_e C_e is symmetric.  Not the reference's semantics (its
// static parent buffer, FEM/include/matvec.h:403-457); used for solves at scale.
Traversal build_conforming(const MeshParams &mp, const std::vector<Oct> &elems, const NodeSet &ns);

// Integer-exact synthetic inputs (also restated in numpy by the tests).
std::vector<Oct> sphere_tree(const MeshParams &mp, int base_level, int max_level, const double *center,
                             double radius);
std::vector<Oct> gauss_points(const MeshParams &mp, size_t n, int ncomp, double sigma, u64 seed);
double keyed_uniform(u64 seed, const u32 *coords, int dim);

// 1D Lagrange interpolation from a parent to child `c` (0/1) at equispaced
// nodes of order p: ip[k*(p+1)+i] = l_k(child point i)  (RefElement ip_1D_c,
// FEM/src/refel.cpp:158-165).
void interp_1d(int p, int c, double *ip);

}  // namespace dkt
