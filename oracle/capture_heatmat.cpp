// Captures the reference's HeatEq::HeatMat<3> element operator
// (FEM/examples/src/heatMat.cpp:46-117) through the feMatrix shim
// (include/dkt_fematrix.hpp: capture_level_matrices, the same code the shim
// runs before it builds a DKT_OP_CAPTURED plan), writes the per-level
// matrices as .npy, and hands them to dkt_plan_create(DKT_OP_CAPTURED) as a
// dry run on the same tree.  Test infrastructure only (links the reference
// build of oracle/Makefile and libdkt.so).
//
// Usage: capture_heatmat OUT.npy [maxdepth=30]
#include <mpi.h>

#include <cstdio>
#include <string>
#include <vector>

#include "hcurvedata.h"
#include "heatMat.h"
#include "oda.h"
#include "point.h"
#include "treeNode.h"
#include "tsort.h"

#include "dkt_fematrix.hpp"
#include "npy_io.h"

int main(int argc, char **argv) {
  MPI_Init(&argc, &argv);
  if (argc < 2) {
    fprintf(stderr, "usage: %s OUT.npy [maxdepth]\n", argv[0]);
    return 2;
  }
  const int maxdepth = argc > 2 ? std::stoi(argv[2]) : 30;
  m_uiMaxDepth = maxdepth;
  _InitializeHcurve(3);
  MPI_Comm comm = MPI_COMM_WORLD;
  // A uniform level-2 tree: 64 cells.
  std::vector<ot::TreeNode<unsigned int, 3>> tree;
  for (unsigned k = 0; k < 4; k++)
    for (unsigned j = 0; j < 4; j++)
      for (unsigned i = 0; i < 4; i++) {
        const unsigned s = maxdepth - 2;
        std::array<unsigned int, 3> c{i << s, j << s, k << s};
        tree.emplace_back(c, 2);
      }
  ot::SFC_Tree<unsigned int, 3>::distTreeSort(tree, 0.3, comm);
  ot::DA<3> da(tree.data(), (unsigned)tree.size(), comm, 1);
  HeatEq::HeatMat<3> heat(&da, 1);
  heat.setProblemDimensions(Point<3>(0.0, 0.0, 0.0), Point<3>(1.0, 1.0, 1.0));
  const std::vector<double> K = dkt_shim::capture_level_matrices(
      [&heat](const double *in, double *out, double *coords, double s) { heat.elementalMatVec(in, out, coords, s); },
      3, 1, maxdepth, 1.0);
  npyio::save(argv[1], K.data(), {(size_t)maxdepth + 1, 8, 8});

  // The captured matrices through the ABI (host-side plan, no device).
  std::vector<uint32_t> cells(tree.size() * 4);
  for (size_t e = 0; e < tree.size(); e++) {
    for (int d = 0; d < 3; d++) cells[e * 4 + d] = tree[e].getX(d);
    cells[e * 4 + 3] = tree[e].getLevel();
  }
#ifdef HILBERT_ORDERING
  const int curve = DKT_CURVE_HILBERT;
#else
  const int curve = DKT_CURVE_MORTON;
#endif
  dkt_octs *o = nullptr;
  dkt_da *dda = nullptr;
  dkt_plan *pl = nullptr;
  dkt_shim::check(dkt_octs_create(3, maxdepth, curve, cells.data(), tree.size(), &o));
  dkt_shim::check(dkt_tree_sort(o));
  dkt_shim::check(dkt_da_create(o, 1, &dda));
  dkt_shim::check(dkt_plan_create(dda, DKT_OP_CAPTURED, K.data(), DKT_PLAN_DRYRUN, &pl));
  const size_t n_dkt = dkt_da_count(dda, 1);
  printf("captured levels 0..%d; DA nodes reference %zu, dkt %zu\n", maxdepth, (size_t)da.getLocalNodalSz(), n_dkt);
  dkt_plan_destroy(pl);
  dkt_da_destroy(dda);
  dkt_octs_destroy(o);
  MPI_Finalize();
  return n_dkt == (size_t)da.getLocalNodalSz() ? 0 : 1;
}
