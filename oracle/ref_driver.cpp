// Driver that runs the reference Dendro-KT implementation (compiled from its own
// sources where they lie, see oracle/Makefile) to produce golden vectors.
//
// Test infrastructure only.  It builds the reference's own tree (distTreeSort /
// distTreeBalancing, src/tsort.cpp:201-878), DA (src/oda.cpp:53-151) and runs
// the reference traversal matvec fem::matvec (FEM/include/matvec.h:231-548)
// with a choice of elemental operators:
//   ident   identity elemental op (as test/testMatvec.cpp:187-198)
//   poisson Poisson stiffness K_e = h^(dim-2) K_ref, K_ref the unit-cube Q_p
//           stiffness built from closed-form 1D matrices (a true K-dim
//           operator; the reference's HeatMat<4> is not one)
//   mass    Q_p mass matrix h^dim M_ref
//   heat    the reference's own HeatEq::HeatMat<3> (FEM/examples/src/heatMat.cpp:46-117)
//
// Usage:
//   ref_driver --dim D --maxdepth M --order P --in octs.npy --mode {tree|balance}
//              --out DIR [--matvec OP:X:MASK]... [--leaves] [--matrix OP:MASK]
//              [--time OP:REPS:SEED]
//   octs.npy : uint32 [n, D+1] (anchor coords..., level).  mode=tree sorts the
//              given complete tree; mode=balance runs distTreeBalancing with
//              maxPtsPerRegion=1 (octants or points at level M).
//   X        : ones | index | hash<SEED>
//   MASK     : raw (fem::matvec) | masked (feMatrix::matVec with Dirichlet pre/post)
#include <mpi.h>
#include <chrono>
#include <cmath>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "treeNode.h"
#include "tsort.h"
#include "nsort.h"
#include "oda.h"
#include "hcurvedata.h"
#include "feMatrix.h"
#include "matvec.h"
#include "heatMat.h"
#include "point.h"

#include "npy_io.h"

using C = unsigned int;

struct Args {
  int dim = 3, maxdepth = 30, order = 1;
  std::string in, mode = "tree", out = ".";
  std::vector<std::string> matvecs, matrices, times;
  bool leaves = false;
};

// 1D closed-form stiffness/mass on [0,1] with equispaced Lagrange nodes.
static void oneD(int p, std::vector<double> &K1, std::vector<double> &M1) {
  if (p == 1) {
    K1 = {1, -1, -1, 1};
    M1 = {1.0 / 3, 1.0 / 6, 1.0 / 6, 1.0 / 3};
  } else if (p == 2) {
    K1 = {7.0 / 3, -8.0 / 3, 1.0 / 3, -8.0 / 3, 16.0 / 3, -8.0 / 3, 1.0 / 3, -8.0 / 3, 7.0 / 3};
    M1 = {4.0 / 30, 2.0 / 30, -1.0 / 30, 2.0 / 30, 16.0 / 30, 2.0 / 30, -1.0 / 30, 2.0 / 30, 4.0 / 30};
  } else {
    fprintf(stderr, "order %d not supported by driver operators\n", p);
    exit(2);
  }
}

// Dense tensor-product reference matrix: sum over terms of Kronecker products,
// lexicographic node index with axis 0 fastest.
static std::vector<double> kronSum(int dim, int p, bool stiffness) {
  std::vector<double> K1, M1;
  oneD(p, K1, M1);
  const int n = p + 1;
  int nPe = 1;
  for (int d = 0; d < dim; d++) nPe *= n;
  std::vector<double> A(nPe * nPe, 0.0);
  for (int i = 0; i < nPe; i++)
    for (int j = 0; j < nPe; j++) {
      double s = 0.0;
      const int nterms = stiffness ? dim : 1;
      for (int t = 0; t < nterms; t++) {
        double prod = 1.0;
        int ii = i, jj = j;
        for (int d = 0; d < dim; d++) {
          const int a = ii % n, b = jj % n;
          ii /= n;
          jj /= n;
          const double *m = (stiffness && d == t) ? K1.data() : M1.data();
          prod *= m[a * n + b];
        }
        s += prod;
      }
      A[i * nPe + j] = s;
    }
  return A;
}

template <unsigned dim>
class DrvMat : public feMatrix<DrvMat<dim>, dim> {
 public:
  std::function<void(const double *, double *, double *, double)> op;
  DrvMat(ot::DA<dim> *da) : feMatrix<DrvMat<dim>, dim>(da, 1) {}
  void elementalMatVec(const VECType *in, VECType *out, double *coords, double scale) override {
    op(in, out, coords, scale);
  }
  bool preMatVec(const VECType *in, VECType *out, double scale) {
    std::vector<unsigned int> b;
    feMat<dim>::m_uiOctDA->getBoundaryNodeIndices(b);
    for (unsigned int i : b) out[i] = 0.0;
    return true;
  }
  bool postMatVec(const VECType *in, VECType *out, double scale) {
    std::vector<unsigned int> b;
    feMat<dim>::m_uiOctDA->getBoundaryNodeIndices(b);
    for (unsigned int i : b) out[i] = 0.0;
    return true;
  }
};

static std::vector<std::string> split(const std::string &s, char c) {
  std::vector<std::string> r;
  size_t a = 0;
  while (true) {
    size_t b = s.find(c, a);
    r.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
    if (b == std::string::npos) break;
    a = b + 1;
  }
  return r;
}

template <unsigned dim>
struct Runner {
  Args a;
  MPI_Comm comm = MPI_COMM_WORLD;
  std::vector<ot::TreeNode<C, dim>> tree;
  ot::DA<dim> *da = nullptr;
  size_t N = 0;
  int nPe = 0;
  std::vector<uint32_t> nodeInts;  // [N, dim+1]

  std::function<void(const double *, double *, double *, double)> makeOp(const std::string &name,
                                                                         HeatEq::HeatMat<3> **heatOut) {
    const int p = a.order;
    if (name == "ident") {
      const int n = nPe;
      return [n](const double *in, double *out, double *, double) {
        for (int i = 0; i < n; i++) out[i] = in[i];
      };
    }
    if (name == "poisson" || name == "mass") {
      const bool stiff = (name == "poisson");
      auto A = std::make_shared<std::vector<double>>(kronSum(dim, p, stiff));
      const int n = nPe;
      const int hp = stiff ? (int)dim - 2 : (int)dim;
      return [A, n, hp](const double *in, double *out, double *coords, double) {
        const double h = coords[(n - 1) * dim + 0] - coords[0];
        const double s = std::pow(h, hp);
        for (int i = 0; i < n; i++) {
          double acc = 0.0;
          for (int j = 0; j < n; j++) acc += (*A)[i * n + j] * in[j];
          out[i] = s * acc;
        }
      };
    }
    if (name == "heat") {
      if (dim != 3) {
        fprintf(stderr, "heat op is the reference's 3D HeatMat only\n");
        exit(2);
      }
      auto *hm = new HeatEq::HeatMat<3>((ot::DA<3> *)da, 1);
      hm->setProblemDimensions(Point<3>(0.0, 0.0, 0.0), Point<3>(1.0, 1.0, 1.0));
      *heatOut = hm;
      return [hm](const double *in, double *out, double *coords, double scale) {
        hm->elementalMatVec(in, out, coords, scale);
      };
    }
    fprintf(stderr, "unknown op %s\n", name.c_str());
    exit(2);
  }

  void makeX(const std::string &spec, std::vector<double> &x) {
    x.assign(N, 0.0);
    if (spec == "ones") {
      for (size_t i = 0; i < N; i++) x[i] = 1.0;
    } else if (spec == "index") {
      for (size_t i = 0; i < N; i++) x[i] = (double)i;
    } else if (spec.rfind("hash", 0) == 0) {
      uint64_t seed = std::stoull(spec.substr(4));
      for (size_t i = 0; i < N; i++) x[i] = npyio::keyed_uniform(seed, &nodeInts[i * (dim + 1)], dim);
    } else {
      fprintf(stderr, "bad x spec %s\n", spec.c_str());
      exit(2);
    }
  }

  void applyOnce(const std::string &opname, bool masked, const std::vector<double> &x, std::vector<double> &y) {
    HeatEq::HeatMat<3> *hm = nullptr;
    auto op = makeOp(opname, &hm);
    y.assign(N, 0.0);
    if (masked) {
      // feMatrix::matVec with HeatMat-style Dirichlet hooks (heatMat.cpp:120-139):
      // zero the ghosted input and the output at domain-boundary nodes.
      DrvMat<dim> m(da);
      m.op = op;
      m.matVec(x.data(), y.data(), 1.0);
    } else {
      fem::matvec<double, ot::TreeNode<C, dim>, RefElement>(
          x.data(), y.data(), da->getTNCoords(), da->getTotalNodalSz(), *da->getTreePartFront(),
          *da->getTreePartBack(), op, 1.0, da->getReferenceElement());
    }
    delete hm;
  }

  int run() {
    std::vector<size_t> shp;
    auto raw = npyio::load<uint32_t>(a.in, shp);
    if (shp.size() != 2 || shp[1] != dim + 1) {
      fprintf(stderr, "input must be [n, dim+1]\n");
      return 2;
    }
    std::vector<ot::TreeNode<C, dim>> pts;
    pts.reserve(shp[0]);
    for (size_t i = 0; i < shp[0]; i++) {
      std::array<C, dim> c;
      for (unsigned d = 0; d < dim; d++) c[d] = raw[i * (dim + 1) + d];
      pts.emplace_back(c, raw[i * (dim + 1) + dim]);
    }
    auto t0 = std::chrono::steady_clock::now();
    if (a.mode == "balance") {
      ot::SFC_Tree<C, dim>::distTreeBalancing(pts, tree, 1, 0.3, comm);
    } else if (a.mode == "tree") {
      ot::SFC_Tree<C, dim>::distTreeSort(pts, 0.3, comm);
      tree = pts;
    } else {
      fprintf(stderr, "bad mode\n");
      return 2;
    }
    auto t1 = std::chrono::steady_clock::now();
    std::vector<uint32_t> el(tree.size() * (dim + 1));
    for (size_t i = 0; i < tree.size(); i++) {
      for (unsigned d = 0; d < dim; d++) el[i * (dim + 1) + d] = tree[i].getX(d);
      el[i * (dim + 1) + dim] = tree[i].getLevel();
    }
    npyio::save(a.out + "/elems.npy", el.data(), {tree.size(), (size_t)dim + 1});

    da = new ot::DA<dim>(tree.data(), (unsigned)tree.size(), comm, a.order);
    auto t2 = std::chrono::steady_clock::now();
    N = da->getLocalNodalSz();
    nPe = da->getNumNodesPerElement();
    const auto *tn = da->getTNCoords() + da->getLocalNodeBegin();
    nodeInts.resize(N * (dim + 1));
    for (size_t i = 0; i < N; i++) {
      for (unsigned d = 0; d < dim; d++) nodeInts[i * (dim + 1) + d] = tn[i].getX(d);
      nodeInts[i * (dim + 1) + dim] = tn[i].getLevel();
    }
    npyio::save(a.out + "/nodes.npy", nodeInts.data(), {N, (size_t)dim + 1});
    std::vector<unsigned int> b;
    da->getBoundaryNodeIndices(b);
    std::vector<uint32_t> bb(b.begin(), b.end());
    npyio::save(a.out + "/bdy.npy", bb.data(), {bb.size()});
    printf("elements %zu nodes %zu global %lld bdy %zu tree_s %.3f da_s %.3f\n", tree.size(), N,
           (long long)da->getGlobalNodeSz(), bb.size(),
           std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count());

    for (const auto &spec : a.matvecs) {
      auto f = split(spec, ':');
      std::vector<double> x, y;
      makeX(f[1], x);
      applyOnce(f[0], f[2] == "masked", x, y);
      npyio::save(a.out + "/y_" + f[0] + "_" + f[1] + "_" + f[2] + ".npy", y.data(), {N});
    }
    if (a.leaves) {
      // Record each traversal leaf: its node coordinates (scaled ints) and the
      // gathered input with vecIn[i] = i.
      std::vector<double> recIn, recCoords;
      const int n = nPe;
      auto op = [&](const double *in, double *out, double *coords, double) {
        for (int i = 0; i < n; i++) recIn.push_back(in[i]), out[i] = 0.0;
        for (int i = 0; i < n * (int)dim; i++) recCoords.push_back(coords[i]);
      };
      std::vector<double> x(N), y(N);
      for (size_t i = 0; i < N; i++) x[i] = (double)i;
      fem::matvec<double, ot::TreeNode<C, dim>, RefElement>(
          x.data(), y.data(), da->getTNCoords(), da->getTotalNodalSz(), *da->getTreePartFront(),
          *da->getTreePartBack(), op, 1.0, da->getReferenceElement());
      const size_t L = recIn.size() / n;
      const double scale = (double)(1u << m_uiMaxDepth);
      std::vector<uint32_t> anchors(L * (dim + 1));
      for (size_t l = 0; l < L; l++) {
        const double *c = &recCoords[l * n * dim];
        const double h = c[(n - 1) * dim] - c[0];
        for (unsigned d = 0; d < dim; d++) anchors[l * (dim + 1) + d] = (uint32_t)std::llround(c[d] * scale);
        anchors[l * (dim + 1) + dim] = (uint32_t)std::llround(m_uiMaxDepth - std::log2(h * scale));
      }
      npyio::save(a.out + "/leaf_cells.npy", anchors.data(), {L, (size_t)dim + 1});
      npyio::save(a.out + "/leaf_in.npy", recIn.data(), {L, (size_t)n});
    }
    for (const auto &spec : a.matrices) {
      auto f = split(spec, ':');
      std::vector<double> A(N * N), x(N), y;
      for (size_t j = 0; j < N; j++) {
        std::fill(x.begin(), x.end(), 0.0);
        x[j] = 1.0;
        applyOnce(f[0], f[1] == "masked", x, y);
        for (size_t i = 0; i < N; i++) A[i * N + j] = y[i];
      }
      npyio::save(a.out + "/A_" + f[0] + "_" + f[1] + ".npy", A.data(), {N, N});
    }
    for (const auto &spec : a.times) {
      auto f = split(spec, ':');
      const int reps = std::stoi(f[1]);
      std::vector<double> x, y;
      makeX("hash" + f[2], x);
      applyOnce(f[0], false, x, y);  // warm-up
      auto s0 = std::chrono::steady_clock::now();
      for (int r = 0; r < reps; r++) applyOnce(f[0], false, x, y);
      auto s1 = std::chrono::steady_clock::now();
      const double per = std::chrono::duration<double>(s1 - s0).count() / reps;
      printf("time op %s reps %d ms_per_matvec %.6f mdof_per_s %.6f\n", f[0].c_str(), reps, per * 1e3,
             N / per / 1e6);
    }
    delete da;
    return 0;
  }
};

int main(int argc, char **argv) {
  MPI_Init(&argc, &argv);
  Args a;
  for (int i = 1; i < argc; i++) {
    std::string k = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) {
        fprintf(stderr, "missing value for %s\n", k.c_str());
        exit(2);
      }
      return argv[++i];
    };
    if (k == "--dim") a.dim = std::stoi(next());
    else if (k == "--maxdepth") a.maxdepth = std::stoi(next());
    else if (k == "--order") a.order = std::stoi(next());
    else if (k == "--in") a.in = next();
    else if (k == "--mode") a.mode = next();
    else if (k == "--out") a.out = next();
    else if (k == "--matvec") a.matvecs.push_back(next());
    else if (k == "--matrix") a.matrices.push_back(next());
    else if (k == "--time") a.times.push_back(next());
    else if (k == "--leaves") a.leaves = true;
    else {
      fprintf(stderr, "unknown arg %s\n", k.c_str());
      return 2;
    }
  }
  m_uiMaxDepth = a.maxdepth;
  _InitializeHcurve(a.dim);
  int rc = 2;
  if (a.dim == 2) { Runner<2> r; r.a = a; rc = r.run(); }
  else if (a.dim == 3) { Runner<3> r; r.a = a; rc = r.run(); }
  else if (a.dim == 4) { Runner<4> r; r.a = a; rc = r.run(); }
  MPI_Finalize();
  return rc;
}
