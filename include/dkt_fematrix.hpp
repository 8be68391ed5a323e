/*
 * dkt_fematrix.hpp -- C++ shim that keeps feMatrix::matVec and the user's
 * elementalMatVec as they are and runs the traversal on the device through
 * the C ABI (dkt.h).  Header-only, C++14 (the reference's dialect).
 *
 * Replaces the body of feMatrix<LeafT, dim>::matVec (FEM/include/feMatrix.h:
 * 189-259).  The reference there: copies `in` to a ghosted buffer, lets the
 * leaf class's preMatVec override it (HeatMat: Dirichlet zeros,
 * FEM/examples/src/heatMat.cpp:120-128), runs fem::matvec with
 * elementalMatVec bound as the element operator (std::bind of
 * (in, out, coords, scale), feMatrix.h:225-226, matvec.h:33-38), copies the
 * result to `out` and lets postMatVec override it (heatMat.cpp:130-139).
 * DeviceOperator::matVec does the same with the traversal replaced by
 * dkt_plan_matvec (raw, masked = 0): the hooks still run on the host, on the
 * leaf object, so any user hook keeps its meaning.
 *
 * The element operator is captured, not re-implemented: for every level
 * 0..maxdepth the shim calls the leaf's elementalMatVec on the unit vectors of
 * the element at the origin of that level (coordinates as fem::matvec forms
 * them, matvec.h:459-465, with `scale` passed through untouched as the
 * reference passes it) and hands the per-level matrices to
 * dkt_plan_create(DKT_OP_CAPTURED).  This is exact for operators that depend on
 * the element through its size only (HeatMat, mass and stiffness operators);
 * an operator that depends on the element's position stays on the host
 * traversal.  Plans are captured once per scale value and kept.
 *
 * Integration into the reference (one member and three lines):
 *
 *   // in feMatrix<LeafT, dim>:
 *   std::unique_ptr<dkt_shim::DeviceOperator<dim>> m_device;  // set by the application
 *   void matVec(const VECType *in, VECType *out, double scale) {
 *     if (m_device) { m_device->matVec(asLeaf(), in, out, scale); return; }
 *     ... the reference's host traversal ...
 *   }
 *
 * Vectors are the reference's local nodal vectors (one rank: no ghosts), in
 * the DA's node order, which dkt_da_create reproduces.
 */
#ifndef DKT_FEMATRIX_HPP
#define DKT_FEMATRIX_HPP

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "dkt.h"

namespace dkt_shim {

inline void check(int rc) {
  if (rc != 0) throw std::runtime_error(std::string("dkt: ") + dkt_last_error());
}
inline void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Node coordinates of the element at the origin of level `lev`, as fem::matvec
// hands them to the element operator: dim * nPe doubles in [0,1], node n in
// lexicographic order (axis 0 fastest) at 2^-lev * i_d / p.
inline std::vector<double> element_coords(int dim, int p, int lev) {
  int npe = 1;
  for (int d = 0; d < dim; d++) npe *= p + 1;
  const double h = 1.0 / (double)(1ull << lev);
  std::vector<double> c((size_t)dim * npe);
  for (int n = 0; n < npe; n++) {
    int r = n;
    for (int d = 0; d < dim; d++, r /= p + 1) c[(size_t)n * dim + d] = h * (double)(r % (p + 1)) / (double)p;
  }
  return c;
}

// Per-level element matrices of an element operator `op(in, out, coords,
// scale)` (the signature of feMatrix::elementalMatVec): K[lev][r][c] = out[r]
// for in = e_c on the element at the origin of level lev, lev = 0..maxdepth;
// row-major, out = K in, the layout DKT_OP_CAPTURED takes.
template <typename EleOp>
std::vector<double> capture_level_matrices(EleOp &&op, int dim, int p, int maxdepth, double scale) {
  int npe = 1;
  for (int d = 0; d < dim; d++) npe *= p + 1;
  std::vector<double> K((size_t)(maxdepth + 1) * npe * npe);
  std::vector<double> in(npe), out(npe);
  for (int lev = 0; lev <= maxdepth; lev++) {
    std::vector<double> coords = element_coords(dim, p, lev);
    for (int c = 0; c < npe; c++) {
      std::fill(in.begin(), in.end(), 0.0);
      std::fill(out.begin(), out.end(), 0.0);
      in[c] = 1.0;
      op(in.data(), out.data(), coords.data(), scale);
      for (int r = 0; r < npe; r++) K[((size_t)lev * npe + r) * npe + c] = out[r];
    }
  }
  return K;
}

// A DA's mesh on the device, and the captured operators of one leaf class.
template <unsigned int dim>
class DeviceOperator {
 public:
  // cells: the DA's elements, uint32 [n][dim + 1] (TreeNode anchor
  // coordinates, level); order: the DA's element order; maxdepth: the
  // reference's m_uiMaxDepth; curve: DKT_CURVE_HILBERT when the reference is
  // built with HILBERT_ORDERING, else DKT_CURVE_MORTON.
  DeviceOperator(const uint32_t *cells, size_t n, int order, int maxdepth,
#ifdef HILBERT_ORDERING
                 int curve = DKT_CURVE_HILBERT
#else
                 int curve = DKT_CURVE_MORTON
#endif
                 )
      : order_(order), maxdepth_(maxdepth) {
    dkt_octs *o = nullptr;
    check(dkt_octs_create((int)dim, maxdepth, curve, cells, n, &o));
    const int rc_sort = dkt_tree_sort(o);
    const int rc_da = rc_sort == 0 ? dkt_da_create(o, order, &da_) : rc_sort;
    dkt_octs_destroy(o);
    check(rc_da);
    n_ = dkt_da_count(da_, 1);
    check_hip(hipMalloc((void **)&x_, n_ * sizeof(double)), "hipMalloc");
    check_hip(hipMalloc((void **)&y_, n_ * sizeof(double)), "hipMalloc");
  }
  ~DeviceOperator() {
    for (auto &kv : plans_) dkt_plan_destroy(kv.second);
    (void)hipFree(x_);
    (void)hipFree(y_);
    if (da_) dkt_da_destroy(da_);
  }
  DeviceOperator(const DeviceOperator &) = delete;
  DeviceOperator &operator=(const DeviceOperator &) = delete;

  size_t nodes() const { return n_; }
  dkt_da *da() const { return da_; }

  // The device plan of leaf's element operator at this scale (captured on
  // first use).
  template <typename LeafT>
  dkt_plan *plan(LeafT &leaf, double scale) {
    auto it = plans_.find(scale);
    if (it != plans_.end()) return it->second;
    const std::vector<double> K = capture_level_matrices(
        [&leaf](const double *in, double *out, double *coords, double s) { leaf.elementalMatVec(in, out, coords, s); },
        (int)dim, order_, maxdepth_, scale);
    dkt_plan *pl = nullptr;
    check(dkt_plan_create(da_, DKT_OP_CAPTURED, K.data(), 0, &pl));
    plans_[scale] = pl;
    return pl;
  }

  // feMatrix<LeafT, dim>::matVec(in, out, scale) with the traversal on the
  // device (host vectors of nodes() doubles; see the file comment).
  template <typename LeafT>
  void matVec(LeafT &leaf, const double *in, double *out, double scale = 1.0) {
    dkt_plan *pl = plan(leaf, scale);
    hin_.assign(in, in + n_);
    leaf.preMatVec(in, hin_.data(), scale);
    check_hip(hipMemcpy(x_, hin_.data(), n_ * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    check(dkt_plan_matvec(pl, 0, x_, y_, 1, 0, nullptr));
    hout_.resize(n_);
    check_hip(hipMemcpy(hout_.data(), y_, n_ * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    std::copy(hout_.begin(), hout_.end(), out);
    leaf.postMatVec(hout_.data(), out, scale);
  }

 private:
  int order_, maxdepth_;
  dkt_da *da_ = nullptr;
  size_t n_ = 0;
  double *x_ = nullptr, *y_ = nullptr;
  std::map<double, dkt_plan *> plans_;
  std::vector<double> hin_, hout_;
};

}  // namespace dkt_shim

#endif /* DKT_FEMATRIX_HPP */
