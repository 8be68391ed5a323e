// Compiles three of the reference's translation units in place and adds
// explicit instantiations of the entry points the driver calls.  The reference
// declares `template struct SFC_Tree<...>`, `template class DA<...>` and
// `template class HeatMat<...>` in its headers, before the member definitions in
// these .cpp files; clang (unlike gcc) then emits only the members defined at
// that point.  Test infrastructure only.
#include "src/tsort.cpp"
#include "src/oda.cpp"
#include "FEM/examples/src/heatMat.cpp"

#define DKT_INST_DIM(D)                                                                              \
  template void ot::SFC_Tree<unsigned int, D>::distTreeSort(std::vector<ot::TreeNode<unsigned int, D>> &, \
                                                          double, MPI_Comm);                         \
  template void ot::SFC_Tree<unsigned int, D>::distTreeBalancing(                                    \
      std::vector<ot::TreeNode<unsigned int, D>> &, std::vector<ot::TreeNode<unsigned int, D>> &, RankI, \
      double, MPI_Comm);                                                                             \
  template void ot::SFC_Tree<unsigned int, D>::distTreeConstruction(                                 \
      std::vector<ot::TreeNode<unsigned int, D>> &, std::vector<ot::TreeNode<unsigned int, D>> &, RankI, \
      double, MPI_Comm);                                                                             \
  template ot::DA<D>::DA(const ot::TreeNode<unsigned int, D> *, unsigned int, MPI_Comm, unsigned int,   \
                         unsigned int, double);                                                      \
  template ot::DA<D>::~DA();                                                                         \
  template HeatEq::HeatMat<D>::HeatMat(ot::DA<D> *, unsigned int);                                   \
  template HeatEq::HeatMat<D>::~HeatMat();                                                           \
  template bool HeatEq::HeatMat<D>::preMatVec(const VECType *, VECType *, double);                   \
  template bool HeatEq::HeatMat<D>::postMatVec(const VECType *, VECType *, double);

using ot::RankI;
#ifdef __clang__
DKT_INST_DIM(2)
DKT_INST_DIM(3)
DKT_INST_DIM(4)
#endif
