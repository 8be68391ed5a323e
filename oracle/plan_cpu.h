/*
 * CPU restatement of the matvec over the flattened traversal (see
 * plan_cpu.c) -- test infrastructure and the CPU baseline; the product never
 * links it.
 */
#ifndef DKT_PLAN_CPU_H
#define DKT_PLAN_CPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* y = A x over the maps of dkt_da_traversal (leaf_cells [L][dim+1], child [L],
 * slot [L][nPe], hang_of [L], pslot [H][nPe]) and dkt_da_stale_terms (st_*),
 * op 0 stiffness, 1 mass, 2 identity; nthreads OpenMP threads.  Returns 0 on
 * success. */
int oracle_plan_matvec(int dim, int p, size_t n_nodes, size_t L, const uint32_t *leaf_cells, const uint8_t *child,
                       const int32_t *slot, const int32_t *hang_of, const int32_t *pslot, const int64_t *st_off,
                       const uint8_t *st_rank, const int32_t *st_node, const double *st_w, int op, const double *x,
                       double *y, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
