/*
 * CPU restatement of the reference Dendro-KT matvec traversal -- test
 * infrastructure (the oracle).  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg load it, as the checker; the product never does.
 */
#ifndef DKT_MATVEC_ORACLE_H
#define DKT_MATVEC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* op: 0 Poisson stiffness h^(dim-2) K_ref, 1 mass h^dim M_ref, 2 identity.
 * curve: 0 Morton, 1 Hilbert (child visiting order, as the reference build
 * option HILBERT_ORDERING).  nodes: uint32 [n][dim+1] in DA order.
 * masked: zero x and y at domain-boundary nodes (feMatrix::matVec with
 * HeatMat's Dirichlet hooks).  Returns 0 on success. */
int oracle_matvec(int dim, int maxdepth, int order, int curve, const uint32_t *nodes, size_t n, int op,
                  const double *x, double *y, int masked);

/* Resets the emulated static buffers to their first-call state (zeros). */
void oracle_reset(void);

#ifdef __cplusplus
}
#endif
#endif
