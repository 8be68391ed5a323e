/* A plain C caller of the drop-in boundary (include/dkt.h), with no Python or
 * torch: builds the level-7 3D sphere tree (Hilbert order), its DA
 * and a host-side (dry-run) tiled plan, prints the counts and exits 0 on
 * success.  Built and run by tests/test_abi.py. */
#include <stdio.h>

#include "dkt.h"

#define CHECK(call)                                           \
  do {                                                        \
    if ((call) != 0) {                                        \
      fprintf(stderr, "%s: %s\n", #call, dkt_last_error());  \
      return 1;                                               \
    }                                                         \
  } while (0)

int main(void) {
  const double center[3] = {0.5, 0.5, 0.5};
  dkt_octs *cells = NULL, *tree = NULL;
  dkt_da *da = NULL;
  dkt_plan *plan = NULL;
  double info[23];
  CHECK(dkt_tree_sphere(3, 30, 1, 3, 7, center, 0.3, &cells));
  CHECK(dkt_tree_balance(cells, 1, &tree));
  CHECK(dkt_da_create(tree, 1, &da));
  CHECK(dkt_plan_create(da, DKT_OP_STIFFNESS, NULL, DKT_PLAN_DRYRUN, &plan));
  CHECK(dkt_plan_info(plan, info, 23));
  printf("elements %zu nodes %zu boundary %zu tiles %.0f\n", dkt_da_count(da, 0), dkt_da_count(da, 1),
         dkt_da_count(da, 2), info[5]);
  /* a dry-run plan refuses to run a matvec, with a message */
  if (dkt_plan_matvec(plan, 0, NULL, NULL, 1, 0, NULL) == 0 || dkt_last_error()[0] == '\0') return 2;
  dkt_plan_destroy(plan);
  dkt_da_destroy(da);
  dkt_octs_destroy(tree);
  dkt_octs_destroy(cells);
  return 0;
}
