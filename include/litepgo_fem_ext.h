/* litepgo_fem_ext.h -- entry points of the MI355X library that have no counterpart in the reference.
 * They tune or inspect the device execution and never change results beyond floating-point reassociation
 * (the assembly order stays the element order for every configuration). */
#ifndef LITEPGO_FEM_EXT_H
#define LITEPGO_FEM_EXT_H

#include "litepgo_fem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Launch shape of the assembly kernel: lanes per vertex (1, 2, 4), waves per workgroup (1, 2, 4), vertices per
 * tile (<= 64 * waves / lanes_per_vertex) and the LDS budget for a tile's Hessian segment, in doubles.  Used by
 * the autotuner; the default is (2, 1, 32, 6144). */
lpg_status lpg_tet_energy_set_launch(lpg_tet_energy* energy, int32_t lanes_per_vertex, int32_t waves,
                                     int32_t vertices_per_tile, int32_t lds_budget_doubles);
/* cfg4 <- {lanes_per_vertex, waves, vertices_per_tile, lds_budget_doubles}. */
lpg_status lpg_tet_energy_get_launch(const lpg_tet_energy* energy, int32_t* cfg4);

/* Per-phase device timing with HIP events recorded on the evaluation's stream.  set_timing(1) enables it and
 * resets the totals; kernel_times fills ms[4] with cumulative milliseconds of {projection prepass + eigen-clamp,
 * assembly kernel(s), energy reduction, whole evaluation from the first kernel to the last} and *evals with the
 * number of timed evaluations. */
lpg_status lpg_tet_energy_set_timing(lpg_tet_energy* energy, int32_t enable);
lpg_status lpg_tet_energy_kernel_times(const lpg_tet_energy* energy, double* ms, int32_t* evals);

/* Largest vertex degree (distinct neighbours incl. itself) of the mesh. */
int32_t lpg_tet_energy_max_degree(const lpg_tet_energy* energy);

#ifdef __cplusplus
}
#endif

#endif /* LITEPGO_FEM_EXT_H */
