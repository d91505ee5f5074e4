/* litepgo_fem_ext.h -- entry points of the MI355X library that have no counterpart in the reference.
 * They tune or inspect the device execution and never change a bit of the results: every Hessian value, gradient
 * entry and energy is summed in an order fixed by the mesh alone (each vertex's incidence list, split over 4 lanes
 * in a fixed pattern), whatever the launch configuration. */
#ifndef LITEPGO_FEM_EXT_H
#define LITEPGO_FEM_EXT_H

#include "litepgo_fem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Launch shape of the assembly kernel: lanes per vertex (always 4: it fixes the summation order), waves per
 * workgroup in {1, 2, 4} (1 only when a vertex has more than 255 neighbours), vertices per tile
 * (<= 16 * waves) and the LDS budget, in doubles, for a tile's Hessian segment (a single vertex above it is
 * assembled in HBM instead, in the same summation order).  The budget plus 4 doubles per vertex slot must fit one
 * workgroup's LDS.  Used by the autotuner; the default is (4, 1, 16, 4096).  Other lane counts are rejected with
 * LPG_INVALID_ARGUMENT. */
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

/* Per-element Lame parameters from Young's modulus and Poisson ratio (device arrays of n entries), enqueued on
 * `stream` with no synchronisation: theta_dev (n x 2) <- (E nu / ((1 + nu)(1 - 2 nu)), E / (2 (1 + nu))).
 * lame_vjp gives the reverse product (gE, gnu) = (g_lambda, g_mu) d(lambda, mu)/d(E, nu) for the cotangent
 * gtheta_dev (n x 2); gE_dev or gnu_dev may be NULL.  No counterpart in the reference (its materials take
 * (lambda, mu) directly); used by the torch bridge to differentiate with respect to E and nu fields. */
lpg_status lpg_lame_device(const double* E_dev, const double* nu_dev, int64_t n, double* theta_dev, void* stream);
lpg_status lpg_lame_vjp_device(const double* E_dev, const double* nu_dev, const double* gtheta_dev, int64_t n,
                               double* gE_dev, double* gnu_dev, void* stream);

/* lpg_tet_energy_eval_device without any device allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  Arguments,
 * results (bit for bit) and errors are those of lpg_tet_energy_eval_device; the per-evaluation scratch (status and
 * energy slot, vertex energies, reduction partials and, for the Projected Hessian, the projection slots, list and
 * low-rank store) comes from one static 64 MiB device array per GPU, held until the process ends, instead of the
 * handle's buffers.  Limits: the handle's mesh must already be on the device (one evaluation through an allocating
 * entry point, or lpg_tet_energy_set_launch, puts it there; LPG_LOGIC_ERROR otherwise); a mesh whose scratch
 * exceeds 64 MiB (about 8 bytes per vertex, plus 8 per tet and 416 per tet that needs the eigen-analysis with the
 * Projected Hessian) fails with LPG_INVALID_ARGUMENT, as does an evaluation that lists more tets for the
 * eigen-analysis than the array holds.  Every call on one GPU uses the same array, whatever the handle: calls that
 * overlap in time on one GPU (other streams, other threads) race; calls in sequence on one stream are safe.  Graph
 * mode does not apply (the evaluation is enqueued eagerly and synchronised once to read the status). */
lpg_status lpg_tet_energy_eval_device_noalloc(lpg_tet_energy* energy, const double* x_dev, const double* theta_dev,
                                              int32_t want_gradient, int32_t hessian_kind, double* value,
                                              double* gradient_dev, double* hessian_values_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LITEPGO_FEM_EXT_H */
