/* litepgo_fem_batch.h -- many independent meshes evaluated together, and launch control for repeated calls.
 *
 * None of these entry points has a counterpart in the reference: the reference batches the elements of one mesh
 * (src/core/energy/scalar/mesh_energy.cpp:30-42) and has no multi-mesh batcher.  A batch handle is an
 * lpg_tet_energy over the union of the meshes (block-diagonal pattern, mesh-local vertex numbering shifted by the
 * mesh's vertex offset); every lpg_tet_energy_* entry point applies to it, and each mesh's energy, gradient rows
 * and Hessian values equal that mesh's own evaluation (the union-equals-sum property of
 * tests/.../volumetric_fem_energy_tet_p1_gtest.cpp:226-240, per mesh).  Tiles never cross mesh boundaries and
 * each mesh's energy is reduced separately. */
#ifndef LITEPGO_FEM_BATCH_H
#define LITEPGO_FEM_BATCH_H

#include "litepgo_fem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* num_vertices[m], num_tets[m] per mesh; tets: concatenated connectivity with mesh-local vertex ids; rest
 * (3 per vertex) and theta (2 per tet) concatenated in mesh order.  Validation as lpg_tet_energy_create, per
 * mesh. */
lpg_status lpg_tet_batch_create(int32_t num_meshes, const int32_t* num_vertices, const int32_t* num_tets,
                                const int32_t* tets, const double* rest_positions, const double* theta,
                                lpg_tet_energy** out);
int32_t lpg_tet_batch_num_meshes(const lpg_tet_energy* energy);
/* Offsets (num_meshes + 1 entries each) of every mesh's vertices, tets and Hessian values in the union; a mesh's
 * DOFs start at 3 * vertex_offsets[m]. */
lpg_status lpg_tet_batch_offsets(const lpg_tet_energy* energy, int32_t* vertex_offsets, int32_t* tet_offsets,
                                 int32_t* nnz_offsets);
/* One evaluation of every mesh: energies_dev[num_meshes] (device) receives the per-mesh energies; gradient and
 * Hessian values as lpg_tet_energy_eval_device over the union.  Synchronises on `stream` to read the status. */
lpg_status lpg_tet_batch_eval_device(lpg_tet_energy* energy, const double* x_dev, const double* theta_dev,
                                     int32_t want_gradient, int32_t hessian_kind, double* energies_dev,
                                     double* gradient_dev, double* hessian_values_dev, void* stream);

/* The evaluation of lpg_tet_batch_eval_device / lpg_tet_energy_eval_device split in two, so that a caller can
 * enqueue the evaluations of many handles (one per mesh, say) on one stream and synchronise once.  No counterpart
 * in the reference, whose MeshEnergy::evaluate is synchronous (src/core/energy/scalar/mesh_energy.cpp:30-42).
 * _enqueue checks the arguments, enqueues the kernels and the read-back of the status and returns without
 * synchronising; energies_dev as lpg_tet_batch_eval_device (null: the value is reported by _wait only).  The
 * handle stays busy (other calls fail with LPG_BUSY) until lpg_tet_energy_eval_wait, which synchronises the stream,
 * maps the status to an error as the synchronous call does, and, when energies_dev was null, stores the energy
 * (for a batch handle, the sum over its meshes) in *value (may be null; not written when energies_dev was given).
 * The buffers must stay valid until then.  Graph mode does not apply. */
lpg_status lpg_tet_energy_eval_enqueue(lpg_tet_energy* energy, const double* x_dev, const double* theta_dev,
                                       int32_t want_gradient, int32_t hessian_kind, double* energies_dev,
                                       double* gradient_dev, double* hessian_values_dev, void* stream);
lpg_status lpg_tet_energy_eval_wait(lpg_tet_energy* energy, double* value);

/* Per-mesh feasible steps: alphas_dev[num_meshes] (device) receives each mesh's max_feasible_step of x along dx
 * (lpg_tet_energy_max_feasible_step restricted to the mesh's tets).  Synchronises on `stream` to read the status. */
lpg_status lpg_tet_batch_max_feasible_step_device(lpg_tet_energy* energy, const double* x_dev, const double* dx_dev,
                                                  double* alphas_dev, void* stream);

/* Graph mode (any handle): the first device evaluation with a given set of buffers, kind and launch plan runs
 * eagerly; a repeat of it is captured into a hipGraph (kernels plus the status read-back) and later repeats
 * replay the graph, one launch per evaluation.  Changing buffers, kind or plan falls back to the eager path.  A
 * replay returns as soon as its last kernel has written the energy, status word and list count to the handle's
 * pinned memory, without synchronising the stream: every kernel that writes the gradient or Hessian values has
 * finished by then, and later work on `stream` stays ordered after the replay. */
lpg_status lpg_tet_energy_set_graph(lpg_tet_energy* energy, int32_t enable);

/* Autotuner: times the assembly on this handle's own data (x_dev, the given kind) for each candidate launch
 * shape (waves, vertices per tile, LDS budget; 4 lanes per vertex) and, for batches, persistent grid size, keeps
 * the fastest, and reports it in chosen[5] = {lanes, waves, vertices per tile, LDS budget, persistent grid} with its
 * mean assembly time (ms, repeats timed runs each).  The choice is cached per process, keyed by the device, the
 * Hessian kind and the handle's mesh-size histogram (mesh count, tets per mesh in power-of-two bins, vertex degree
 * class, plastic parts): a later call on a handle with the same key applies the cached choice without timing
 * (repeats > 0) or times again and replaces it (repeats < 0, |repeats| runs).  No shape changes a bit of the
 * results. */
lpg_status lpg_tet_energy_autotune(lpg_tet_energy* energy, const double* x_dev, int32_t hessian_kind,
                                   int32_t repeats, int32_t* chosen, double* best_ms);

#ifdef __cplusplus
}
#endif

#endif
