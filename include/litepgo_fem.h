/* litepgo_fem.h -- C ABI of the MI355X (gfx950) implementation of litepgo's volumetric FEM hot path:
 * neo-Hookean Tet P1 energy, gradient and 12x12 element Hessian (Exact or PSD-projected), assembled into the
 * global sparse Hessian.  All arithmetic is FP64; all indices are int32, as in the reference's SpMat
 * (src/core/common/types.hpp:18-21).
 *
 * Conventions (identical to the reference):
 *   - unknowns are displacements u = x - X, DOF 3*v + c for vertex v and axis c (src/core/dof/vertex_dof_map.hpp:43-50);
 *   - theta is per element, (lambda, mu) (src/core/fem/material/volumetric/neo_hookean.hpp:9);
 *   - the Hessian pattern holds both triangles, sorted unique per column (energy_assembler.cpp:38-46).  It is
 *     structurally symmetric, so the CSC arrays (outer, inner) of the reference are bitwise the CSR arrays
 *     (row_ptr, col_idx) exported here, and H values are stored in that order;
 *   - the energy is E = sum_e w_e W(F_e) with w_e = det(Dm_e)/6 (volumetric_element.hpp:23-39).
 *
 * Errors follow the reference's exception classes (see lpg_status).  Overlapping calls on one handle are rejected
 * with LPG_BUSY, as MeshEnergy's BusyScope does (src/core/energy/scalar/mesh_energy.cpp:49).  A handle stays usable
 * after any error.
 *
 * Device-pointer entry points take a hipStream_t as void*; NULL means the default stream.  They return after the
 * work is finished (the status is read back).
 */
#ifndef LITEPGO_FEM_H
#define LITEPGO_FEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lpg_status {
    LPG_OK = 0,
    LPG_INVALID_ARGUMENT = 1, /* std::invalid_argument / std::out_of_range: bad x, theta, geometry or sizes */
    LPG_DOMAIN_ERROR = 2,     /* std::domain_error: det(F) <= 0, non-finite results, overflow */
    LPG_LOGIC_ERROR = 3,      /* std::logic_error: broken invariants */
    LPG_BUSY = 4,             /* overlapping call on one handle (BusyScope's std::logic_error) */
    LPG_DEVICE_ERROR = 5      /* HIP runtime failure or no usable gfx950 device */
} lpg_status;

/* HessianKind (src/core/energy/scalar/energy.hpp:6-13); LPG_HESSIAN_NONE = std::nullopt. */
typedef enum lpg_hessian_kind {
    LPG_HESSIAN_NONE = -1,
    LPG_HESSIAN_EXACT = 0,
    LPG_HESSIAN_PROJECTED = 1
} lpg_hessian_kind;

typedef struct lpg_tet_energy lpg_tet_energy;

/* Message of the last failed call on this thread (create included). Copies at most len-1 bytes. */
size_t lpg_last_error(char* buf, size_t len);

/* Replaces VolumetricFemEnergy<NeoHookean, TetP1Basis, TetCentroidQuadrature>(map, rest_positions, Fp, theta,
 * batch_size) (src/core/fem/assembly/volumetric_fem_energy.hpp:29-65), its MeshEnergy base
 * (src/core/energy/scalar/mesh_energy.cpp:30-42) and EnergyAssembler (src/core/dof/energy_assembler.cpp:22-67).
 *   tets  : num_tets x 4 vertex ids (host), rest: num_vertices x 3 rest positions (host),
 *   theta : num_tets x 2 (lambda, mu) (host), Fp: num_tets x 9 column-major or NULL for the identity.
 *   batch_size is accepted for interface parity (> 0 required); it does not change results.
 * Snapshots everything; nothing borrowed is kept.  Rest geometry is validated as VolumetricElement3D does
 * (invalid_argument on det(Dm) <= 0); theta is validated at evaluation, as NeoHookean does.
 * A plastic part Fp != I enters as MultiplicativeKinematics does (multiplicative_kinematics.cpp:9-51): Fe = F Fp^-1,
 * W = det(Fp) psi(Fe); a non-finite Fp (LPG_INVALID_ARGUMENT) or one with det(Fp) <= 0 or a non-finite inverse
 * (LPG_DOMAIN_ERROR) is accepted here and fails every evaluation, as the reference fails when evaluating. */
lpg_status lpg_tet_energy_create(int32_t num_vertices, int32_t num_tets, const int32_t* tets,
                                 const double* rest_positions, const double* theta, const double* Fp,
                                 int32_t batch_size, lpg_tet_energy** out);
void lpg_tet_energy_destroy(lpg_tet_energy* energy);

/* Energy::num_dofs (src/core/energy/scalar/energy.hpp:24-26) and sizes of the fixed pattern. */
int32_t lpg_tet_energy_num_dofs(const lpg_tet_energy* energy);
int32_t lpg_tet_energy_num_tets(const lpg_tet_energy* energy);
int32_t lpg_tet_energy_nnz(const lpg_tet_energy* energy);

/* FixedPatternEnergy's pattern (src/core/energy/scalar/fixed_pattern_energy.hpp:4-15): host copies of
 * row_ptr[num_dofs + 1] and col_idx[nnz] (== the reference's CSC outer/inner). */
lpg_status lpg_tet_energy_pattern(const lpg_tet_energy* energy, int32_t* row_ptr, int32_t* col_idx);

/* EnergyAssembler::hessian_offsets_ in the reference convention (src/core/dof/energy_assembler.cpp:56-67):
 * offsets[144 e + i + 12 j] = position of (row idx_e[i], col idx_e[j]) in the value array.  Host copy. */
lpg_status lpg_tet_energy_slot_map(const lpg_tet_energy* energy, int32_t* offsets);

/* Energy::eval (src/core/energy/scalar/energy.cpp:12-34) on host buffers: x[num_dofs] in, value, gradient
 * [num_dofs] (if want_gradient) and hessian_values[nnz] (unless kind is NONE) out. */
lpg_status lpg_tet_energy_eval(lpg_tet_energy* energy, const double* x, int32_t want_gradient,
                               int32_t hessian_kind, double* value, double* gradient, double* hessian_values);

/* The same on device buffers.  theta_dev (num_tets x 2) overrides the snapshot theta for this call when non-NULL.
 * value is a host pointer; gradient_dev / hessian_values_dev may be NULL when not requested. */
lpg_status lpg_tet_energy_eval_device(lpg_tet_energy* energy, const double* x_dev, const double* theta_dev,
                                      int32_t want_gradient, int32_t hessian_kind, double* value,
                                      double* gradient_dev, double* hessian_values_dev, void* stream);

/* Number of elements whose projection changed them (indefinite material tangent) in the last Projected eval. */
int32_t lpg_tet_energy_last_indefinite_count(const lpg_tet_energy* energy);

/* Energy::max_feasible_step (energy/scalar/energy.hpp:34-36, energy.cpp:35-45): the largest alpha in [0, 1] such
 * that every tet keeps det F > 0 on the whole segment x + t dx, t in [0, alpha] (a conservative bound: 0.9 of the
 * bisected first unsafe t, positive_det_feasible_step.cpp:13-84), the minimum over tets
 * (mesh_energy.cpp:88-107, volumetric_fem_energy.hpp:156-185, neo_hookean.cpp:130-145); 1 when nothing limits it.
 * x must be feasible (det F > 0 everywhere: LPG_DOMAIN_ERROR otherwise); non-finite x or dx: LPG_INVALID_ARGUMENT.
 * Host arrays of num_dofs doubles; the _device form takes device arrays and enqueues on `stream`. */
lpg_status lpg_tet_energy_max_feasible_step(lpg_tet_energy* energy, const double* x, const double* dx, double* alpha);
lpg_status lpg_tet_energy_max_feasible_step_device(lpg_tet_energy* energy, const double* x_dev, const double* dx_dev,
                                                   double* alpha, void* stream);

/* Differentiation form (device buffers).  Per-element parameter derivatives have layout num_tets x 2
 * (d/dlambda, d/dmu).  theta_dev may be NULL (snapshot theta).
 *   dE_dtheta      : dE/dtheta_e = w_e (L^2/2, (|F|^2-3)/2 - L), L = log det F (psi is linear in theta,
 *                    neo_hookean.cpp:33-36);
 *   force_vjp_theta: out_e = (d g / d theta_e)^T w, the parameter VJP of the force-material response
 *                    (fem/assembly/detail/volumetric_fem_force_response.hpp:155-179, energy/response/mesh_response.cpp:70-98);
 *   force_jvp_theta: out = sum_e (d g / d theta_e) dtheta_e (num_dofs);
 *   hvp            : out = H(x) v for the Exact Hessian, matrix-free (num_dofs). */
lpg_status lpg_tet_energy_dE_dtheta_device(lpg_tet_energy* energy, const double* x_dev, const double* theta_dev,
                                           double* out_dev, void* stream);
lpg_status lpg_tet_energy_force_vjp_theta_device(lpg_tet_energy* energy, const double* x_dev,
                                                 const double* theta_dev, const double* w_dev, double* out_dev,
                                                 void* stream);
lpg_status lpg_tet_energy_force_jvp_theta_device(lpg_tet_energy* energy, const double* x_dev,
                                                 const double* theta_dev, const double* dtheta_dev, double* out_dev,
                                                 void* stream);
lpg_status lpg_tet_energy_hvp_device(lpg_tet_energy* energy, const double* x_dev, const double* theta_dev,
                                     const double* v_dev, double* out_dev, void* stream);
/* Force response to the plastic parts (VolumetricFemForceFpResponse, fem/assembly/volumetric_fem_force_fp_response.hpp:
 * detail::VolumetricFemForceResponse with ForceInput::Plastic, fem/assembly/detail/volumetric_fem_force_response.hpp:
 * 81-179, and NeoHookean::eval_d2W_dF_dFp, neo_hookean.cpp:103-128): f(Fp) = dE/dx at fixed x and theta, as a
 * function of the per-tet plastic parts.  fp_dev: num_tets x 9, column-major vec(Fp) per tet (det Fp > 0), or NULL
 * for the handle's own Fp.  theta_dev may be NULL (snapshot theta).
 *   force_fp    : out = f (num_dofs);
 *   force_vjp_fp: out_e = (d f / d vec(Fp_e))^T w (num_tets x 9, column-major);
 *   force_jvp_fp: out = sum_e (d f / d vec(Fp_e)) dFp_e (num_dofs), dfp_dev num_tets x 9 column-major. */
lpg_status lpg_tet_energy_force_fp_device(lpg_tet_energy* energy, const double* x_dev, const double* theta_dev,
                                          const double* fp_dev, double* out_dev, void* stream);
lpg_status lpg_tet_energy_force_vjp_fp_device(lpg_tet_energy* energy, const double* x_dev, const double* theta_dev,
                                              const double* fp_dev, const double* w_dev, double* out_dev,
                                              void* stream);
lpg_status lpg_tet_energy_force_jvp_fp_device(lpg_tet_energy* energy, const double* x_dev, const double* theta_dev,
                                              const double* fp_dev, const double* dfp_dev, double* out_dev,
                                              void* stream);
/* Reverse product of (E, f = dE/dx) in one call, the fused backward of litepgo.torch's energy and response
 * adapters (src/python/torch.py:139-142 gives upstream * gradient, :154-157 the response VJP):
 *   gx     = s f + H w                       (num_dofs; H the Exact Hessian),
 *   gtheta = s dE/dtheta + (df/dtheta)^T w   (num_tets x 2),
 * with s = *gE_dev (a device scalar; 0 when gE_dev is NULL) and w = w_dev (a zero vector when NULL).  gx_dev or
 * gtheta_dev may be NULL to skip that output.  One status read-back for both outputs. */
lpg_status lpg_tet_energy_vjp_device(lpg_tet_energy* energy, const double* x_dev, const double* theta_dev,
                                     const double* gE_dev, const double* w_dev, double* gx_dev, double* gtheta_dev,
                                     void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LITEPGO_FEM_H */
