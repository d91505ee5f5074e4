// Launchers of the equilibrium solver's device kernels (lpg_solver.hip).
#pragma once

#include "lpg_internal.hpp"

namespace lpg {

// y = F (H p) + shift F p on the vertex-ordered CSR values H (F: free-DOF mask; p zero on fixed DOFs).
void launch_spmv(const DeviceMesh& m, const double* H, const uint8_t* fixed, const double* p, const double* shift,
                 double* y, hipStream_t s);
// diag = diagonal of H on free DOFs (0 on fixed ones).
void launch_diag(const DeviceMesh& m, const double* H, const uint8_t* fixed, double* diag, hipStream_t s);
// a *= f over n entries.
void launch_scale(double* a, double f, int64_t n, hipStream_t s);
// out[0] = a.b (and out[1] = c.d when c is given), fixed-order; part holds dot_partials() doubles.
void launch_dot(const double* a, const double* b, const double* c, const double* d, int n, double* part, double* out,
                hipStream_t s);
int dot_partials();
// out[0] = max |a_i| over free DOFs.
void launch_free_maxabs(const double* a, const uint8_t* fixed, int n, double* out, hipStream_t s);
// Fused preconditioned CG (two kernels per iteration, 3x3 block-Jacobi preconditioner).  Device scalars: [0, 1] r.z
// of the last two iterations (a ring indexed by parity), 3 r.r (after launch_pcg_finish), 4 breakdown flag.
// `part` holds pcg_partials() doubles of per-workgroup partial sums passed from each kernel to the next.
// binv: 6 doubles per vertex, the symmetric inverse of the free part of the vertex's diagonal block (+ shift).
void launch_block_jacobi(const DeviceMesh& m, const double* H, const uint8_t* fixed, const double* shift, double* binv,
                         hipStream_t s);
// r = -F g, d = 0, z = B^-1 r (partials of r.z, r.r); scalars[4] = 0.
void launch_pcg_begin(const DeviceMesh& m, const double* g, const uint8_t* fixed, const double* binv, double* r,
                      double* d, double* z, double* scalars, double* part, hipStream_t s);
// scalars[3] = r.r of the last begin / update.
// With rb (pinned host memory, device-visible): rb[0] = r.r, rb[1] = breakdown flag, then the word at rb + 2 = seq.
void launch_pcg_finish(const DeviceMesh& m, double* scalars, const double* part, hipStream_t s, double* rb = nullptr,
                       unsigned seq = 0);
// Iteration it >= 1: p = z + beta p_old (it = 1: p = z), q = F (H p) + shift F p (partials of p.q).
void launch_pcg_apply(const DeviceMesh& m, const double* H, const uint8_t* fixed, const double* z, const double* p_old,
                      double* p, double* q, const double* shift, int it, double* scalars, double* part,
                      hipStream_t s);
// Iteration it: d += alpha p, r -= alpha q, z = B^-1 r (partials of the new r.z and r.r).
void launch_pcg_update(const DeviceMesh& m, const double* p, const double* q, double* d, double* r, double* z,
                       const double* binv, int it, double* scalars, double* part, hipStream_t s);
int pcg_partials();
// y = x + alpha v.
void launch_axpy(const double* x, double alpha, const double* v, double* y, int n, hipStream_t s);
// y = F (a x + b w) (x or w may be null: zero).
void launch_free_combine(const double* x, double a, const double* w, double b, const uint8_t* fixed, double* y, int n,
                         hipStream_t s);

} // namespace lpg
