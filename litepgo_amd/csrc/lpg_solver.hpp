// Launchers of the equilibrium solver's device kernels (lpg_solver.hip).
#pragma once

#include "lpg_internal.hpp"

namespace lpg {

// y = F (H p) + shift F p on the vertex-ordered CSR values H (F: free-DOF mask; p zero on fixed DOFs).
void launch_spmv(const DeviceMesh& m, const double* H, const uint8_t* fixed, const double* p, const double* shift,
                 double* y, hipStream_t s);
// minv = 1 / (diag(H) + shift) on free DOFs with a positive value, else 0.
void launch_jacobi(const DeviceMesh& m, const double* H, const uint8_t* fixed, const double* shift, double* minv,
                   hipStream_t s);
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
// Preconditioned CG pieces; scalars: [r.z, p.q, new r.z, r.r, breakdown flag].
void launch_pcg_start(const double* g, const uint8_t* fixed, const double* minv, double* r, double* d, double* z,
                      double* p, int n, hipStream_t s);
void launch_pcg_step(const double* q, double* r, double* d, double* z, const double* p, const double* minv,
                     double* scalars, int n, hipStream_t s);
void launch_pcg_direction(const double* z, double* p, double* scalars, int n, hipStream_t s);
// y = x + alpha v.
void launch_axpy(const double* x, double alpha, const double* v, double* y, int n, hipStream_t s);
// y = F (a x + b w) (x or w may be null: zero).
void launch_free_combine(const double* x, double a, const double* w, double b, const uint8_t* fixed, double* y, int n,
                         hipStream_t s);

} // namespace lpg
