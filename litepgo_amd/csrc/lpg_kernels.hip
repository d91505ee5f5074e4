// HIP kernels for gfx950: neo-Hookean Tet P1 energy / gradient / Hessian assembly and its derivatives.
//
// Layout in HBM (lpg_internal.hpp): rest positions Xr[V] (double4), the displacement vector x[3V], connectivity
// conn[T] (int4), theta[T] (double2), per-vertex sorted neighbour lists and incidence lists (tet id + the position of
// each of the tet's vertices in this vertex's neighbour list, which is also its column block in the vertex's rows),
// and the CSR row pointer.  No per-tet geometry is stored: Dm^-1 and w = det(Dm)/6 are recomputed from the four
// rest positions (~50 FP64 operations), which saves 80 bytes of HBM traffic per tet.
//
// Assembly is vertex-owned: the rows 3v..3v+2 of H form one contiguous segment of 9 |nbr(v)| values, and only
// the lanes of vertex v write it.  Four lanes per vertex split its incident tets (lane s takes incidences s, s + 4,
// ...), rebuild F, J, F^-T and log J in registers, and add the three rows of the tet's 12x12 Hessian that belong to
// v, block by block, to the segment staged in LDS.  No atomics on HBM, and every H value is written to HBM once,
// coalesced.  The order in which each value is summed is fixed by the vertex's incidence list (element order, the
// reference's dof_contributions_ order of src/core/dof/energy_assembler.cpp:48-54) and the fixed 4-lane split, not
// by the launch configuration, so results are bitwise identical for every launch shape, tile size and persistent
// grid, and a mesh evaluated inside a batch gives the bits of its own evaluation (the reference guarantees the same
// across thread counts and batch sizes: assembly at energy_assembler.cpp:123-170).  The lanes' partial sums are
// not the reference's plain element-order sums, so values agree with the reference's to rounding.
//
// The element Hessian uses the closed form of B^T A B for Tet P1 and neo-Hookean (A from neo_hookean.cpp:43-60):
//   H_ab = w [ mu (g_a . g_b) I3 + lambda q_a q_b^T - p q_b q_a^T ],  q_b = F^-T g_b,  p = lambda log J - mu,
// where g_b is row b of G = dN/dxi Dm^-1.  The gradient block is w (mu F g_a + p q_a) and the energy
// w (mu/2 (|F|^2 - 3) - mu log J + lambda/2 log^2 J).
//
// The Projected Hessian clamps the spectrum of the 12x12 element matrix (mesh_energy.cpp:16-27).  With the
// orthonormal Helmert basis P (4x3) of the complement of (1,1,1,1), H_e = Q M Q^T with Q = P (x) I3 and the 9x9
// M = w (C (x) I3) A (C (x) I3)^T, C = P^T G; hence proj(H_e) = Q proj(M) Q^T, and by Sylvester's law M is PSD
// exactly when A is.  A prepass certifies A positive definite in closed form (eig9::tangent_certified_pd); only the
// other elements get the 9x9 eigen-clamp of M, and the assembly reconstructs their blocks from proj(M).
#include "lpg_internal.hpp"
#include "lpg_eig9.hpp"
#include "lpg_math.hpp"

#include <algorithm>
#include <cmath>

// No implicit FMA contraction: every fused multiply-add below is an explicit fma(), so each value is computed by the
// same operation sequence in every template instantiation and launch shape, and results are bitwise reproducible
// across configurations.
#pragma clang fp contract(off)

namespace lpg {
namespace {

__device__ __forceinline__ double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
    return fma(a2, b2, fma(a1, b1, a0 * b0));
}

constexpr int kWave = 64;

// Helmert basis of 1-perp in R^4: P(a, m) (orthonormal columns orthogonal to (1, 1, 1, 1)).  Indexed with
// compile-time indices only, so its zeros fold away.
constexpr double kP[4][3] = {
    {0.70710678118654752440, 0.40824829046386301637, 0.28867513459481288225},
    {-0.70710678118654752440, 0.40824829046386301637, 0.28867513459481288225},
    {0.0, -0.81649658092772603273, 0.28867513459481288225},
    {0.0, 0.0, -0.86602540378443864676},
};

// Tet inputs of one evaluation: rest and current vertex data, material.
struct TetIn {
    double X[4][3];
    double u[4][3];
    double lam, mu;
    // Plastic part of the tet (row-major Fp^-1, then det Fp: 10 doubles), or null for Fp = I.
    const double* fp = nullptr;
};

__device__ __forceinline__ void load_vertices(const double* __restrict__ Xr, const double* __restrict__ x,
                                              const int ids[4], TetIn& t) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const double4 X = reinterpret_cast<const double4*>(Xr)[ids[b]];
        t.X[b][0] = X.x;
        t.X[b][1] = X.y;
        t.X[b][2] = X.z;
        const double* p = x + 3 * static_cast<size_t>(ids[b]);
        t.u[b][0] = p[0];
        t.u[b][1] = p[1];
        t.u[b][2] = p[2];
    }
}

// (lambda, mu) of tet e from the per-call override, else from the snapshot taken at construction.
__device__ __forceinline__ void load_theta(const double* __restrict__ snapshot, const double* __restrict__ override_,
                                           int e, TetIn& t) {
    const double2 th = reinterpret_cast<const double2*>(override_ ? override_ : snapshot)[e];
    t.lam = th.x;
    t.mu = th.y;
}

// Element-wise kernels: connectivity (int4) and theta (double2) of tet e are read coalesced, vertex data gathered.
__device__ __forceinline__ void load_tet(const double* __restrict__ Xr, const int32_t* __restrict__ conn,
                                         const double* __restrict__ snapshot, const double* __restrict__ x,
                                         const double* __restrict__ override_, int e, int ids[4], TetIn& t,
                                         const double* __restrict__ fp = nullptr) {
    t.fp = fp ? fp + 10 * static_cast<size_t>(e) : nullptr;
    const int4 c = reinterpret_cast<const int4*>(conn)[e];
    ids[0] = c.x;
    ids[1] = c.y;
    ids[2] = c.z;
    ids[3] = c.w;
    load_vertices(Xr, x, ids, t);
    load_theta(snapshot, override_, e, t);
}

// Rest geometry: Dm = [X1-X0 | X2-X0 | X3-X0], Dm^-1 = cof(Dm)^T / det, w = det / 6.  The host applies the same
// operation sequence when it validates the mesh (det > 0, finite Dm^-1), so a tet accepted there is valid here.
// With a plastic part Fp (multiplicative_kinematics.cpp:9-51), Fe = F Fp^-1 = Fp^-1 + Du (Dm^-1 Fp^-1) and
// W = det(Fp) psi(Fe): the element is a plain neo-Hookean tet with rest edges Fp Dm, so Di holds Dm^-1 Fp^-1, w is
// scaled by det Fp and F0 = Fp^-1 replaces the identity in F.
struct Geo {
    double Di[9]; // Dm^-1 (Fp^-1) row-major
    double w;
    bool plastic = false;
    double F0[9];  // Fp^-1 row-major (plastic only)
};

__device__ __forceinline__ void rest_geometry(const TetIn& t, Geo& g) {
    double Dm[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            Dm[3 * r + k] = t.X[k + 1][r] - t.X[0][r];
    double C[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            C[3 * i + j] = fma(Dm[3 * i1 + j1], Dm[3 * i2 + j2], -(Dm[3 * i1 + j2] * Dm[3 * i2 + j1]));
        }
    const double det = dot3(Dm[0], Dm[1], Dm[2], C[0], C[1], C[2]);
    const double inv = 1.0 / det;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            g.Di[3 * i + j] = C[3 * j + i] * inv;
    g.w = det * (1.0 / 6.0);
    g.plastic = t.fp != nullptr;
    if (g.plastic) {
#pragma unroll
        for (int q = 0; q < 9; ++q)
            g.F0[q] = t.fp[q];
        double Dp[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Dp[3 * i + j] = dot3(g.Di[3 * i], g.Di[3 * i + 1], g.Di[3 * i + 2], g.F0[j], g.F0[3 + j], g.F0[6 + j]);
#pragma unroll
        for (int q = 0; q < 9; ++q)
            g.Di[q] = Dp[q];
        g.w *= t.fp[9];
    }
}

// Deformation gradient and its inverse transpose: F = I + Du Dm^-1 (Fp^-1 + Du Dm^-1 Fp^-1 with a plastic part),
// a = F^-T = cof(F) / det F.
struct Kin {
    double F[9]; // row-major
    double a[9]; // row-major, a(i, j) = F^-T(i, j)
    double J, L;
};

__device__ __forceinline__ void kinematics(const TetIn& t, const Geo& geo, Kin& k) {
    const double* Di = geo.Di;
    double Du[9]; // Du(i, c) = (u_{c+1} - u_0)_i
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            Du[3 * i + c] = t.u[c + 1][i] - t.u[0][i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = geo.plastic ? geo.F0[3 * i + j] : ((i == j) ? 1.0 : 0.0);
#pragma unroll
            for (int c = 0; c < 3; ++c)
                s = fma(Du[3 * i + c], Di[3 * c + j], s);
            k.F[3 * i + j] = s;
        }
    double C[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            C[3 * i + j] = fma(k.F[3 * i1 + j1], k.F[3 * i2 + j2], -(k.F[3 * i1 + j2] * k.F[3 * i2 + j1]));
        }
    k.J = dot3(k.F[0], k.F[1], k.F[2], C[0], C[1], C[2]);
    const double inv = 1.0 / k.J;
#pragma unroll
    for (int q = 0; q < 9; ++q)
        k.a[q] = C[q] * inv;
    k.L = log(k.J);
}

// Rows of G: g_0 = -(g_1 + g_2 + g_3), g_b = row b-1 of Dm^-1.
__device__ __forceinline__ void shape_gradients(const double Di[9], double G[4][3]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        G[1][j] = Di[j];
        G[2][j] = Di[3 + j];
        G[3][j] = Di[6 + j];
        G[0][j] = -(Di[j] + Di[3 + j] + Di[6 + j]);
    }
}

__device__ __forceinline__ bool kin_ok(const Kin& k) {
    return k.J > 0.0 && isfinite(k.J) && isfinite(k.L);
}

__device__ __forceinline__ bool theta_ok(double lam, double mu) {
    return isfinite(lam) && isfinite(mu) && lam >= 0.0 && mu > 0.0;
}

// Adds into a value owned by this lane: an LDS atomic add (one instruction, no read-back) when staged in LDS,
// a plain read-modify-write when accumulating in HBM.
template <bool GLOBAL_ACC> __device__ __forceinline__ void accum(double* p, double v) {
    if (GLOBAL_ACC)
        *p += v;
    else
        __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Accessors of the incidence records: cp(r, j) is the column-block position of the vertex at rotated slot j = 1..3,
// slot(r) the vertex's slot a in the tet's own order.
template <bool WIDE> struct IncT;
template <> struct IncT<false> {
    using type = IncNarrow;
    __device__ static int cp(const IncNarrow& r, int j) {
        return (r.packed >> (8 * (j - 1))) & 0xff;
    }
    __device__ static int slot(const IncNarrow& r) {
        return (r.packed >> 24) & 3;
    }
};
template <> struct IncT<true> {
    using type = IncWide;
    __device__ static int cp(const IncWide& r, int j) {
        return j == 1 ? (r.cp12 & 0xffff) : j == 2 ? (r.cp12 >> 16) : (r.cp3a & 0xffff);
    }
    __device__ static int slot(const IncWide& r) {
        return (r.cp3a >> 16) & 3;
    }
};

// 1 / sqrt(x) for finite x > 0 from the hardware estimate refined by Newton steps (within an ulp or so; only the
// eigen-clamp's rotation angles use it, where c^2 + s^2 = 1 to rounding is what matters).
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = y * fma(-hx * y, y, 1.5);
    return y * fma(-hx * y, y, 1.5);
}

// 1 / d from the hardware reciprocal estimate refined by two Newton steps (within an ulp of the quotient for the
// normal range used here; 0 gives NaN and the caller's finiteness checks fire).
__device__ __forceinline__ double recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// Row `a` of a 4x3 register array as a masked sum.  A select chain here is rewritten by the compiler into a
// dynamically indexed private array (scratch memory); multiplying by 0.0 / 1.0 masks keeps it in registers.
__device__ __forceinline__ void select_row(const double G[4][3], int a, double out[3]) {
    const double m0 = a == 0 ? 1.0 : 0.0, m1 = a == 1 ? 1.0 : 0.0, m2 = a == 2 ? 1.0 : 0.0, m3 = a == 3 ? 1.0 : 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
        out[j] = fma(m3, G[3][j], fma(m2, G[2][j], fma(m1, G[1][j], m0 * G[0][j])));
}

// ---------------------------------------------------------------------------------------------------------------
// Assembly kernel: one workgroup per vertex tile.  The LPV (= 4) lanes of a vertex split its incidence list (lane s
// takes incidences s, s + LPV, s + 2 LPV, ...).  A lane evaluates its tet in the vertex order (a, a^1, a^2, a^3),
// which puts the vertex at local slot 0, so the three rows of the element Hessian that belong to the vertex are the
// blocks H_0j.  The diagonal block H_00 and the gradient and energy contributions are summed per lane in registers
// and combined in lane order at the end of the tile; the off-diagonal blocks H_0j (j = 1..3) are added to the
// vertex's H segment in LDS with ds_add_f64, block by block, by all lanes at once.  Lanes of one vertex collide on an
// address only when their tets share the neighbour; colliding adds within one LDS instruction are applied in
// ascending lane order, so every value is summed in (round, block, lane) order whatever the tile size or the waves
// per workgroup.  Tiles above the LDS budget accumulate in HBM with the lanes of a vertex taking turns per block,
// the same order.  KIND: -1 none, 0 Exact, 1 Projected.  PLASTIC: the handle has plastic parts Fp.
// ---------------------------------------------------------------------------------------------------------------
struct AssembleParams {
    const double* Xr;
    const double* theta_snap;
    const int32_t* inc_ptr;
    const void* inc;
    const int32_t* nbr_ptr;
    const int32_t* nbr;
    const int32_t* self_pos;
    const int32_t* tile_vbeg;
    const int32_t* tile_ids; // tiles to process (index into tile_vbeg), or null for 0..num_tiles-1
    const double* x;
    const double* theta; // per-call override or null
    double* g;
    double* H;
    double* vertex_E; // per-vertex sum of the energies of the tets whose slot 0 it is
    unsigned* status;
    const int32_t* proj_slot;
    const double* proj_H;
    const double* fp = nullptr; // plastic parts (10 per tet) or null
    int proj_capacity = 0;      // list entries proj_H holds (an overflowing evaluation is repeated by the host)
};

// Rest and current positions of a vertex.
struct VertexData {
    double X[3];
    double u[3];
};

// The values of one incidence, vertex v at rotated slot 0 of its tet, neighbours at slots 1..3: the gradient block
// gc = w (mu F g_0 + p q_0), the tet's energy ec when v is the tet's slot 0 (a == 0), the 6 distinct entries of
// the symmetric diagonal block H_00, and each off-diagonal block H_0j passed to emit(j, block[9]) as soon as it is
// formed.  With Ds = [x1-x0 | x2-x0 | x3-x0] (deformed edges), Dm the rest edges and F = Ds Dm^-1 (Fp^-1):
//   g_j (j = 1..3) = rows of Dm^-1 (times Fp^-1),  g_0 = -(g_1 + g_2 + g_3),
//   q_j = rows of Ds^-1 = F^-T g_j,  q_0 = -(q_1 + q_2 + q_3),  J = det Ds / (det Dm det Fp),  p = lambda log J - mu,
//   H_0j = w [mu (g_0 . g_j) I + lambda q_0 q_j^T - p q_j q_0^T],  H_00 = w [mu |g_0|^2 I + (lambda - p) q_0 q_0^T],
//   F g_0 = Ds (g_1 . g_0, g_2 . g_0, g_3 . g_0)^T  (Dm^-1 has rows g_j),
// (neo_hookean.cpp:33-60 for psi, P and A; B^T A B of dof_derivatives.cpp:22-51 in closed form for Tet P1).
template <int KIND, bool PLASTIC, class Emit>
__device__ __forceinline__ void incidence_row(const VertexData& v0, const VertexData (&nb)[3], double lam, double mu,
                                              const double* __restrict__ fp, bool energy, double diag[6],
                                              double gc[3], double& ec, unsigned& flags, Emit&& emit) {
    double Dm[9], Cm[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            Dm[3 * r + c] = nb[c].X[r] - v0.X[r];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            Cm[3 * i + j] = fma(Dm[3 * i1 + j1], Dm[3 * i2 + j2], -(Dm[3 * i1 + j2] * Dm[3 * i2 + j1]));
        }
    const double det_m = dot3(Dm[0], Dm[1], Dm[2], Cm[0], Cm[1], Cm[2]);
    const double inv_m = recip(det_m);
    double G[3][3]; // rows g_1..g_3
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            G[b][j] = Cm[3 * j + b] * inv_m;
    double w = det_m * (1.0 / 6.0);
    double inv_mp = inv_m;
    if (PLASTIC) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double g0 = G[b][0], g1 = G[b][1], g2 = G[b][2];
#pragma unroll
            for (int j = 0; j < 3; ++j)
                G[b][j] = dot3(g0, g1, g2, fp[j], fp[3 + j], fp[6 + j]);
        }
        w *= fp[9];
        inv_mp = inv_m * recip(fp[9]);
    }
    double g0[3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
        g0[j] = -(G[0][j] + G[1][j] + G[2][j]);
    double Ds[9], Cs[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            Ds[3 * r + c] = Dm[3 * r + c] + (nb[c].u[r] - v0.u[r]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            Cs[3 * i + j] = fma(Ds[3 * i1 + j1], Ds[3 * i2 + j2], -(Ds[3 * i1 + j2] * Ds[3 * i2 + j1]));
        }
    const double det_s = dot3(Ds[0], Ds[1], Ds[2], Cs[0], Cs[1], Cs[2]);
    const double inv_s = recip(det_s);
    const double J = det_s * inv_mp;
    const double L = log_pos(J);
    if (!(det_m > 0.0 && J > 0.0 && isfinite(J) && isfinite(L)))
        flags |= kStatusDetF;
    if (!theta_ok(lam, mu))
        flags |= kStatusBadTheta;
    double q[3][3], q0[3];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            q[b][j] = Cs[3 * j + b] * inv_s;
#pragma unroll
    for (int j = 0; j < 3; ++j)
        q0[j] = -(q[0][j] + q[1][j] + q[2][j]);
    // d_j = g_0 . g_j (j = 1..3): the off-diagonal mu terms and Dm^-1 g_0.
    double d[3];
#pragma unroll
    for (int b = 0; b < 3; ++b)
        d[b] = dot3(g0[0], g0[1], g0[2], G[b][0], G[b][1], G[b][2]);
    const double p_ = fma(lam, L, -mu);
    const double wmu = w * mu, wlam = w * lam, wp = w * p_;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double Fg0 = dot3(Ds[3 * i], Ds[3 * i + 1], Ds[3 * i + 2], d[0], d[1], d[2]);
        gc[i] = fma(wmu, Fg0, wp * q0[i]);
    }
    if (energy) {
        // |F|^2 with F = Ds [g_1; g_2; g_3]: psi = mu/2 (|F|^2 - 3) - mu L + lambda/2 L^2.
        double sq = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double Fij = dot3(Ds[3 * i], Ds[3 * i + 1], Ds[3 * i + 2], G[0][j], G[1][j], G[2][j]);
                sq = fma(Fij, Fij, sq);
            }
        ec = w * fma(0.5 * lam * L, L, fma(0.5 * mu, sq - 3.0, -(mu * L)));
    }
    if (KIND < 0)
        return;
    double xa[3], ya[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        xa[i] = wlam * q0[i];
        ya[i] = wp * q0[i];
    }
    {
        const double c1 = wmu * dot3(g0[0], g0[1], g0[2], g0[0], g0[1], g0[2]);
        double ca[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            ca[i] = xa[i] - ya[i];
        diag[0] = fma(ca[0], q0[0], c1);
        diag[1] = ca[0] * q0[1];
        diag[2] = ca[0] * q0[2];
        diag[3] = fma(ca[1], q0[1], c1);
        diag[4] = ca[1] * q0[2];
        diag[5] = fma(ca[2], q0[2], c1);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const double s1 = wmu * d[j];
        double blk[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double val = fma(xa[i], q[j][k], -(q[j][i] * ya[k]));
                blk[3 * i + k] = i == k ? val + s1 : val;
            }
        emit(j, blk);
    }
}

// Rest and current positions of an incidence's three other vertices and its tet's (lambda, mu).
struct NbrData {
    double X[3][3];
    double u[3][3];
    double lam, mu;
};

__device__ __forceinline__ void load_nbr_data(const AssembleParams& p, const int (&id)[3], int tet, NbrData& d) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const double4 X = reinterpret_cast<const double4*>(p.Xr)[id[j]];
        d.X[j][0] = X.x;
        d.X[j][1] = X.y;
        d.X[j][2] = X.z;
        const double* pu = p.x + 3 * static_cast<size_t>(id[j]);
        d.u[j][0] = pu[0];
        d.u[j][1] = pu[1];
        d.u[j][2] = pu[2];
    }
    const double2 th = reinterpret_cast<const double2*>(p.theta ? p.theta : p.theta_snap)[tet];
    d.lam = th.x;
    d.mu = th.y;
}

// One tile: zero the tile's H segment in LDS (GLOBAL_ACC: in HBM, for a segment above the LDS budget), accumulate,
// write it back coalesced, write the gradient rows and per-vertex energies, OR the status flags.
// LDS layout in doubles: [gradient: 3 per vertex][energy: 1 per vertex][H segment of the tile].
// GLOBAL_ACC tiles add with plain read-modify-writes to HBM, the lanes of a vertex taking turns per block.
template <int LPV, int WAVES, bool WANT_G, int KIND, bool WIDE, bool GLOBAL_ACC, bool PLASTIC>
__device__ __forceinline__ void assemble_tile(const AssembleParams& p, int tile, double* lds) {
    using Inc = typename IncT<WIDE>::type;
    constexpr int kThreads = kWave * WAVES;
    constexpr int kPerWave = kWave / LPV;   // vertex groups never straddle two waves
    constexpr int kVerts = WAVES * kPerWave; // lanes kPerWave * LPV .. 63 of each wave stay idle
    const Inc* __restrict__ inc = static_cast<const Inc*>(p.inc);
    const int vb = p.tile_vbeg[tile], ve = p.tile_vbeg[tile + 1];
    const int nb = p.nbr_ptr[vb];
    // Rows 3v..3v+2 hold 9 |nbr(v)| values, so the tile's segment starts at 9 nbr_ptr[vb].
    const int seg0 = KIND >= 0 ? 9 * nb : 0;
    const int seg1 = KIND >= 0 ? 9 * p.nbr_ptr[ve] : 0;
    double* gl = lds;
    double* el = gl + 3 * kVerts;
    double* acc = GLOBAL_ACC ? p.H + seg0 : el + kVerts;
    if (KIND >= 0)
        for (int i = threadIdx.x; i < seg1 - seg0; i += kThreads)
            acc[i] = 0.0;
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    // Lanes are numbered sub-major (lane = sub * kPerWave + vertex): the LDS serves a 64-bit access 16 lanes at a
    // time, and with the LPV lanes of one vertex in different groups of 16 their adds to the same block (tets that
    // share an edge) no longer collide within a group.  Colliding adds still apply in ascending sub order.
    const int sub = lane / kPerWave;
    const int lv = (static_cast<int>(threadIdx.x) / kWave) * kPerWave + lane % kPerWave;
    const int v = vb + lv;
    const bool owner = lane < kPerWave * LPV && v < ve;
    __syncthreads();
    unsigned flags = 0;
    double gs0 = 0.0, gs1 = 0.0, gs2 = 0.0, es = 0.0;
    double ds[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double* row = acc;
    int rowlen = 0;
    if (owner) {
        const int nv0 = p.nbr_ptr[v];
        rowlen = 3 * (p.nbr_ptr[v + 1] - nv0);
        row = acc + 9 * (nv0 - nb);
        const int32_t* my_nbr = p.nbr + nv0;
        // The vertex's own rest and current position: slot 0 of every tet in its list.
        const double4 Xv = reinterpret_cast<const double4*>(p.Xr)[v];
        const double* xv = p.x + 3 * static_cast<size_t>(v);
        const double uv0 = xv[0], uv1 = xv[1], uv2 = xv[2];
        if (sub == 0 && !(isfinite(uv0) && isfinite(uv1) && isfinite(uv2)))
            flags |= kStatusNonFiniteX;
        const int kb = p.inc_ptr[v], ke = p.inc_ptr[v + 1];
        // Software pipeline over the lane's rounds m (incidence k = kb + sub + m LPV), three stages deep: while
        // round m computes, the vertex data of round m+1, the neighbour ids of round m+2 and the record of round
        // m+3 are in flight.  Rounds alternate between two vertex-data buffers (the loop body holds two rounds),
        // so no register copy waits on a load that is still in flight.  Indices past the lane's last incidence
        // are clamped to it: every load stays in bounds and is branch-free, and a clamped round computes nothing.
        Inc pend_r0{}, pend_r1{};
        int pend_s0 = -1, pend_s1 = -1;
        bool pend_all = false;
        // Rounds (m = (k - kb - sub) / LPV) of listed incidences after the first two, one bit each; a listed round
        // at m >= 64 (a vertex with more than 256 incidences) sets `pend_all` instead.
        unsigned long long pend_more = 0ull;
        const auto clampk = [&](int k) { return k < ke ? k : ke - 1; };
        const auto load_ids = [&](const Inc& r, int (&id)[3]) {
#pragma unroll
            for (int j = 1; j < 4; ++j)
                id[j - 1] = my_nbr[IncT<WIDE>::cp(r, j)];
        };
        const auto load_data = [&](const Inc& r, const int (&id)[3], NbrData& d, int& slot) {
            load_nbr_data(p, id, r.tet, d);
            if (KIND == 1)
                slot = p.proj_slot[r.tet];
        };
        // One round: the incidence's values, the register sums, and the off-diagonal blocks into the segment.
        // Listed incidences (Projected) keep the records and projection slots of up to two per lane for the
        // second pass; later ones are remembered by their round index m (bit m of `pend_more`).
        const VertexData me{{Xv.x, Xv.y, Xv.z}, {uv0, uv1, uv2}};
        const auto round = [&](const Inc& r, const NbrData& d, int slot, bool active, int m) {
            if (!active)
                return;
            int cps[3];
            VertexData nbd[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                cps[j] = IncT<WIDE>::cp(r, j + 1);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    nbd[j].X[c] = d.X[j][c];
                    nbd[j].u[c] = d.u[j][c];
                }
            }
            const double* fp = PLASTIC ? p.fp + 10 * static_cast<size_t>(r.tet) : nullptr;
            // Projected: a listed tet's low-rank term is added by the second pass below; its exact blocks are added
            // here unless the term replaces them (slot <= -2).
            const bool add = KIND == 0 || (KIND == 1 && slot >= -1);
            if (KIND == 1 && slot != -1 && (slot >= 0 ? slot : -2 - slot) < p.proj_capacity) {
                const int li = slot >= 0 ? slot : -2 - slot;
                if (pend_s0 < 0) {
                    pend_s0 = li;
                    pend_r0 = r;
                } else if (pend_s1 < 0) {
                    pend_s1 = li;
                    pend_r1 = r;
                } else if (m < 64) {
                    pend_more |= 1ull << m;
                } else {
                    pend_all = true;
                }
            }
            double gc[3], ec = 0.0, dg[6];
            incidence_row<KIND, PLASTIC>(me, nbd, d.lam, d.mu, fp, IncT<WIDE>::slot(r) == 0, dg, gc, ec, flags,
                                         [&](int j, const double(&blk)[9]) {
                                             if (KIND < 0 || !add)
                                                 return;
                                             double* b = row + 3 * cps[j];
#pragma unroll
                                             for (int s = 0; s < (GLOBAL_ACC ? LPV : 1); ++s) {
                                                 if (!GLOBAL_ACC || sub == s) {
#pragma unroll
                                                     for (int i = 0; i < 3; ++i)
#pragma unroll
                                                         for (int kk = 0; kk < 3; ++kk)
                                                             accum<GLOBAL_ACC>(b + i * rowlen + kk, blk[3 * i + kk]);
                                                 }
                                             }
                                         });
            gs0 += gc[0];
            gs1 += gc[1];
            gs2 += gc[2];
            es += ec;
            if (KIND >= 0 && add)
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    ds[c] += dg[c];
        };
        if (kb < ke) {
            const int k = kb + sub;
            Inc r0 = inc[clampk(k)], r1 = inc[clampk(k + LPV)], r2 = inc[clampk(k + 2 * LPV)];
            int id1[3];
            {
                int id0[3];
                load_ids(r0, id0);
                load_ids(r1, id1);
                NbrData dA, dB;
                int sA = -1, sB = -1;
                load_data(r0, id0, dA, sA);
                for (int k0 = kb;; k0 += 2 * LPV) {
                    const int kc = k0 + sub;
                    int id2[3], id3[3];
                    // Round m (buffer A); stages for rounds m+1 (data), m+2 (ids), m+3 (record).
                    load_data(r1, id1, dB, sB);
                    load_ids(r2, id2);
                    const Inc r3 = inc[clampk(kc + 3 * LPV)];
                    round(r0, dA, sA, kc < ke, (k0 - kb) / LPV);
                    if (k0 + LPV >= ke)
                        break;
                    // Round m+1 (buffer B); stages for rounds m+2, m+3, m+4.
                    load_data(r2, id2, dA, sA);
                    load_ids(r3, id3);
                    const Inc r4 = inc[clampk(kc + 4 * LPV)];
                    round(r1, dB, sB, kc + LPV < ke, (k0 - kb) / LPV + 1);
                    if (k0 + 2 * LPV >= ke)
                        break;
                    r0 = r2;
                    r1 = r3;
                    r2 = r4;
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        id1[j] = id3[j];
                }
            }
        }
        if (KIND == 1) {
            // Second pass: listed tets add their low-rank term sum_v u_v u_v^T (stored by the eigen-clamp kernel):
            // for slot a, block (a, a) (6 distinct entries) and blocks (a, a^j), j = 1..3.  Keeping these loads out
            // of the first pass keeps them from sharing registers with the exact values there.  Entries go two at a
            // time: both entries' counts and first two vectors are loaded in one round (one memory latency for the
            // pair; a count above two loads the remaining vectors afterwards), then entry 0's sums and blocks are
            // added before entry 1's, so the order of additions is the entries' order.
            // The first two listed entries are in registers; later ones are taken from `pend_more` in round order
            // (their records and slots loaded again); `pend_all` rescans every round instead.
            const int rescan_steps = (ke - kb + 2 * LPV - 1) / (2 * LPV);
            for (int q = 0;; ++q) {
                const bool more = pend_all ? q < rescan_steps : (q == 0 || pend_more != 0ull);
                if (!eig9::any_lane(more))
                    break;
                Inc rr[2] = {pend_r0, pend_r1};
                int li[2] = {q == 0 ? pend_s0 : -1, q == 0 ? pend_s1 : -1};
                if (pend_all || q > 0) {
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        int k = ke;
                        if (pend_all) {
                            k = kb + sub + (2 * q + e) * LPV;
                        } else if (pend_more != 0ull) {
                            const int m = __builtin_ctzll(pend_more);
                            pend_more &= pend_more - 1ull;
                            k = kb + sub + m * LPV;
                        }
                        li[e] = -1;
                        if (k < ke) {
                            rr[e] = inc[k];
                            const int slot = p.proj_slot[rr[e].tet];
                            li[e] = slot >= 0 ? slot : (slot == -1 ? -1 : -2 - slot);
                            if (li[e] >= p.proj_capacity)
                                li[e] = -1;
                        }
                    }
                }
                // Element DOF positions of the entry's rows: slot a first, then a^1, a^2, a^3.
                double cnt[2] = {0.0, 0.0};
                double uv[2][2][12];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    if (li[e] >= 0) {
                        const int a = IncT<WIDE>::slot(rr[e]);
                        const double* src = p.proj_H + static_cast<size_t>(li[e]) * kProjPerTet;
                        cnt[e] = src[0];
#pragma unroll
                        for (int vv = 0; vv < 2; ++vv)
#pragma unroll
                            for (int j = 0; j < 4; ++j)
#pragma unroll
                                for (int k = 0; k < 3; ++k)
                                    uv[e][vv][3 * j + k] = src[4 + 12 * vv + 3 * (a ^ j) + k];
                    } else {
#pragma unroll
                        for (int vv = 0; vv < 2; ++vv)
#pragma unroll
                            for (int i = 0; i < 12; ++i)
                                uv[e][vv][i] = 0.0;
                    }
                }
                asm volatile("" ::: "memory");
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const bool add = li[e] >= 0;
                    double pv[6 + 27];
#pragma unroll
                    for (int c = 0; c < 6 + 27; ++c)
                        pv[c] = 0.0;
                    const auto outer = [&](const double (&u)[12]) {
                        pv[0] = fma(u[0], u[0], pv[0]);
                        pv[1] = fma(u[0], u[1], pv[1]);
                        pv[2] = fma(u[0], u[2], pv[2]);
                        pv[3] = fma(u[1], u[1], pv[3]);
                        pv[4] = fma(u[1], u[2], pv[4]);
                        pv[5] = fma(u[2], u[2], pv[5]);
#pragma unroll
                        for (int j = 0; j < 3; ++j)
#pragma unroll
                            for (int i = 0; i < 3; ++i)
#pragma unroll
                                for (int k = 0; k < 3; ++k)
                                    pv[6 + 9 * j + 3 * i + k] = fma(u[i], u[3 + 3 * j + k], pv[6 + 9 * j + 3 * i + k]);
                    };
                    const int n = add ? static_cast<int>(cnt[e]) : 0;
                    // u_0 is always stored (zero when the count is 0); u_1 only when the count is above one.
                    outer(uv[e][0]);
                    if (n > 1)
                        outer(uv[e][1]);
                    if (n > 2) {
                        const int a = IncT<WIDE>::slot(rr[e]);
                        const double* src = p.proj_H + static_cast<size_t>(li[e]) * kProjPerTet;
                        for (int vv = 2; vv < n; ++vv) {
                            double u[12];
#pragma unroll
                            for (int j = 0; j < 4; ++j)
#pragma unroll
                                for (int k = 0; k < 3; ++k)
                                    u[3 * j + k] = src[4 + 12 * vv + 3 * (a ^ j) + k];
                            outer(u);
                        }
                    }
                    int cps[3] = {0, 0, 0};
                    if (add) {
#pragma unroll
                        for (int j = 1; j < 4; ++j)
                            cps[j - 1] = IncT<WIDE>::cp(rr[e], j);
#pragma unroll
                        for (int c = 0; c < 6; ++c)
                            ds[c] += pv[c];
                    }
                    // Block by block, the lanes of a vertex taking turns within a block in HBM accumulation: the
                    // (block, lane) order of the LDS adds.
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        double* blk = row + 3 * cps[j];
#pragma unroll
                        for (int s = 0; s < (GLOBAL_ACC ? LPV : 1); ++s) {
                            if (add && (!GLOBAL_ACC || sub == s)) {
#pragma unroll
                                for (int i = 0; i < 3; ++i)
#pragma unroll
                                    for (int kk = 0; kk < 3; ++kk)
                                        accum<GLOBAL_ACC>(blk + i * rowlen + kk, pv[6 + 9 * j + 3 * i + kk]);
                            }
                        }
                    }
                }
            }
        }
    }
    // Per-vertex diagonal block, gradient and energy: the lanes' partial sums in lane order.
    const int self = owner ? p.self_pos[v] : 0;
#pragma unroll
    for (int s = 0; s < LPV; ++s) {
        if (owner && sub == s) {
            // A vertex no tet uses has an empty row (and zero sums): its `row` is the next row's start.
            if (KIND >= 0 && rowlen > 0) {
                double* blk = row + 3 * self;
                blk[0] += ds[0];
                blk[1] += ds[1];
                blk[2] += ds[2];
                blk[rowlen] += ds[1];
                blk[rowlen + 1] += ds[3];
                blk[rowlen + 2] += ds[4];
                blk[2 * rowlen] += ds[2];
                blk[2 * rowlen + 1] += ds[4];
                blk[2 * rowlen + 2] += ds[5];
            }
            if (s == 0) {
                gl[3 * lv] = gs0;
                gl[3 * lv + 1] = gs1;
                gl[3 * lv + 2] = gs2;
                el[lv] = es;
            } else {
                gl[3 * lv] += gs0;
                gl[3 * lv + 1] += gs1;
                gl[3 * lv + 2] += gs2;
                el[lv] += es;
            }
        }
        __syncthreads();
    }
    if (KIND >= 0) {
        if (!GLOBAL_ACC) {
            // Four LDS reads in flight per lane before their stores: one LDS latency per four values.  The values are
            // stored non-temporally: the Hessian (the bulk of the kernel's traffic) is written once and not read
            // back, and keeping it out of L2 leaves L2 to the vertex data and (lambda, mu) that several incidences
            // share.
            const int n = seg1 - seg0;
            int i = threadIdx.x;
            for (; i + 3 * kThreads < n; i += 4 * kThreads) {
                double val[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    val[q] = acc[i + q * kThreads];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (!isfinite(val[q]))
                        flags |= kStatusNonFiniteOut;
                    __builtin_nontemporal_store(val[q], p.H + seg0 + i + q * kThreads);
                }
            }
            for (; i < n; i += kThreads) {
                const double val = acc[i];
                if (!isfinite(val))
                    flags |= kStatusNonFiniteOut;
                __builtin_nontemporal_store(val, p.H + seg0 + i);
            }
        } else {
            for (int i = threadIdx.x; i < seg1 - seg0; i += kThreads)
                if (!isfinite(acc[i]))
                    flags |= kStatusNonFiniteOut;
        }
    }
    const int nv = ve - vb;
    for (int q = threadIdx.x; q < nv; q += kThreads)
        p.vertex_E[vb + q] = el[q];
    if (WANT_G)
        for (int q = threadIdx.x; q < 3 * nv; q += kThreads) {
            const double val = gl[q];
            if (!isfinite(val))
                flags |= kStatusNonFiniteOut;
            p.g[3 * static_cast<size_t>(vb) + q] = val;
        }
    if (flags)
        atomicOr(p.status, flags);
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).  Remapping block b to the slot
// (b mod 8) q + min(b mod 8, r) + b / 8 (q = G / 8, r = G mod 8) gives every XCD one contiguous run of tiles, so
// the vertex data shared by neighbouring tiles is fetched into one L2 instead of several.
__device__ __forceinline__ int xcd_slot(int b, int G) {
    const int q = G >> 3, r = G & 7, x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

// Waves per SIMD the register allocation must allow: 2 (a 256-VGPR budget).  A tile's LDS segment (about 18 KB
// for 16 vertices of degree 15) already holds a CU to 8 single-wave workgroups, i.e. 2 waves per SIMD, so a smaller
// budget would only force the two-round pipeline's buffers to spill.
constexpr int kAssembleMinWaves = 2;

template <int LPV, int WAVES, bool WANT_G, int KIND, bool WIDE, bool GLOBAL_ACC, bool PLASTIC>
__global__ __launch_bounds__(kWave* WAVES, kAssembleMinWaves) void k_assemble(AssembleParams p) {
    extern __shared__ double lds[];
    const int slot = xcd_slot(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
    const int tile = p.tile_ids ? p.tile_ids[slot] : slot;
    assemble_tile<LPV, WAVES, WANT_G, KIND, WIDE, GLOBAL_ACC, PLASTIC>(p, tile, lds);
}

// Persistent form: a fixed grid walks the tile list with a grid stride (the batched launch).
template <int LPV, int WAVES, bool WANT_G, int KIND, bool WIDE, bool GLOBAL_ACC, bool PLASTIC>
__global__ __launch_bounds__(kWave* WAVES, kAssembleMinWaves) void k_assemble_persistent(AssembleParams p, int num_tiles) {
    extern __shared__ double lds[];
    for (int t = blockIdx.x; t < num_tiles; t += gridDim.x) {
        const int tile = p.tile_ids ? p.tile_ids[t] : t;
        assemble_tile<LPV, WAVES, WANT_G, KIND, WIDE, GLOBAL_ACC, PLASTIC>(p, tile, lds);
        __syncthreads();
    }
}

__global__ void k_clear(unsigned* status, unsigned* counter) {
    if (threadIdx.x == 0) {
        status[0] = 0u;
        if (counter)
            counter[0] = 0u;
    }
}

// Energy reduction, independent of the tiling: level 1 sums fixed chunks of at most 1024 per-vertex values
// (chunk boundaries are relative to the first vertex of each mesh), level 2 sums each mesh's chunk partials.  Both
// levels use a fixed strided order per lane and a fixed tree, so the energy of a mesh has the same bits whatever the
// launch shape, and whether the mesh is evaluated alone or inside a batch.  The tree: an xor butterfly within each
// wave (all lanes end with the same bits), then the N / 64 wave totals in wave order (one barrier).
template <int N> __device__ __forceinline__ double tree_sum(double* red, double s) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
        s += __shfl_xor(s, off);
    if (threadIdx.x % kWave == 0)
        red[threadIdx.x / kWave] = s;
    __syncthreads();
    double t = red[0];
#pragma unroll
    for (int w = 1; w < N / kWave; ++w)
        t += red[w];
    return t;
}

__global__ __launch_bounds__(256) void k_energy_chunks(const double* __restrict__ values,
                                                       const int32_t* __restrict__ chunk_begin,
                                                       double* __restrict__ chunk_sum) {
    __shared__ double red[256 / kWave];
    const int c = blockIdx.x;
    const int b = chunk_begin[c], e = chunk_begin[c + 1];
    double s = 0.0;
    for (int i = b + static_cast<int>(threadIdx.x); i < e; i += 256)
        s += values[i];
    const double total = tree_sum<256>(red, s);
    if (threadIdx.x == 0)
        chunk_sum[c] = total;
}

// Hands the evaluation's energy, status word and list counter to host-visible pinned memory (readback) from the
// evaluation's last kernel, in place of a device-to-host copy, and zeroes the status word and the counter for the
// next evaluation (the host clears them itself only after an operation that leaves them set).  The status word goes
// last, after a system-scope fence: a host that sees it change (the host sets it to all ones before a replay) also
// sees the energy and the counter.
__device__ __forceinline__ void write_readback(double energy, unsigned st, unsigned* status, const unsigned* counter,
                                               void* readback) {
    double* rb = static_cast<double*>(readback);
    unsigned* rbu = static_cast<unsigned*>(readback);
    rb[0] = energy;
    rbu[3] = counter ? __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    __hip_atomic_store(status, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (counter)
        __hip_atomic_store(const_cast<unsigned*>(counter), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence_system();
    __hip_atomic_store(&rbu[2], st, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// readback (one segment only): the read-back above from this kernel, the evaluation's last.
__global__ __launch_bounds__(256) void k_energy_segments(const double* __restrict__ chunk_sum,
                                                         const int32_t* __restrict__ seg_chunk,
                                                         double* __restrict__ out, unsigned* status,
                                                         const unsigned* counter, void* readback) {
    __shared__ double red[256 / kWave];
    const int m = blockIdx.x;
    double s = 0.0;
    for (int i = seg_chunk[m] + static_cast<int>(threadIdx.x); i < seg_chunk[m + 1]; i += 256)
        s += chunk_sum[i];
    const double total = tree_sum<256>(red, s);
    if (threadIdx.x == 0) {
        out[m] = total;
        unsigned st = __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!isfinite(total)) {
            atomicOr(status, kStatusNonFiniteOut);
            st |= kStatusNonFiniteOut;
        }
        if (readback)
            write_readback(total, st, status, counter, readback);
    }
}

// The read-back as a kernel of its own (an evaluation of several segments): the energy slot, status and counter
// as the earlier kernels left them.
__global__ void k_readback(const double* energy, unsigned* status, const unsigned* counter, void* readback) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        write_readback(*energy, __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), status,
                       counter, readback);
}

// The total over segments (a batch evaluated through the single-energy entry points): one workgroup, fixed order.
__global__ __launch_bounds__(256) void k_energy_total(const double* __restrict__ seg, int n, double* __restrict__ out) {
    __shared__ double red[256 / kWave];
    double s = 0.0;
    for (int i = static_cast<int>(threadIdx.x); i < n; i += 256)
        s += seg[i];
    const double total = tree_sum<256>(red, s);
    if (threadIdx.x == 0)
        out[0] = total;
}

// ---------------------------------------------------------------------------------------------------------------
// Projection: prepass (closed-form definiteness certificate of the 9x9 material tangent A) and the 9x9 eigen-clamp
// of M.
// ---------------------------------------------------------------------------------------------------------------
struct ProjectParams {
    const double* Xr;
    const int32_t* conn;
    const double* theta_snap;
    const double* x;
    const double* theta;
    int num_tets;
    unsigned* status;
    unsigned* counter;
    int32_t* proj_slot;
    int32_t* proj_list;
    double* proj_H;
    const double* fp = nullptr; // plastic parts (10 per tet) or null
    int proj_capacity = 0;      // list entries proj_H holds
};

__device__ __forceinline__ int sym_index(int r, int c) {
    const int lo = r < c ? r : c, hi = r < c ? c : r;
    return lo * 9 - lo * (lo - 1) / 2 + (hi - lo);
}

constexpr int kPrepassThreads = 1024;

__global__ __launch_bounds__(kPrepassThreads) void k_project_prepass(ProjectParams p) {
    __shared__ unsigned wave_count[kPrepassThreads / kWave];
    __shared__ unsigned block_base;
    const int e = blockIdx.x * kPrepassThreads + threadIdx.x;
    bool indefinite = false;
    if (e < p.num_tets) {
        int ids[4];
        TetIn t;
        load_tet(p.Xr, p.conn, p.theta_snap, p.x, p.theta, e, ids, t, p.fp);
        Geo geo;
        rest_geometry(t, geo);
        Kin k;
        kinematics(t, geo, k);
        // Elements whose tangent is not certified positive definite (indefinite, or too close to singular to tell)
        // get the eigen-clamp; for a positive semi-definite M it returns M itself.
        if (kin_ok(k) && theta_ok(t.lam, t.mu))
            indefinite = !eig9::tangent_certified_pd(k.F, k.J, t.lam, t.mu, fma(t.lam, k.L, -t.mu));
    }
    // One global atomic per workgroup: the wave counts are combined in LDS, and the indefinite lanes take
    // consecutive slots in (wave, lane) order after the workgroup's base.  A single counter hit once per wave
    // serializes ~T/64 atomics at one L2 address; per workgroup it is 16x fewer.
    const unsigned long long mask = __ballot(indefinite);
    const int lane = static_cast<int>(__lane_id());
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    if (lane == 0)
        wave_count[wave] = static_cast<unsigned>(__popcll(mask));
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned total = 0;
        for (int w = 0; w < kPrepassThreads / kWave; ++w)
            total += wave_count[w];
        block_base = total ? atomicAdd(p.counter, total) : 0u;
    }
    __syncthreads();
    int32_t slot = -1;
    if (indefinite) {
        unsigned base = block_base;
        for (int w = 0; w < wave; ++w)
            base += wave_count[w];
        slot = static_cast<int32_t>(base + __popcll(mask & ((1ull << lane) - 1ull)));
        p.proj_list[slot] = e;
    }
    if (e < p.num_tets)
        p.proj_slot[e] = slot;
}

// The reduced element Hessian M = w (C (x) I3) A (C (x) I3)^T of tet e, packed upper: with c_m = sum_a P(a,m) g_a
// and r_m = F^-T c_m, M[(3m+i), (3n+k)] = w [mu (c_m . c_n) d_ik + lambda r_m,i r_n,k - p r_n,i r_m,k].
// ADD: M += the reduced Hessian (one rounding per entry) instead of M = it.
template <bool ADD = false>
__device__ __forceinline__ void reduced_hessian(const ProjectParams& p, int e, double (&M)[45]) {
    int ids[4];
    TetIn t;
    load_tet(p.Xr, p.conn, p.theta_snap, p.x, p.theta, e, ids, t, p.fp);
    Geo geo;
    rest_geometry(t, geo);
    Kin k;
    kinematics(t, geo, k);
    double G[4][3];
    shape_gradients(geo.Di, G);
    const double pr = fma(t.lam, k.L, -t.mu);
    double cm[3][3], rm[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            cm[m][j] = fma(kP[3][m], G[3][j], fma(kP[2][m], G[2][j], fma(kP[1][m], G[1][j], kP[0][m] * G[0][j])));
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            rm[m][i] = dot3(k.a[3 * i], k.a[3 * i + 1], k.a[3 * i + 2], cm[m][0], cm[m][1], cm[m][2]);
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int c = r; c < 9; ++c) {
            const int m = r / 3, i = r % 3, n = c / 3, kk = c % 3;
            double val = fma(t.lam * rm[m][i], rm[n][kk], -(pr * rm[n][i] * rm[m][kk]));
            if (i == kk)
                val = fma(t.mu, dot3(cm[m][0], cm[m][1], cm[m][2], cm[n][0], cm[n][1], cm[n][2]), val);
            if (ADD)
                M[sym_index(r, c)] = fma(geo.w, val, M[sym_index(r, c)]);
            else
                M[sym_index(r, c)] = geo.w * val;
        }
}

// Per-lane spill store of the eigen-clamp (eig9::kSpill doubles: the Householder reflectors and the tridiagonal
// matrix), lane-interleaved in LDS so that the 64 lanes of a wave touch 64 consecutive doubles per access.
struct LdsSpill {
    double* base; // &lds[lane]
    __device__ void put(int i, double v) { base[i * 64] = v; }
    __device__ double get(int i) const { return base[i * 64]; }
};

// One thread per listed element.  A single pass (no grid-stride loop) keeps the compiler from hoisting the
// element-independent constants of the eigen-analysis out of a loop and spilling them to scratch memory.
__global__ __launch_bounds__(64, 2) void k_project_jacobi(ProjectParams p) {
    __shared__ double spill_lds[eig9::kSpill * 64];
    const unsigned count = *p.counter;
    {
        const unsigned idx = blockIdx.x * 64 + threadIdx.x;
        if (idx >= count)
            return;
        if (idx >= static_cast<unsigned>(p.proj_capacity)) {
            // The store is too small for this evaluation's list: the host grows it and evaluates again.
            if (idx == static_cast<unsigned>(p.proj_capacity))
                atomicOr(p.status, kStatusProjOverflow);
            return;
        }
        const int e = p.proj_list[idx];
        double M[45];
        reduced_hessian(p, e, M);
        double scale = 0.0;
#pragma unroll
        for (int q = 0; q < 45; ++q)
            scale = fmax(scale, fabs(M[q]));
        // proj(M) = (add_M ? M : 0) + sum_v wts_v x_v x_v^T from the eigenvalues and the smaller eigen-set's vectors
        // (M is destroyed by the reduction).  With Q = P (x) I3, proj(H_e) = Q proj(M) Q^T = (add_M ? H_e : 0)
        // + sum_v u_v u_v^T, u_v = sqrt(wts_v) Q x_v: the assembly adds H_e itself, so only the u_v are stored.
        double X[4][9], wts[4];
        int ns = 0;
        bool add_M = false;
        LdsSpill spill{spill_lds + threadIdx.x};
        const bool converged = eig9::psd_part(
            M, X, wts, ns, add_M, scale, [](double x) { return rsqrt_nr(x); }, [](double x) { return recip(x); },
            spill);
        double* out = p.proj_H + kProjPerTet * static_cast<size_t>(idx);
        out[0] = static_cast<double>(ns);
        bool fin = true;
#pragma unroll
        for (int v = 0; v < kProjVectors; ++v) {
            if (!eig9::any_lane(v < ns || v == 0))
                break;
            const double sw = sqrt(wts[v]);
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    double u = 0.0;
#pragma unroll
                    for (int m = 0; m < 3; ++m)
                        if (kP[a][m] != 0.0)
                            u = fma(kP[a][m], X[v][3 * m + i], u);
                    u *= sw;
                    fin = fin && isfinite(u);
                    if (v < ns || v == 0)
                        out[4 + 12 * v + 3 * a + i] = u;
                }
        }
        if (!add_M)
            p.proj_slot[e] = -2 - static_cast<int32_t>(idx);
        if (!converged || !fin)
            atomicOr(p.status, kStatusProjection);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Differentiation kernels.
// ---------------------------------------------------------------------------------------------------------------
struct ElemParams {
    const double* Xr;
    const int32_t* conn;
    const double* theta_snap;
    const double* x;
    const double* theta;
    const double* w;      // force cotangent (num_dofs)
    const double* scale;  // device scalar multiplying dE/dtheta, or null for scale_val
    double scale_val;
    double* out;          // T x 2
    unsigned* status;
    int num_tets;
    const double* fp = nullptr; // plastic parts (10 per tet) or null
};

// Per-tet theta cotangent: out_e = s dE/dtheta_e + (d g_e / d theta_e)^T w, with
//   dE/dtheta_e = w_e (L^2 / 2, (|F|^2 - 3) / 2 - L)
//   d g_e[3a+i] / d lambda = w_e L q_a,i,  d / d mu = w_e ((F g_a)_i - q_a,i)
// (psi and P = mu (F - F^-T) + lambda L F^-T of neo_hookean.cpp:88-101, P linear in theta).
// HAS_S drops the second term's absence; HAS_W the first's: <true, false> is dE/dtheta (s = 1), <false, true>
// the force VJP, <true, true> both in one pass.
template <bool HAS_S, bool HAS_W> __global__ __launch_bounds__(256) void k_theta_vjp(ElemParams p) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= p.num_tets)
        return;
    int ids[4];
    TetIn t;
    load_tet(p.Xr, p.conn, p.theta_snap, p.x, p.theta, e, ids, t, p.fp);
    Geo geo;
    rest_geometry(t, geo);
    Kin k;
    kinematics(t, geo, k);
    unsigned flags = 0;
    if (!kin_ok(k))
        flags |= kStatusDetF;
    if (!theta_ok(t.lam, t.mu))
        flags |= kStatusBadTheta;
    double2 r = make_double2(0.0, 0.0);
    if (HAS_S) {
        const double s = p.scale ? *p.scale : p.scale_val;
        double sq = 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q)
            sq = fma(k.F[q], k.F[q], sq);
        r.x = s * (geo.w * (0.5 * k.L * k.L));
        r.y = s * (geo.w * (0.5 * (sq - 3.0) - k.L));
    }
    if (HAS_W) {
        double G[4][3];
        shape_gradients(geo.Di, G);
        double sl = 0.0, sm = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const double* wb = p.w + 3 * static_cast<size_t>(ids[b]);
            const double w0 = wb[0], w1 = wb[1], w2 = wb[2];
            double q[3], Fg[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                q[i] = dot3(k.a[3 * i], k.a[3 * i + 1], k.a[3 * i + 2], G[b][0], G[b][1], G[b][2]);
                Fg[i] = dot3(k.F[3 * i], k.F[3 * i + 1], k.F[3 * i + 2], G[b][0], G[b][1], G[b][2]);
            }
            const double qw = dot3(q[0], q[1], q[2], w0, w1, w2);
            const double fw = dot3(Fg[0], Fg[1], Fg[2], w0, w1, w2);
            sl += qw;
            sm += fw - qw;
        }
        r.x += geo.w * k.L * sl;
        r.y += geo.w * sm;
    }
    if (!(isfinite(r.x) && isfinite(r.y)))
        flags |= kStatusNonFiniteOut;
    reinterpret_cast<double2*>(p.out)[e] = r;
    if (flags)
        atomicOr(p.status, flags);
}

// Force response to the plastic part (VolumetricFemForceFpResponse = detail::VolumetricFemForceResponse with
// ForceInput::Plastic, volumetric_fem_force_response.hpp:81-179; NeoHookean::eval_d2W_dF_dFp, neo_hookean.cpp:103-128):
// f(Fp) = dE/dx at fixed x and theta as a function of the per-tet plastic parts (column-major vec(Fp), 9 per tet).
// With Fe = F Fp^-1, psi_Fe = mu Fe + p Fe^-T (p = lambda log det Fe - mu), w' = w det Fp and g'_b = Fp^-T g_b (g_b
// the rest shape gradients), the element force is f_b = w' psi_Fe g'_b, and for a direction dFp
//   df_b = w' D g'_b,  D = tr(Fp^-1 dFp) psi_Fe + A:dFe - psi_Fe Fp^-T dFp^T,  dFe = -Fe dFp Fp^-1,
// A:X = mu X + lambda (Fe^-T : X) Fe^-T - p Fe^-T X^T Fe^-T (the tangent of neo_hookean.cpp:43-60).  The transpose
// product for a force cotangent w (V' = sum_b w_b g'_b^T, c = psi_Fe : V') is
//   (df/dFp)^T w = w' [c I - Fe^T (A:V') - V'^T psi_Fe] Fp^-T.
struct FpParams {
    const double* Xr;
    const int32_t* conn;
    const double* theta_snap;
    const double* x;
    const double* theta;
    const double* fp_snap; // the handle's Fp^-1 (row-major) and det Fp per tet, or null (identity)
    const double* fp;      // per-call Fp (column-major vec, 9 per tet) or null
    const double* w;       // VJP: force cotangent (3 per vertex)
    const double* dfp;     // element mode: direction (9 per tet), or null for the force itself
    const int32_t* tet_inc; // element mode: incidence index of each (tet, slot), 4 per tet
    double* out;           // VJP: 9 per tet; element mode: 3 per incidence
    unsigned* status;
    int num_tets;
};

// Per-tet state of the Fp response: rest gradients g_b (rows of G), Fp^-1 (row-major), w' = w det Fp, Fe, Fe^-T,
// p and psi_Fe (row-major).
struct FpState {
    double G[4][3];
    double Fi[9];
    double wp, p, mu, lam;
    double Fe[9], a[9], psi[9];
};

__device__ __forceinline__ void fp_state(const FpParams& p, int e, int ids[4], FpState& s, unsigned& flags) {
    TetIn t;
    load_tet(p.Xr, p.conn, p.theta_snap, p.x, p.theta, e, ids, t, nullptr);
    Geo geo;
    rest_geometry(t, geo); // rest geometry without the plastic part: Di = Dm^-1, w = det Dm / 6
    shape_gradients(geo.Di, s.G);
    double detp = 1.0;
    if (p.fp) {
        double Fp[9]; // row-major from the column-major input
        bool fp_finite = true;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                Fp[3 * i + j] = p.fp[9 * static_cast<size_t>(e) + i + 3 * j];
                fp_finite = fp_finite && isfinite(Fp[3 * i + j]);
            }
        double C[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                C[3 * i + j] = fma(Fp[3 * i1 + j1], Fp[3 * i2 + j2], -(Fp[3 * i1 + j2] * Fp[3 * i2 + j1]));
            }
        detp = dot3(Fp[0], Fp[1], Fp[2], C[0], C[1], C[2]);
        // MultiplicativeKinematics (multiplicative_kinematics.cpp:9-16): finite Fp, then finite det Fp > 0.
        if (!fp_finite)
            flags |= kStatusNonFiniteFp;
        else if (!(detp > 0.0) || !isfinite(detp))
            flags |= kStatusBadDetFp;
        const double inv = 1.0 / detp;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                s.Fi[3 * i + j] = C[3 * j + i] * inv;
    } else if (p.fp_snap) {
#pragma unroll
        for (int q = 0; q < 9; ++q)
            s.Fi[q] = p.fp_snap[10 * static_cast<size_t>(e) + q];
        detp = p.fp_snap[10 * static_cast<size_t>(e) + 9];
    } else {
#pragma unroll
        for (int q = 0; q < 9; ++q)
            s.Fi[q] = (q % 4 == 0) ? 1.0 : 0.0;
    }
    s.wp = geo.w * detp;
    s.mu = t.mu;
    s.lam = t.lam;
    if (!theta_ok(t.lam, t.mu))
        flags |= kStatusBadTheta;
    // F = I + Du Dm^-1 (total), Fe = F Fp^-1.
    double F[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double v = (i == j) ? 1.0 : 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                v = fma(t.u[c + 1][i] - t.u[0][i], geo.Di[3 * c + j], v);
            F[3 * i + j] = v;
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            s.Fe[3 * i + j] = dot3(F[3 * i], F[3 * i + 1], F[3 * i + 2], s.Fi[j], s.Fi[3 + j], s.Fi[6 + j]);
    double C[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            C[3 * i + j] = fma(s.Fe[3 * i1 + j1], s.Fe[3 * i2 + j2], -(s.Fe[3 * i1 + j2] * s.Fe[3 * i2 + j1]));
        }
    const double J = dot3(s.Fe[0], s.Fe[1], s.Fe[2], C[0], C[1], C[2]);
    const double L = log(J);
    if (!(J > 0.0 && isfinite(J) && isfinite(L)))
        flags |= kStatusDetF;
    const double inv = 1.0 / J;
    s.p = fma(t.lam, L, -t.mu);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        s.a[q] = C[q] * inv;
        s.psi[q] = fma(t.mu, s.Fe[q], s.p * s.a[q]);
    }
}

// A:X for the neo-Hookean tangent at Fe (row-major 3x3 arguments).
__device__ __forceinline__ void tangent_apply(const FpState& s, const double X[9], double Y[9]) {
    double ax = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q)
        ax = fma(s.a[q], X[q], ax);
    double XTa[9]; // X^T a
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            XTa[3 * i + j] = dot3(X[i], X[3 + i], X[6 + i], s.a[j], s.a[3 + j], s.a[6 + j]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double aXTa = dot3(s.a[3 * i], s.a[3 * i + 1], s.a[3 * i + 2], XTa[j], XTa[3 + j], XTa[6 + j]);
            Y[3 * i + j] = fma(s.mu, X[3 * i + j], fma(s.lam * ax, s.a[3 * i + j], -(s.p * aXTa)));
        }
}

// g'_b = Fp^-T g_b.
__device__ __forceinline__ void plastic_gradient(const FpState& s, int b, double gp[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
        gp[i] = dot3(s.Fi[i], s.Fi[3 + i], s.Fi[6 + i], s.G[b][0], s.G[b][1], s.G[b][2]);
}

__global__ __launch_bounds__(256) void k_fp_vjp(FpParams p) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= p.num_tets)
        return;
    unsigned flags = 0;
    int ids[4];
    FpState s;
    fp_state(p, e, ids, s, flags);
    double V[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}; // sum_b w_b g'_b^T
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        double gp[3];
        plastic_gradient(s, b, gp);
        const double* wb = p.w + 3 * static_cast<size_t>(ids[b]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                V[3 * i + j] = fma(wb[i], gp[j], V[3 * i + j]);
    }
    double c = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q)
        c = fma(s.psi[q], V[q], c);
    double Y[9];
    tangent_apply(s, V, Y);
    double Mt[9]; // c I - Fe^T Y - V^T psi
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double fy = dot3(s.Fe[i], s.Fe[3 + i], s.Fe[6 + i], Y[j], Y[3 + j], Y[6 + j]);
            const double vp = dot3(V[i], V[3 + i], V[6 + i], s.psi[j], s.psi[3 + j], s.psi[6 + j]);
            Mt[3 * i + j] = ((i == j) ? c : 0.0) - fy - vp;
        }
    double* out = p.out + 9 * static_cast<size_t>(e);
    bool fin = true;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            // (M Fp^-T)(i, j) = sum_k M(i, k) Fp^-1(j, k); column-major output.
            const double v = s.wp * dot3(Mt[3 * i], Mt[3 * i + 1], Mt[3 * i + 2], s.Fi[3 * j], s.Fi[3 * j + 1],
                                         s.Fi[3 * j + 2]);
            fin = fin && isfinite(v);
            out[i + 3 * j] = v;
        }
    if (!fin)
        flags |= kStatusNonFiniteOut;
    if (flags)
        atomicOr(p.status, flags);
}

// Element contributions: the force f_b itself (dfp null) or df_b for dfp, the 3-vector of slot b stored at the
// incidence of (tet, b), so that each vertex's contributions are contiguous.
__global__ __launch_bounds__(256) void k_fp_elem(FpParams p) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= p.num_tets)
        return;
    unsigned flags = 0;
    int ids[4];
    FpState s;
    fp_state(p, e, ids, s, flags);
    double D[9];
    if (!p.dfp) {
#pragma unroll
        for (int q = 0; q < 9; ++q)
            D[q] = s.psi[q];
    } else {
        double dF[9]; // row-major from column-major
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                dF[3 * i + j] = p.dfp[9 * static_cast<size_t>(e) + i + 3 * j];
        double tr = 0.0; // tr(Fp^-1 dFp)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            tr += dot3(s.Fi[3 * i], s.Fi[3 * i + 1], s.Fi[3 * i + 2], dF[i], dF[3 + i], dF[6 + i]);
        double T1[9], dFe[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                T1[3 * i + j] = dot3(s.Fe[3 * i], s.Fe[3 * i + 1], s.Fe[3 * i + 2], dF[j], dF[3 + j], dF[6 + j]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                dFe[3 * i + j] = -dot3(T1[3 * i], T1[3 * i + 1], T1[3 * i + 2], s.Fi[j], s.Fi[3 + j], s.Fi[6 + j]);
        double AdFe[9];
        tangent_apply(s, dFe, AdFe);
        double PF[9]; // psi Fp^-T
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                PF[3 * i + j] = dot3(s.psi[3 * i], s.psi[3 * i + 1], s.psi[3 * i + 2], s.Fi[3 * j], s.Fi[3 * j + 1],
                                     s.Fi[3 * j + 2]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                // (psi Fp^-T dFp^T)(i, j) = sum_k PF(i, k) dFp(j, k)
                const double pfd = dot3(PF[3 * i], PF[3 * i + 1], PF[3 * i + 2], dF[3 * j], dF[3 * j + 1], dF[3 * j + 2]);
                D[3 * i + j] = fma(tr, s.psi[3 * i + j], AdFe[3 * i + j] - pfd);
            }
    }
    const int4 kinc = reinterpret_cast<const int4*>(p.tet_inc)[e];
    bool fin = true;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        double gp[3];
        plastic_gradient(s, b, gp);
        const int k = b == 0 ? kinc.x : b == 1 ? kinc.y : b == 2 ? kinc.z : kinc.w;
        double* out = p.out + 3 * static_cast<size_t>(k);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double v = s.wp * dot3(D[3 * i], D[3 * i + 1], D[3 * i + 2], gp[0], gp[1], gp[2]);
            fin = fin && isfinite(v);
            out[i] = v;
        }
    }
    if (!fin)
        flags |= kStatusNonFiniteOut;
    if (flags)
        atomicOr(p.status, flags);
}

// out_v = sum of the 3-vectors of v's incidences (contiguous, 3 doubles each).  Eight lanes share a vertex and
// stride over its range; their partial sums are combined by a fixed xor tree, so the result does not depend on
// timing.
constexpr int kElemGatherLanes = 8;

__global__ __launch_bounds__(256) void k_elem_gather(const int32_t* __restrict__ inc_ptr,
                                                     const double* __restrict__ elem, double* __restrict__ out,
                                                     int num_vertices) {
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int v = gid / kElemGatherLanes;
    const int sub = gid % kElemGatherLanes;
    // A group is either wholly past the end or wholly active, so the shuffles below read active lanes only.
    if (v >= num_vertices)
        return;
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    const int ke = inc_ptr[v + 1];
    for (int k = inc_ptr[v] + sub; k < ke; k += kElemGatherLanes) {
        const double* src = elem + 3 * static_cast<size_t>(k);
        y0 += src[0];
        y1 += src[1];
        y2 += src[2];
    }
#pragma unroll
    for (int m = kElemGatherLanes / 2; m > 0; m >>= 1) {
        y0 += __shfl_xor(y0, m, kElemGatherLanes);
        y1 += __shfl_xor(y1, m, kElemGatherLanes);
        y2 += __shfl_xor(y2, m, kElemGatherLanes);
    }
    if (sub == 0) {
        out[3 * static_cast<size_t>(v)] = y0;
        out[3 * static_cast<size_t>(v) + 1] = y1;
        out[3 * static_cast<size_t>(v) + 2] = y2;
    }
}

// Vertex-gather products over each vertex's incidence list:
//   MODE 0: (H v)_v                                (Exact Hessian-vector product)
//   MODE 1: sum_e (d g_e / d theta_e) dtheta_e     (force JVP in theta; v = dtheta, T x 2)
//   MODE 2: s f_v + (H v)_v                        (s = *scale or 0, f = dE/dx; v may be null)
// kGatherLanes lanes share a vertex and stride over its incidences; their partial sums are combined by a fixed
// xor butterfly, so every lane of the group ends with the same, launch-invariant value.
struct GatherParams {
    const double* Xr;
    const int32_t* conn;
    const double* theta_snap;
    const int32_t* inc_ptr;
    const int32_t* inc_tet; // stride (bytes) given by inc_stride
    int inc_stride;
    const double* x;
    const double* theta;
    const double* v;
    const double* scale;
    double* out;     // num_dofs
    unsigned* status;
    int num_vertices;
    const double* fp = nullptr; // plastic parts (10 per tet) or null
};

constexpr int kGatherLanes = 8;
constexpr int kGatherThreads = 256;

template <int MODE> __global__ __launch_bounds__(kGatherThreads) void k_vertex_gather(GatherParams p) {
    static_assert(kGatherThreads % kGatherLanes == 0 && kWave % kGatherLanes == 0, "groups stay inside a wave");
    const int gid = blockIdx.x * kGatherThreads + threadIdx.x;
    const int v = gid / kGatherLanes;
    const int sub = gid % kGatherLanes;
    // A group is either wholly past the end or wholly active, so the shuffles below read active lanes only.
    if (v >= p.num_vertices)
        return;
    unsigned flags = 0;
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    const bool has_v = MODE != 2 || p.v != nullptr;
    const double s = (MODE == 2 && p.scale) ? *p.scale : 0.0;
    const int kb = p.inc_ptr[v], ke = p.inc_ptr[v + 1];
    for (int kk = kb + sub; kk < ke; kk += kGatherLanes) {
        const int e = *reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(p.inc_tet) +
                                                         static_cast<size_t>(kk) * p.inc_stride);
        int ids[4];
        TetIn t;
        load_tet(p.Xr, p.conn, p.theta_snap, p.x, p.theta, e, ids, t, p.fp);
        Geo geo;
        rest_geometry(t, geo);
        const int a = (ids[0] == v) ? 0 : (ids[1] == v) ? 1 : (ids[2] == v) ? 2 : 3;
        Kin k;
        kinematics(t, geo, k);
        if (!kin_ok(k))
            flags |= kStatusDetF;
        if (!theta_ok(t.lam, t.mu))
            flags |= kStatusBadTheta;
        double G[4][3];
        shape_gradients(geo.Di, G);
        double ga[3];
        select_row(G, a, ga);
        double qa[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            qa[i] = dot3(k.a[3 * i], k.a[3 * i + 1], k.a[3 * i + 2], ga[0], ga[1], ga[2]);
        if (MODE == 1 || MODE == 2) {
            double Fg[3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                Fg[i] = dot3(k.F[3 * i], k.F[3 * i + 1], k.F[3 * i + 2], ga[0], ga[1], ga[2]);
            if (MODE == 1) {
                const double2 dt = reinterpret_cast<const double2*>(p.v)[e];
                const double dl = dt.x * k.L;
                y0 = fma(geo.w, fma(dl, qa[0], dt.y * (Fg[0] - qa[0])), y0);
                y1 = fma(geo.w, fma(dl, qa[1], dt.y * (Fg[1] - qa[1])), y1);
                y2 = fma(geo.w, fma(dl, qa[2], dt.y * (Fg[2] - qa[2])), y2);
            } else {
                // f_a = w (mu F g_a + (lambda L - mu) q_a)
                const double pr = fma(t.lam, k.L, -t.mu);
                const double sw = s * geo.w;
                y0 = fma(sw, fma(t.mu, Fg[0], pr * qa[0]), y0);
                y1 = fma(sw, fma(t.mu, Fg[1], pr * qa[1]), y1);
                y2 = fma(sw, fma(t.mu, Fg[2], pr * qa[2]), y2);
            }
        }
        if ((MODE == 0 || MODE == 2) && has_v) {
            const double pr = fma(t.lam, k.L, -t.mu);
            // sum_b H_ab v_b = w sum_b [mu (g_a.g_b) v_b + lambda q_a (q_b.v_b) - p q_b (q_a.v_b)]
            double sv[3] = {0.0, 0.0, 0.0}, sq[3] = {0.0, 0.0, 0.0};
            double sqv = 0.0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double* vb = p.v + 3 * static_cast<size_t>(ids[b]);
                const double vb0 = vb[0], vb1 = vb[1], vb2 = vb[2];
                double qb[3];
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    qb[i] = dot3(k.a[3 * i], k.a[3 * i + 1], k.a[3 * i + 2], G[b][0], G[b][1], G[b][2]);
                const double gab = dot3(ga[0], ga[1], ga[2], G[b][0], G[b][1], G[b][2]);
                sv[0] = fma(gab, vb0, sv[0]);
                sv[1] = fma(gab, vb1, sv[1]);
                sv[2] = fma(gab, vb2, sv[2]);
                sqv += dot3(qb[0], qb[1], qb[2], vb0, vb1, vb2);
                const double qav = dot3(qa[0], qa[1], qa[2], vb0, vb1, vb2);
                sq[0] = fma(qb[0], qav, sq[0]);
                sq[1] = fma(qb[1], qav, sq[1]);
                sq[2] = fma(qb[2], qav, sq[2]);
            }
            y0 = fma(geo.w, fma(t.mu, sv[0], fma(t.lam * qa[0], sqv, -(pr * sq[0]))), y0);
            y1 = fma(geo.w, fma(t.mu, sv[1], fma(t.lam * qa[1], sqv, -(pr * sq[1]))), y1);
            y2 = fma(geo.w, fma(t.mu, sv[2], fma(t.lam * qa[2], sqv, -(pr * sq[2]))), y2);
        }
    }
#pragma unroll
    for (int m = kGatherLanes / 2; m > 0; m >>= 1) {
        y0 += __shfl_xor(y0, m, kGatherLanes);
        y1 += __shfl_xor(y1, m, kGatherLanes);
        y2 += __shfl_xor(y2, m, kGatherLanes);
    }
    if (sub == 0) {
        double* o = p.out + 3 * static_cast<size_t>(v);
        o[0] = y0;
        o[1] = y1;
        o[2] = y2;
        if (!(isfinite(y0) && isfinite(y1) && isfinite(y2)))
            flags |= kStatusNonFiniteOut;
    }
    if (flags)
        atomicOr(p.status, flags);
}

// Lame parameters from Young's modulus and Poisson ratio, per element:
//   lambda = E nu / ((1 + nu)(1 - 2 nu)),  mu = E / (2 (1 + nu)),
// and the reverse product (g_E, g_nu) = (g_lambda, g_mu) d(lambda, mu)/d(E, nu) with
//   dlambda/dE = nu / D, dlambda/dnu = E (1 + 2 nu^2) / D^2, dmu/dE = 1 / (2 (1 + nu)), dmu/dnu = -E / (2 (1 + nu)^2),
// D = (1 + nu)(1 - 2 nu).
__global__ __launch_bounds__(256) void k_lame(const double* __restrict__ E, const double* __restrict__ nu, long long n,
                                              double2* __restrict__ theta) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += 256ll * gridDim.x) {
        const double e = E[i], v = nu[i];
        const double op = 1.0 + v;
        theta[i] = make_double2(e * v / (op * (1.0 - 2.0 * v)), e / (2.0 * op));
    }
}

__global__ __launch_bounds__(256) void k_lame_vjp(const double* __restrict__ E, const double* __restrict__ nu,
                                                  const double2* __restrict__ gtheta, long long n,
                                                  double* __restrict__ gE, double* __restrict__ gnu) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += 256ll * gridDim.x) {
        const double e = E[i], v = nu[i];
        const double2 g = gtheta[i];
        const double op = 1.0 + v;
        const double D = op * (1.0 - 2.0 * v);
        const double inv_D = 1.0 / D, inv_op = 1.0 / op;
        if (gE)
            gE[i] = g.x * (v * inv_D) + g.y * (0.5 * inv_op);
        if (gnu)
            gnu[i] = g.x * (e * fma(2.0 * v, v, 1.0) * inv_D * inv_D) - g.y * (0.5 * e * inv_op * inv_op);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Feasible step: the largest alpha in [0, 1] (times the reference's 0.9 safety factor) such that det F(x + t dx) > 0
// for every tet and every t in [0, alpha] (Energy::max_feasible_step, energy.cpp:35-45).
// ---------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ double triple(const double a[3], const double b[3], const double c[3]) {
    // a . (b x c)
    return a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
}

// Largest safe step along A + t dA keeping det > 0 (positive_det_feasible_step.cpp:13-84): columns scaled by their
// largest magnitude, det(A + t dA) = c0 + c1 t + c2 t^2 + c3 t^3, its stationary points split [0, 1] into monotone
// pieces, the first piece whose right end is not safely positive is bisected to the last safe t, and 0.9 of it is
// returned.  A and dA are row-major 3x3 (column j = entries j, 3 + j, 6 + j).  Requires det A > 0.
// The bisection runs in steps (bisect) so that the caller can stop a lane whose step cannot be the minimum.
struct DetStep {
    double c0, c1, c2, c3;
    double lo = 0.0, hi = 1.0;
    bool active = false; // bisecting [lo, hi]
    double result = 1.0; // the step once not active
    static constexpr double kEps = 2.220446049250313e-16;
    __device__ double value(double t) const {
        return fma(fma(fma(c3, t, c2), t, c1), t, c0);
    }
    __device__ double tolerance(double t) const {
        return 64.0 * kEps * (((fabs(c3) * t + fabs(c2)) * t + fabs(c1)) * t + c0);
    }
    __device__ void setup(const double A[9], const double dA[9]) {
        double a[3][3], d[3][3]; // a[j] = column j of A / scale_j
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double scale = 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i)
                scale = fmax(scale, fmax(fabs(A[3 * i + j]), fabs(dA[3 * i + j])));
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                a[j][i] = A[3 * i + j] / scale;
                d[j][i] = dA[3 * i + j] / scale;
            }
        }
        c0 = triple(a[0], a[1], a[2]);
        c1 = triple(d[0], a[1], a[2]) + triple(a[0], d[1], a[2]) + triple(a[0], a[1], d[2]);
        c2 = triple(d[0], d[1], a[2]) + triple(d[0], a[1], d[2]) + triple(a[0], d[1], d[2]);
        c3 = triple(d[0], d[1], d[2]);
        if (!(c0 > 0.0)) {
            result = 0.0;
            return;
        }
        // Stationary points of the cubic inside (0, 1), in increasing order (2 marks an absent one); with 0 and 1
        // they split [0, 1] into the pieces [0, k1], [k1, k2], [k2, 1].  Registers only (no dynamically indexed
        // array).
        const auto inside = [](double t) { return t > 0.0 && t < 1.0 ? t : 2.0; };
        double r1 = 2.0, r2 = 2.0;
        const double qa = 3.0 * c3, qb = 2.0 * c2, qc = c1;
        if (qa == 0.0) {
            if (qb != 0.0)
                r1 = inside(-qc / qb);
        } else {
            double disc = fma(qb, qb, -4.0 * qa * qc);
            const double err = 8.0 * kEps * (qb * qb + fabs(4.0 * qa * qc));
            if (disc >= -err) {
                disc = fmax(0.0, disc);
                const double q = -0.5 * (qb + copysign(sqrt(disc), qb));
                if (q == 0.0) {
                    r1 = inside(-qb / (2.0 * qa));
                } else {
                    r1 = inside(q / qa);
                    r2 = inside(qc / q);
                }
            }
        }
        const double k1 = fmin(r1, r2), k2 = fmax(r1, r2);
        // The first piece whose right end is not safely positive is bisected to the last safe t.
        bool found = false;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double end = i == 0 ? k1 : (i == 1 ? k2 : 1.0);
            if (found || end > 1.0)
                continue;
            if (value(end) > tolerance(end)) {
                lo = end;
            } else {
                hi = end;
                found = true;
            }
        }
        active = found;
    }
    // Up to `steps` halvings (all of them for steps < 0); the bracket is final once the midpoint equals an end (at
    // most ~1100 halvings for doubles in [0, 1]), and the step is then 0.9 lo.
    __device__ void bisect(int steps) {
        for (int k = 0; active && k != steps; ++k) {
            const double mid = lo + (hi - lo) / 2.0;
            if (mid == lo || mid == hi) {
                active = false;
                result = 0.9 * lo;
                break;
            }
            if (value(mid) > tolerance(mid))
                lo = mid;
            else
                hi = mid;
        }
    }
    // An upper bound of the final step: 0.9 hi while bisecting (lo stays below hi), else the step itself.
    __device__ double upper() const {
        return active ? 0.9 * hi : result;
    }
};

struct StepParams {
    const double* Xr;
    const int32_t* conn;
    const double* theta_snap;
    const double* x;
    const double* dx;
    const int32_t* seg_tbeg; // num_segments + 1 tet offsets (one segment for a plain mesh)
    int num_segments;
    unsigned long long* alpha_bits; // per segment, bit patterns of non-negative doubles (ordered as integers)
    unsigned* status;
    int num_tets;
    int num_dofs;
};

__global__ void k_step_init(unsigned long long* alpha_bits, int n, unsigned* status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        alpha_bits[i] = static_cast<unsigned long long>(__double_as_longlong(1.0));
    if (i == 0)
        status[0] = 0u;
}

// Bisection steps after which a wave publishes an upper bound of its minimum and drops the lanes that cannot reach
// the segment's minimum (see the kernel).
constexpr int kStepChecks[2] = {4, 8};

__global__ __launch_bounds__(256) void k_feasible_step(StepParams p) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    int seg = 0;
    unsigned flags = 0;
    DetStep ds;
    // Every entry of x and dx is checked (grid stride), also those of vertices that no tet uses.
    for (int i = e; i < p.num_dofs; i += 256 * gridDim.x)
        if (!(isfinite(p.x[i]) && isfinite(p.dx[i])))
            flags |= kStatusNonFiniteX;
    if (e < p.num_tets) {
        int ids[4];
        TetIn t;
        load_tet(p.Xr, p.conn, p.theta_snap, p.x, nullptr, e, ids, t);
        Geo geo;
        rest_geometry(t, geo);
        Kin k;
        kinematics(t, geo, k);
        if (!kin_ok(k))
            flags |= kStatusDetF;
        if (!theta_ok(t.lam, t.mu))
            flags |= kStatusBadTheta;
        // dF = dDu Dm^-1 with dDu(i, c) = (dx_{c+1} - dx_0)_i.
        double du[4][3];
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                du[b][i] = p.dx[3 * static_cast<size_t>(ids[b]) + i];
        double dF[9];
        bool fin = true;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                double s = 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    s = fma(du[c + 1][i] - du[0][i], geo.Di[3 * c + j], s);
                dF[3 * i + j] = s;
                fin = fin && isfinite(s);
            }
        if (!fin)
            flags |= kStatusNonFiniteStep;
        if (!flags)
            ds.setup(k.F, dF);
        int lo = 0, hi = p.num_segments; // segment of tet e: seg_tbeg[seg] <= e < seg_tbeg[seg + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (p.seg_tbeg[mid] <= e)
                lo = mid;
            else
                hi = mid;
        }
        seg = lo;
    }
    // Wave minimum when the whole wave lies in one segment (one atomic), else one atomic per lane.
    const int seg0 = __shfl(seg, 0);
    const bool uniform = __all(seg == seg0 || e >= p.num_tets);
    // Pruning, which leaves the minimum's bits unchanged.  After a few halvings each lane holds an upper bound of its
    // step (0.9 hi, or the step itself once known); the wave publishes the least of them to the segment's word and
    // reads the word back (B, the least upper bound published so far, not below the segment's minimum).  A lane
    // whose step is at least 0.9 lo > B cannot be the minimum and stops (contributing 1, which changes no minimum);
    // the lane attaining the minimum m has 0.9 lo <= m <= B and continues.  So the words end at the same minimum.
    // Lanes past the last tet and lanes with an error flag contribute 1 (the host raises on the flag).
    const bool lane_ok = e < p.num_tets && !flags;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        ds.bisect(c == 0 ? kStepChecks[0] : kStepChecks[1] - kStepChecks[0]);
        if (!__any(ds.active))
            break;
        double ub = lane_ok ? ds.upper() : 1.0;
        unsigned long long* word = p.alpha_bits + (uniform ? seg0 : seg);
        if (uniform) {
#pragma unroll
            for (int off = kWave / 2; off > 0; off >>= 1)
                ub = fmin(ub, __shfl_xor(ub, off));
        }
        // Read first, and publish only a bound that lowers the word: every wave's atomic on one L2 address would
        // serialize.
        double bound = __longlong_as_double(
            static_cast<long long>(__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
        if ((!uniform || (threadIdx.x % kWave) == 0) && (uniform || e < p.num_tets) && ub < bound)
            atomicMin(word, static_cast<unsigned long long>(__double_as_longlong(ub)));
        if (uniform)
            bound = fmin(bound, ub); // the wave's own least upper bound counts too
        if (ds.active && 0.9 * ds.lo > bound) {
            ds.active = false;
            ds.result = 1.0;
        }
    }
    ds.bisect(-1);
    const double alpha = lane_ok ? ds.result : 1.0;
    if (uniform) {
        double m = alpha;
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1)
            m = fmin(m, __shfl_xor(m, off));
        // Positive doubles order as their bit patterns.  A wave whose minimum does not lower the segment's current
        // bound skips the atomic: one L2 address takes every wave's update otherwise.  (A stale read only costs an
        // unneeded atomic.)
        if ((threadIdx.x % kWave) == 0 && m < 1.0) {
            const unsigned long long mb = static_cast<unsigned long long>(__double_as_longlong(m));
            if (mb < __hip_atomic_load(p.alpha_bits + seg0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMin(p.alpha_bits + seg0, mb);
        }
    } else if (e < p.num_tets && alpha < 1.0) {
        atomicMin(p.alpha_bits + seg, static_cast<unsigned long long>(__double_as_longlong(alpha)));
    }
    if (flags)
        atomicOr(p.status, flags);
}

void check_launch(const char* what) {
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess)
        fail(kDevice, std::string(what) + ": " + hipGetErrorString(err));
}

template <class Kernel> void allow_large_lds(Kernel kern, int waves) {
    const hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               160 * 1024 - kWave * waves * 8);
    if (err != hipSuccess)
        fail(kDevice, std::string("hipFuncSetAttribute: ") + hipGetErrorString(err));
}

template <int LPV, int WAVES, bool WANT_G, int KIND, bool WIDE, bool GACC, bool PLASTIC>
void launch_assemble_g(const AssembleParams& p, int num_tiles, size_t bytes, int persistent_grid,
                       hipStream_t stream) {
    const dim3 block(kWave * WAVES);
    if (persistent_grid > 0) {
        auto kern = k_assemble_persistent<LPV, WAVES, WANT_G, KIND, WIDE, GACC, PLASTIC>;
        static bool configured = false;
        if (!configured) {
            allow_large_lds(kern, WAVES);
            configured = true;
        }
        hipLaunchKernelGGL(kern, dim3(std::min(persistent_grid, num_tiles)), block, bytes, stream, p, num_tiles);
    } else {
        auto kern = k_assemble<LPV, WAVES, WANT_G, KIND, WIDE, GACC, PLASTIC>;
        static bool configured = false;
        if (!configured) {
            allow_large_lds(kern, WAVES);
            configured = true;
        }
        hipLaunchKernelGGL(kern, dim3(num_tiles), block, bytes, stream, p);
    }
    check_launch("k_assemble");
}

// Wide incidence records (vertex degree above 255) are only instantiated for LPV = 4, WAVES = 1.
template <int LPV, int WAVES, bool WIDE>
void dispatch_kind(const AssembleParams& p, int want_g, int kind, int num_tiles, size_t bytes, bool gacc,
                   int persistent_grid, hipStream_t s) {
#define LPG_CASE(G, K)                                                                                               \
    if (!!want_g == G && kind == K) {                                                                                 \
        if (p.fp)                                                                                                     \
            return gacc ? launch_assemble_g<LPV, WAVES, G, K, WIDE, true, true>(p, num_tiles, bytes, persistent_grid, s) \
                        : launch_assemble_g<LPV, WAVES, G, K, WIDE, false, true>(p, num_tiles, bytes, persistent_grid, s); \
        return gacc ? launch_assemble_g<LPV, WAVES, G, K, WIDE, true, false>(p, num_tiles, bytes, persistent_grid, s)   \
                    : launch_assemble_g<LPV, WAVES, G, K, WIDE, false, false>(p, num_tiles, bytes, persistent_grid, s); \
    }
    LPG_CASE(false, -1)
    LPG_CASE(true, -1)
    LPG_CASE(false, 0)
    LPG_CASE(true, 0)
    LPG_CASE(false, 1)
    LPG_CASE(true, 1)
#undef LPG_CASE
    fail(kLogic, "k_assemble: unsupported configuration");
}

} // namespace

// Compiled launch shapes (lanes per vertex, waves per workgroup).  The lanes of a vertex split its incidence list
// and their sums combine in lane order, so the number of lanes per vertex fixes the summation order of every value:
// it is 4 for every shape, and the shapes differ only in how many vertices (waves) a workgroup takes, which changes
// no bit of the results.  Meshes with a vertex of more than 255 neighbours (16-bit column positions) use (4, 1).
#define LPG_SHAPES(X) X(4, 1) X(4, 2) X(4, 4)

bool launch_shape_supported(int lanes_per_vertex, int waves, bool wide) {
    if (wide)
        return lanes_per_vertex == 4 && waves == 1;
#define LPG_IS(L, W)                                                                                                 \
    if (lanes_per_vertex == L && waves == W)                                                                          \
        return true;
    LPG_SHAPES(LPG_IS)
#undef LPG_IS
    return false;
}

size_t assemble_lds_bytes(const LaunchConfig& lc, const TileCaps& caps, bool global_acc, int kind) {
    const size_t verts = static_cast<size_t>(lc.waves * (kWave / lc.lanes_per_vertex));
    size_t doubles = 4 * verts; // gradient + energy slots
    if (!global_acc && kind >= 0)
        doubles += static_cast<size_t>(caps.segment);
    return doubles * sizeof(double);
}

size_t assemble_lds_limit(const LaunchConfig& lc) {
    return 160 * 1024 - static_cast<size_t>(kWave) * lc.waves * 8;
}

void launch_clear(unsigned* status, unsigned* counter, hipStream_t stream) {
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, stream, status, counter);
    check_launch("k_clear");
}

void launch_assemble(const DeviceMesh& m, const EvalArgs& a, const LaunchConfig& lc, const int32_t* tile_vbeg,
                     const int32_t* tile_ids, int num_tiles, const TileCaps& caps, bool global_acc,
                     int persistent_grid, hipStream_t stream) {
    if (num_tiles <= 0)
        return;
    AssembleParams p{};
    p.Xr = m.Xr;
    p.theta_snap = m.theta;
    p.inc_ptr = m.inc_ptr;
    p.inc = m.inc;
    p.nbr_ptr = m.nbr_ptr;
    p.nbr = m.nbr;
    p.self_pos = m.self_pos;
    p.tile_vbeg = tile_vbeg;
    p.tile_ids = tile_ids;
    p.x = a.x;
    p.theta = a.theta;
    p.g = a.g;
    p.H = a.H;
    p.vertex_E = a.vertex_E;
    p.status = a.status;
    p.proj_slot = a.proj_slot;
    p.proj_H = a.proj_H;
    p.proj_capacity = a.proj_capacity;
    p.fp = m.fp;
    const size_t bytes = assemble_lds_bytes(lc, caps, global_acc, a.kind);
    if (bytes > assemble_lds_limit(lc))
        fail(kLogic, "launch_assemble: staged tile data exceeds the LDS of a workgroup");
    if (m.wide) {
        if (lc.lanes_per_vertex == 4 && lc.waves == 1)
            return dispatch_kind<4, 1, true>(p, a.want_g, a.kind, num_tiles, bytes, global_acc, persistent_grid,
                                             stream);
        fail(kInvalid, "launch_assemble: meshes with vertex degree above 255 use 4 lanes per vertex and one wave");
    }
#define LPG_CFG(L, W)                                                                                                \
    if (lc.lanes_per_vertex == L && lc.waves == W)                                                                    \
        return dispatch_kind<L, W, false>(p, a.want_g, a.kind, num_tiles, bytes, global_acc, persistent_grid, stream);
    LPG_SHAPES(LPG_CFG)
#undef LPG_CFG
    fail(kInvalid, "launch_assemble: unsupported launch configuration");
}

void launch_energy_reduce(const double* vertex_E, const int32_t* chunk_begin, int num_chunks, double* chunk_sum,
                          const int32_t* seg_chunk, int num_segments, double* out, unsigned* status,
                          hipStream_t stream, const unsigned* counter, void* readback) {
    if (readback && num_segments != 1)
        fail(kLogic, "launch_energy_reduce: the fused read-back needs exactly one segment");
    if (num_chunks > 0) {
        hipLaunchKernelGGL(k_energy_chunks, dim3(num_chunks), dim3(256), 0, stream, vertex_E, chunk_begin, chunk_sum);
        check_launch("k_energy_chunks");
    }
    if (num_segments > 0) {
        hipLaunchKernelGGL(k_energy_segments, dim3(num_segments), dim3(256), 0, stream, chunk_sum, seg_chunk, out,
                           status, counter, readback);
        check_launch("k_energy_segments");
    }
}

void launch_readback(const double* energy, unsigned* status, const unsigned* counter, void* readback,
                     hipStream_t stream) {
    hipLaunchKernelGGL(k_readback, dim3(1), dim3(64), 0, stream, energy, status, counter, readback);
    check_launch("k_readback");
}

void launch_energy_total(const double* seg, int n, double* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_energy_total, dim3(1), dim3(256), 0, stream, seg, n, out);
    check_launch("k_energy_total");
}

void launch_project_prepass(const DeviceMesh& m, const EvalArgs& a, hipStream_t stream) {
    if (m.num_tets == 0)
        return;
    ProjectParams p{m.Xr, m.conn, m.theta, a.x, a.theta, m.num_tets, a.status, a.counter, a.proj_slot, a.proj_list, a.proj_H};
    p.fp = m.fp;
    p.proj_capacity = a.proj_capacity;
    hipLaunchKernelGGL(k_project_prepass, dim3((m.num_tets + kPrepassThreads - 1) / kPrepassThreads),
                       dim3(kPrepassThreads), 0, stream, p);
    check_launch("k_project_prepass");
}

void launch_project_jacobi(const DeviceMesh& m, const EvalArgs& a, hipStream_t stream) {
    if (m.num_tets == 0)
        return;
    ProjectParams p{m.Xr, m.conn, m.theta, a.x, a.theta, m.num_tets, a.status, a.counter, a.proj_slot, a.proj_list, a.proj_H};
    p.fp = m.fp;
    p.proj_capacity = a.proj_capacity;
    // One lane per store slot, plus the lane just past the store (it reports an overflowing list): the list length
    // is only known on the device, so the workgroups past it exit at once.
    const int lanes = std::min<int64_t>(m.num_tets, int64_t{a.proj_capacity} + 1);
    const int blocks = (lanes + 63) / 64;
    hipLaunchKernelGGL(k_project_jacobi, dim3(blocks), dim3(64), 0, stream, p);
    check_launch("k_project_jacobi");
}

static void launch_theta_vjp(const DeviceMesh& m, const double* x, const double* theta, const double* scale,
                             double scale_val, bool has_s, const double* w, double* out, unsigned* status,
                             hipStream_t stream) {
    if (m.num_tets == 0)
        return;
    ElemParams p{m.Xr, m.conn, m.theta, x, theta, w, scale, scale_val, out, status, m.num_tets};
    p.fp = m.fp;
    const dim3 grid((m.num_tets + 255) / 256), block(256);
    if (has_s && w)
        hipLaunchKernelGGL((k_theta_vjp<true, true>), grid, block, 0, stream, p);
    else if (has_s)
        hipLaunchKernelGGL((k_theta_vjp<true, false>), grid, block, 0, stream, p);
    else
        hipLaunchKernelGGL((k_theta_vjp<false, true>), grid, block, 0, stream, p);
    check_launch("k_theta_vjp");
}

void launch_dE_dtheta(const DeviceMesh& m, const double* x, const double* theta, double* out, unsigned* status,
                      hipStream_t stream) {
    launch_theta_vjp(m, x, theta, nullptr, 1.0, true, nullptr, out, status, stream);
}

void launch_force_vjp_theta(const DeviceMesh& m, const double* x, const double* theta, const double* w,
                            double* out, unsigned* status, hipStream_t stream) {
    launch_theta_vjp(m, x, theta, nullptr, 0.0, false, w, out, status, stream);
}

static void launch_gather(int mode, const DeviceMesh& m, const double* x, const double* theta, const double* v,
                          const double* scale, double* out, unsigned* status, hipStream_t stream) {
    if (m.num_vertices == 0)
        return;
    GatherParams p{m.Xr,
                   m.conn,
                   m.theta,
                   m.inc_ptr,
                   static_cast<const int32_t*>(m.inc),
                   m.wide ? static_cast<int>(sizeof(IncWide)) : static_cast<int>(sizeof(IncNarrow)),
                   x,
                   theta,
                   v,
                   scale,
                   out,
                   status,
                   m.num_vertices};
    p.fp = m.fp;
    const long long threads = static_cast<long long>(m.num_vertices) * kGatherLanes;
    const dim3 grid(static_cast<unsigned>((threads + kGatherThreads - 1) / kGatherThreads)), block(kGatherThreads);
    if (mode == 0)
        hipLaunchKernelGGL(k_vertex_gather<0>, grid, block, 0, stream, p);
    else if (mode == 1)
        hipLaunchKernelGGL(k_vertex_gather<1>, grid, block, 0, stream, p);
    else
        hipLaunchKernelGGL(k_vertex_gather<2>, grid, block, 0, stream, p);
    check_launch("k_vertex_gather");
}

void launch_force_jvp_theta(const DeviceMesh& m, const double* x, const double* theta, const double* dtheta,
                            double* out, unsigned* status, hipStream_t stream) {
    launch_gather(1, m, x, theta, dtheta, nullptr, out, status, stream);
}

void launch_force_fp(const DeviceMesh& m, const double* x, const double* theta, const double* fp, const double* w,
                     const double* dfp, const int32_t* tet_inc, double* scratch, double* out, unsigned* status,
                     hipStream_t stream) {
    FpParams p{m.Xr, m.conn, m.theta, x, theta, m.fp, fp, w, dfp, tet_inc, w ? out : scratch, status, m.num_tets};
    const dim3 grid((m.num_tets + 255) / 256), block(256);
    if (w) {
        if (m.num_tets > 0) {
            hipLaunchKernelGGL(k_fp_vjp, grid, block, 0, stream, p);
            check_launch("k_fp_vjp");
        }
        return;
    }
    if (m.num_tets > 0) {
        hipLaunchKernelGGL(k_fp_elem, grid, block, 0, stream, p);
        check_launch("k_fp_elem");
    }
    if (m.num_vertices == 0)
        return;
    // Vertices with no incidences (every vertex when there are no tets) get zeros.
    const long long threads = static_cast<long long>(m.num_vertices) * kElemGatherLanes;
    const dim3 ggrid(static_cast<unsigned>((threads + 255) / 256));
    hipLaunchKernelGGL(k_elem_gather, ggrid, dim3(256), 0, stream, m.inc_ptr, scratch, out, m.num_vertices);
    check_launch("k_elem_gather");
}

void launch_hvp(const DeviceMesh& m, const double* x, const double* theta, const double* v, double* out,
                unsigned* status, hipStream_t stream) {
    launch_gather(0, m, x, theta, v, nullptr, out, status, stream);
}

void launch_vjp(const DeviceMesh& m, const double* x, const double* theta, const double* gE, const double* w,
                double* gx, double* gtheta, unsigned* status, hipStream_t stream) {
    if (gx)
        launch_gather(2, m, x, theta, w, gE, gx, status, stream);
    if (gtheta) // with neither cotangent, the kernel writes s dE/dtheta with s = 0
        launch_theta_vjp(m, x, theta, gE, 0.0, gE != nullptr || !w, w, gtheta, status, stream);
}

static unsigned elementwise_blocks(int64_t n) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)));
}

void launch_lame(const double* E, const double* nu, int64_t n, double* theta, hipStream_t stream) {
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_lame, dim3(elementwise_blocks(n)), dim3(256), 0, stream, E, nu, static_cast<long long>(n),
                       reinterpret_cast<double2*>(theta));
    check_launch("k_lame");
}

void launch_lame_vjp(const double* E, const double* nu, const double* gtheta, int64_t n, double* gE, double* gnu,
                     hipStream_t stream) {
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_lame_vjp, dim3(elementwise_blocks(n)), dim3(256), 0, stream, E, nu,
                       reinterpret_cast<const double2*>(gtheta), static_cast<long long>(n), gE, gnu);
    check_launch("k_lame_vjp");
}

void launch_feasible_step(const DeviceMesh& m, const double* x, const double* dx, const int32_t* seg_tbeg,
                          int num_segments, double* alphas, unsigned* status, hipStream_t stream) {
    auto* bits = reinterpret_cast<unsigned long long*>(alphas);
    hipLaunchKernelGGL(k_step_init, dim3((num_segments + 255) / 256), dim3(256), 0, stream, bits, num_segments,
                       status);
    check_launch("k_step_init");
    const int num_dofs = 3 * m.num_vertices;
    if (m.num_tets == 0 && num_dofs == 0)
        return;
    StepParams p{m.Xr, m.conn, m.theta, x, dx, seg_tbeg, num_segments, bits, status, m.num_tets, num_dofs};
    const int blocks = std::max(1, (std::max(m.num_tets, num_dofs) + 255) / 256);
    hipLaunchKernelGGL(k_feasible_step, dim3(std::min(blocks, (m.num_tets + 255) / 256 + 1024)), dim3(256), 0,
                       stream, p);
    check_launch("k_feasible_step");
}

} // namespace lpg
