// HIP kernels for gfx950: neo-Hookean Tet P1 energy / gradient / Hessian assembly and its derivatives.
//
// Layout in HBM (see lpg_internal.hpp): one 128-byte record per tet (Dm^-1, w, lambda, mu, vertex ids), the
// displacement vector x[3V], per-vertex incidence lists (tet id + the column-block position of each of the tet's
// vertices in this vertex's rows), and the CSR row pointer.
//
// Assembly is vertex-owned: the rows 3v..3v+2 of H form one contiguous segment of 9 |nbr(v)| values, and only
// the lanes of vertex v write it.  A lane walks v's incident tets in element order, rebuilds F, J, F^-T and log J
// in registers, and adds the three rows of the tet's 12x12 Hessian that belong to v, block by block, to the
// segment staged in LDS.  Each value is therefore summed in element order by a single lane, as the reference's
// column-owner assembly does (src/core/dof/energy_assembler.cpp:249-273): no atomics on HBM, bitwise
// reproducible results, and every H value is written to HBM exactly once with coalesced stores.
//
// The element Hessian uses the closed form of B^T A B for Tet P1 and neo-Hookean (A from neo_hookean.cpp:43-60):
//   H_ab = w [ mu (g_a . g_b) I3 + lambda q_a q_b^T - p q_b q_a^T ],  q_b = F^-T g_b,  p = lambda log J - mu,
// where g_b is row b of G = dN/dxi Dm^-1.  The gradient block is w (mu F g_a + p q_a) and the energy
// w (mu/2 (|F|^2 - 3) - mu log J + lambda/2 log^2 J).
//
// The Projected Hessian clamps the spectrum of the 12x12 element matrix (mesh_energy.cpp:16-27).  With the
// orthonormal Helmert basis P (4x3) of the complement of (1,1,1,1), H_e = Q M Q^T with Q = P (x) I3 and the 9x9
// M = w (C (x) I3) A (C (x) I3)^T, C = P^T G; hence proj(H_e) = Q proj(M) Q^T, and by Sylvester's law M is PSD
// exactly when A is.  A prepass tests A with an LDL^T factorization; only elements with an indefinite A get a
// 9x9 Jacobi eigen-clamp of M, and the assembly reconstructs their blocks from proj(M).
#include "lpg_internal.hpp"

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

// Helmert basis of 1-perp in R^4: P(a, m).
__device__ __constant__ double kHelmert[4][3] = {
    {0.70710678118654752440, 0.40824829046386301637, 0.28867513459481288225},
    {-0.70710678118654752440, 0.40824829046386301637, 0.28867513459481288225},
    {0.0, -0.81649658092772603273, 0.28867513459481288225},
    {0.0, 0.0, -0.86602540378443864676},
};

struct TetIn {
    double Di[9]; // Dm^-1 row-major
    double w, lam, mu;
    int ids[4];
};

__device__ __forceinline__ void load_tet(const double* __restrict__ rec, const double* __restrict__ theta, int e,
                                         TetIn& t) {
    const double2* r = reinterpret_cast<const double2*>(rec + static_cast<size_t>(e) * kRecDoubles);
    const double2 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3], r4 = r[4], r5 = r[5];
    const int4 id = *reinterpret_cast<const int4*>(r + 6);
    t.Di[0] = r0.x;
    t.Di[1] = r0.y;
    t.Di[2] = r1.x;
    t.Di[3] = r1.y;
    t.Di[4] = r2.x;
    t.Di[5] = r2.y;
    t.Di[6] = r3.x;
    t.Di[7] = r3.y;
    t.Di[8] = r4.x;
    t.w = r4.y;
    if (theta) {
        const double2 th = reinterpret_cast<const double2*>(theta)[e];
        t.lam = th.x;
        t.mu = th.y;
    } else {
        t.lam = r5.x;
        t.mu = r5.y;
    }
    t.ids[0] = id.x;
    t.ids[1] = id.y;
    t.ids[2] = id.z;
    t.ids[3] = id.w;
}

__device__ __forceinline__ void load_u(const double* __restrict__ x, const int ids[4], double u[4][3]) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const double* p = x + 3 * static_cast<size_t>(ids[b]);
        u[b][0] = p[0];
        u[b][1] = p[1];
        u[b][2] = p[2];
    }
}

// Deformation gradient and its inverse transpose: F = I + Du Dm^-1, a = F^-T = cof(F) / det F.
struct Kin {
    double F[9]; // row-major
    double a[9]; // row-major, a(i, j) = F^-T(i, j)
    double J, L;
};

__device__ __forceinline__ void kinematics(const double u[4][3], const double Di[9], Kin& k) {
    double Du[9]; // Du(i, c) = (u_{c+1} - u_0)_i
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            Du[3 * i + c] = u[c + 1][i] - u[0][i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = (i == j) ? 1.0 : 0.0;
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

template <bool WIDE> struct IncT;
template <> struct IncT<false> {
    using type = IncNarrow;
    __device__ static int colpos(const IncNarrow& r, int b) {
        return (r.colpos >> (8 * b)) & 0xff;
    }
};
template <> struct IncT<true> {
    using type = IncWide;
    __device__ static int colpos(const IncWide& r, int b) {
        const uint32_t w = b < 2 ? r.colpos01 : r.colpos23;
        return (w >> (16 * (b & 1))) & 0xffff;
    }
};

// Row `a` of a 4x3 register array as a masked sum.  A select chain here is rewritten by the compiler into a
// dynamically indexed private array (scratch memory); multiplying by 0.0 / 1.0 masks keeps it in registers.
__device__ __forceinline__ void select_row(const double G[4][3], int a, double out[3]) {
    const double m0 = a == 0 ? 1.0 : 0.0, m1 = a == 1 ? 1.0 : 0.0, m2 = a == 2 ? 1.0 : 0.0, m3 = a == 3 ? 1.0 : 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
        out[j] = fma(m3, G[3][j], fma(m2, G[2][j], fma(m1, G[1][j], m0 * G[0][j])));
}

// ---------------------------------------------------------------------------------------------------------------
// Assembly kernel: one workgroup per vertex tile, LPV lanes per vertex (lane s of a vertex owns the column blocks
// whose position is congruent to s modulo LPV).  KIND: -1 none, 0 Exact, 1 Projected.
// ---------------------------------------------------------------------------------------------------------------
struct AssembleParams {
    const double* rec;
    const int32_t* inc_ptr;
    const void* inc;
    const int32_t* row_ptr;
    const int32_t* tile_vbeg;
    const int32_t* tile_ids; // tiles to process (index into tile_vbeg), or null for 0..num_tiles-1
    const double* x;
    const double* theta;
    double* g;
    double* H;
    double* vertex_E; // per-vertex element-order sum of the energies of the tets whose slot 0 it is
    unsigned* status;
    const int32_t* proj_slot;
    const double* proj_M;
};

// One tile: zero the tile's H segment in LDS (or in HBM for GLOBAL_ACC), accumulate, write it back coalesced,
// write the gradient rows and the per-vertex energies, and OR the status flags.
template <int LPV, int WAVES, bool WANT_G, int KIND, bool WIDE, bool GLOBAL_ACC>
__device__ __forceinline__ void assemble_tile(const AssembleParams& p, int tile, double* lds) {
    using Inc = typename IncT<WIDE>::type;
    const Inc* __restrict__ inc = static_cast<const Inc*>(p.inc);
    constexpr int kThreads = kWave * WAVES;
    const int vb = p.tile_vbeg[tile], ve = p.tile_vbeg[tile + 1];
    const int seg0 = KIND >= 0 ? p.row_ptr[3 * vb] : 0;
    const int seg1 = KIND >= 0 ? p.row_ptr[3 * ve] : 0;
    double* acc = GLOBAL_ACC ? p.H + seg0 : lds;
    if (KIND >= 0) {
        for (int i = threadIdx.x; i < seg1 - seg0; i += kThreads)
            acc[i] = 0.0;
        __syncthreads();
    }
    const int sub = threadIdx.x % LPV;
    const int v = vb + static_cast<int>(threadIdx.x) / LPV;
    unsigned flags = 0;
    double esum = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (v < ve) {
        {
            const double* xv = p.x + 3 * static_cast<size_t>(v);
            if (!(isfinite(xv[0]) && isfinite(xv[1]) && isfinite(xv[2])))
                flags |= kStatusNonFiniteX;
        }
        const int rbase = KIND >= 0 ? p.row_ptr[3 * v] - seg0 : 0;
        const int rowlen = KIND >= 0 ? p.row_ptr[3 * v + 1] - p.row_ptr[3 * v] : 0;
        const int kb = p.inc_ptr[v], ke = p.inc_ptr[v + 1];
        for (int k = kb; k < ke; ++k) {
            const Inc r = inc[k];
            const int e = r.tet;
            TetIn t;
            load_tet(p.rec, p.theta, e, t);
            double u[4][3];
            load_u(p.x, t.ids, u);
            const int a = (t.ids[0] == v) ? 0 : (t.ids[1] == v) ? 1 : (t.ids[2] == v) ? 2 : 3;
            Kin kin;
            kinematics(u, t.Di, kin);
            if (!kin_ok(kin))
                flags |= kStatusDetF;
            if (!theta_ok(t.lam, t.mu))
                flags |= kStatusBadTheta;
            double G[4][3];
            shape_gradients(t.Di, G);
            double ga[3];
            select_row(G, a, ga);
            const double p_ = fma(t.lam, kin.L, -t.mu);
            // q_a = F^-T g_a
            double qa[3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                qa[i] = dot3(kin.a[3 * i], kin.a[3 * i + 1], kin.a[3 * i + 2], ga[0], ga[1], ga[2]);
            if (sub == 0) {
                if (WANT_G) {
                    // w (mu F g_a + p q_a)
                    double Fg[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i)
                        Fg[i] = dot3(kin.F[3 * i], kin.F[3 * i + 1], kin.F[3 * i + 2], ga[0], ga[1], ga[2]);
                    g0 = fma(t.w, fma(t.mu, Fg[0], p_ * qa[0]), g0);
                    g1 = fma(t.w, fma(t.mu, Fg[1], p_ * qa[1]), g1);
                    g2 = fma(t.w, fma(t.mu, Fg[2], p_ * qa[2]), g2);
                }
                if (a == 0) {
                    double sq = 0.0;
#pragma unroll
                    for (int q = 0; q < 9; ++q)
                        sq = fma(kin.F[q], kin.F[q], sq);
                    const double psi = fma(0.5 * t.lam * kin.L, kin.L, fma(0.5 * t.mu, sq - 3.0, -(t.mu * kin.L)));
                    esum = fma(t.w, psi, esum);
                }
            }
            if (KIND >= 0) {
                double* row = acc + rbase;
                bool projected = false;
                int slot = -1;
                if (KIND == 1) {
                    slot = p.proj_slot[e];
                    projected = slot >= 0;
                }
                if (!projected) {
                    const double s2 = t.w * t.lam, s3 = t.w * p_, sw = t.w * t.mu;
                    double xa[3], ya[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        xa[i] = s2 * qa[i];
                        ya[i] = s3 * qa[i];
                    }
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int cp = IncT<WIDE>::colpos(r, b);
                        if (LPV == 1 || (cp % LPV) == sub) {
                        double qb[3];
#pragma unroll
                        for (int i = 0; i < 3; ++i)
                            qb[i] = dot3(kin.a[3 * i], kin.a[3 * i + 1], kin.a[3 * i + 2], G[b][0], G[b][1], G[b][2]);
                        const double s1 = sw * dot3(ga[0], ga[1], ga[2], G[b][0], G[b][1], G[b][2]);
                        double* blk = row + 3 * cp;
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int kk = 0; kk < 3; ++kk) {
                                double val = fma(xa[i], qb[kk], -(qb[i] * ya[kk]));
                                if (i == kk)
                                    val += s1;
                                accum<GLOBAL_ACC>(blk + i * rowlen + kk, val);
                            }
                        }
                    }
                } else {
                    // proj(H)_ab = sum_mn P(a,m) P(b,n) Mp[(m,i),(n,k)], Mp packed upper-triangular row-major.
                    const double* Mp = p.proj_M + 45 * static_cast<size_t>(slot);
                    double Pa[3];
                    {
                        double Ph[4][3];
#pragma unroll
                        for (int b = 0; b < 4; ++b)
#pragma unroll
                            for (int m = 0; m < 3; ++m)
                                Ph[b][m] = kHelmert[b][m];
                        select_row(Ph, a, Pa);
                    }
                    // R[n][i][k] = sum_m Pa[m] Mp[(3m+i), (3n+k)]
                    double R[3][3][3];
#pragma unroll
                    for (int n = 0; n < 3; ++n)
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int kk = 0; kk < 3; ++kk) {
                                double s = 0.0;
#pragma unroll
                                for (int m = 0; m < 3; ++m) {
                                    const int rr = 3 * m + i, cc = 3 * n + kk;
                                    const int lo = rr < cc ? rr : cc, hi = rr < cc ? cc : rr;
                                    const int idx = lo * 9 - lo * (lo - 1) / 2 + (hi - lo);
                                    s = fma(Pa[m], Mp[idx], s);
                                }
                                R[n][i][kk] = s;
                            }
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int cp = IncT<WIDE>::colpos(r, b);
                        if (LPV == 1 || (cp % LPV) == sub) {
                        double* blk = row + 3 * cp;
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int kk = 0; kk < 3; ++kk) {
                                const double val = dot3(kHelmert[b][0], kHelmert[b][1], kHelmert[b][2],
                                                        R[0][i][kk], R[1][i][kk], R[2][i][kk]);
                                accum<GLOBAL_ACC>(blk + i * rowlen + kk, val);
                            }
                        }
                    }
                }
            }
        }
        if (sub == 0)
            p.vertex_E[v] = esum;
        if (WANT_G && sub == 0) {
            double* gv = p.g + 3 * static_cast<size_t>(v);
            gv[0] = g0;
            gv[1] = g1;
            gv[2] = g2;
            if (!(isfinite(g0) && isfinite(g1) && isfinite(g2)))
                flags |= kStatusNonFiniteOut;
        }
    }
    if (KIND >= 0) {
        __syncthreads();
        if (!GLOBAL_ACC) {
            for (int i = threadIdx.x; i < seg1 - seg0; i += kThreads) {
                const double val = acc[i];
                if (!isfinite(val))
                    flags |= kStatusNonFiniteOut;
                p.H[seg0 + i] = val;
            }
        } else {
            for (int i = threadIdx.x; i < seg1 - seg0; i += kThreads)
                if (!isfinite(acc[i]))
                    flags |= kStatusNonFiniteOut;
        }
    }
    if (flags)
        atomicOr(p.status, flags);
}

template <int LPV, int WAVES, bool WANT_G, int KIND, bool WIDE, bool GLOBAL_ACC>
__global__ __launch_bounds__(kWave* WAVES) void k_assemble(AssembleParams p) {
    extern __shared__ double lds[];
    const int tile = p.tile_ids ? p.tile_ids[blockIdx.x] : static_cast<int>(blockIdx.x);
    assemble_tile<LPV, WAVES, WANT_G, KIND, WIDE, GLOBAL_ACC>(p, tile, lds);
}

// Persistent form: a fixed grid walks the tile list with a grid stride (the batched launch).
template <int LPV, int WAVES, bool WANT_G, int KIND, bool WIDE, bool GLOBAL_ACC>
__global__ __launch_bounds__(kWave* WAVES) void k_assemble_persistent(AssembleParams p, int num_tiles) {
    extern __shared__ double lds[];
    for (int t = blockIdx.x; t < num_tiles; t += gridDim.x) {
        const int tile = p.tile_ids ? p.tile_ids[t] : t;
        assemble_tile<LPV, WAVES, WANT_G, KIND, WIDE, GLOBAL_ACC>(p, tile, lds);
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
// launch shape, and whether the mesh is evaluated alone or inside a batch.
template <int N> __device__ __forceinline__ double tree_sum(double* red, double s) {
    red[threadIdx.x] = s;
    __syncthreads();
#pragma unroll
    for (int k = N / 2; k > 0; k >>= 1) {
        if (static_cast<int>(threadIdx.x) < k)
            red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    return red[0];
}

__global__ __launch_bounds__(256) void k_energy_chunks(const double* __restrict__ values,
                                                       const int32_t* __restrict__ chunk_begin,
                                                       double* __restrict__ chunk_sum) {
    __shared__ double red[256];
    const int c = blockIdx.x;
    const int b = chunk_begin[c], e = chunk_begin[c + 1];
    double s = 0.0;
    for (int i = b + static_cast<int>(threadIdx.x); i < e; i += 256)
        s += values[i];
    const double total = tree_sum<256>(red, s);
    if (threadIdx.x == 0)
        chunk_sum[c] = total;
}

__global__ __launch_bounds__(256) void k_energy_segments(const double* __restrict__ chunk_sum,
                                                         const int32_t* __restrict__ seg_chunk,
                                                         double* __restrict__ out, unsigned* status) {
    __shared__ double red[256];
    const int m = blockIdx.x;
    double s = 0.0;
    for (int i = seg_chunk[m] + static_cast<int>(threadIdx.x); i < seg_chunk[m + 1]; i += 256)
        s += chunk_sum[i];
    const double total = tree_sum<256>(red, s);
    if (threadIdx.x == 0) {
        out[m] = total;
        if (!isfinite(total))
            atomicOr(status, kStatusNonFiniteOut);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Projection: prepass (PSD test of the 9x9 material tangent A by LDL^T) and 9x9 Jacobi eigen-clamp of M.
// ---------------------------------------------------------------------------------------------------------------
struct ProjectParams {
    const double* rec;
    const double* x;
    const double* theta;
    int num_tets;
    unsigned* status;
    unsigned* counter;
    int32_t* proj_slot;
    int32_t* proj_list;
    double* proj_M;
};

__device__ __forceinline__ int sym_index(int r, int c) {
    const int lo = r < c ? r : c, hi = r < c ? c : r;
    return lo * 9 - lo * (lo - 1) / 2 + (hi - lo);
}

__global__ __launch_bounds__(256) void k_project_prepass(ProjectParams p) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= p.num_tets)
        return;
    TetIn t;
    load_tet(p.rec, p.theta, e, t);
    double u[4][3];
    load_u(p.x, t.ids, u);
    Kin k;
    kinematics(u, t.Di, k);
    int32_t slot = -1;
    if (kin_ok(k) && theta_ok(t.lam, t.mu)) {
        const double pr = fma(t.lam, k.L, -t.mu);
        // A(i+3j, m+3l) = mu d_im d_jl + lambda a_ij a_ml - p a_il a_mj, packed upper (45).
        double A[45];
#pragma unroll
        for (int r = 0; r < 9; ++r)
#pragma unroll
            for (int c = r; c < 9; ++c) {
                const int i = r % 3, j = r / 3, m = c % 3, l = c / 3;
                double val = fma(t.lam * k.a[3 * i + j], k.a[3 * m + l], -(pr * k.a[3 * i + l] * k.a[3 * m + j]));
                if (i == m && j == l)
                    val += t.mu;
                A[sym_index(r, c)] = val;
            }
        double scale = 0.0;
#pragma unroll
        for (int r = 0; r < 9; ++r)
            scale = fmax(scale, fabs(A[sym_index(r, r)]));
        // In-place LDL^T; positive definite iff every pivot is positive.
        bool pd = true;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const double d = A[sym_index(j, j)];
            pd = pd && (d > 1e-13 * scale);
            const double inv = 1.0 / d;
#pragma unroll
            for (int i = j + 1; i < 9; ++i) {
                const double lij = A[sym_index(j, i)] * inv;
#pragma unroll
                for (int c = i; c < 9; ++c)
                    A[sym_index(i, c)] -= lij * A[sym_index(j, c)];
            }
        }
        if (!pd) {
            slot = static_cast<int32_t>(atomicAdd(p.counter, 1u));
            p.proj_list[slot] = e;
        }
    }
    p.proj_slot[e] = slot;
}

// One thread per indefinite element (grid-stride over the compacted list).
__global__ __launch_bounds__(64) void k_project_jacobi(ProjectParams p) {
    const unsigned count = *p.counter;
    for (unsigned idx = blockIdx.x * 64 + threadIdx.x; idx < count; idx += gridDim.x * 64) {
        const int e = p.proj_list[idx];
        TetIn t;
        load_tet(p.rec, p.theta, e, t);
        double u[4][3];
        load_u(p.x, t.ids, u);
        Kin k;
        kinematics(u, t.Di, k);
        double G[4][3];
        shape_gradients(t.Di, G);
        const double pr = fma(t.lam, k.L, -t.mu);
        // c_m = sum_a P(a,m) g_a, r_m = F^-T c_m.
        double cm[3][3], rm[3][3];
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                cm[m][j] = fma(kHelmert[3][m], G[3][j],
                               fma(kHelmert[2][m], G[2][j], fma(kHelmert[1][m], G[1][j], kHelmert[0][m] * G[0][j])));
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                rm[m][i] = dot3(k.a[3 * i], k.a[3 * i + 1], k.a[3 * i + 2], cm[m][0], cm[m][1], cm[m][2]);
        // M[(3m+i), (3n+k)] = w [mu (c_m . c_n) d_ik + lambda r_m,i r_n,k - p r_n,i r_m,k], packed upper.
        double M[45];
#pragma unroll
        for (int r = 0; r < 9; ++r)
#pragma unroll
            for (int c = r; c < 9; ++c) {
                const int m = r / 3, i = r % 3, n = c / 3, kk = c % 3;
                double val = fma(t.lam * rm[m][i], rm[n][kk], -(pr * rm[n][i] * rm[m][kk]));
                if (i == kk)
                    val = fma(t.mu, dot3(cm[m][0], cm[m][1], cm[m][2], cm[n][0], cm[n][1], cm[n][2]), val);
                M[sym_index(r, c)] = t.w * val;
            }
        double V[9][9];
#pragma unroll
        for (int r = 0; r < 9; ++r)
#pragma unroll
            for (int c = 0; c < 9; ++c)
                V[r][c] = (r == c) ? 1.0 : 0.0;
        double scale = 0.0;
#pragma unroll
        for (int q = 0; q < 45; ++q)
            scale = fmax(scale, fabs(M[q]));
        bool converged = false;
        for (int sweep = 0; sweep < 20 && !converged; ++sweep) {
            double off = 0.0;
#pragma unroll
            for (int r = 0; r < 9; ++r)
#pragma unroll
                for (int c = r + 1; c < 9; ++c)
                    off = fmax(off, fabs(M[sym_index(r, c)]));
            if (off <= 1e-17 * scale || off == 0.0) {
                converged = true;
                break;
            }
#pragma unroll
            for (int pp = 0; pp < 9; ++pp)
#pragma unroll
                for (int qq = pp + 1; qq < 9; ++qq) {
                    const double apq = M[sym_index(pp, qq)];
                    if (apq != 0.0) {
                        const double app = M[sym_index(pp, pp)], aqq = M[sym_index(qq, qq)];
                        const double tau = (aqq - app) / (2.0 * apq);
                        const double tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        const double c = 1.0 / sqrt(1.0 + tt * tt), s = tt * c;
#pragma unroll
                        for (int r = 0; r < 9; ++r) {
                            if (r == pp || r == qq)
                                continue;
                            const double mrp = M[sym_index(r, pp)], mrq = M[sym_index(r, qq)];
                            M[sym_index(r, pp)] = fma(c, mrp, -(s * mrq));
                            M[sym_index(r, qq)] = fma(s, mrp, c * mrq);
                        }
                        M[sym_index(pp, pp)] = fma(-tt, apq, app);
                        M[sym_index(qq, qq)] = fma(tt, apq, aqq);
                        M[sym_index(pp, qq)] = 0.0;
#pragma unroll
                        for (int r = 0; r < 9; ++r) {
                            const double vrp = V[r][pp], vrq = V[r][qq];
                            V[r][pp] = fma(c, vrp, -(s * vrq));
                            V[r][qq] = fma(s, vrp, c * vrq);
                        }
                    }
                }
        }
        if (!converged) {
            double off = 0.0;
#pragma unroll
            for (int r = 0; r < 9; ++r)
#pragma unroll
                for (int c = r + 1; c < 9; ++c)
                    off = fmax(off, fabs(M[sym_index(r, c)]));
            converged = off <= 1e-15 * scale;
        }
        double lam[9];
#pragma unroll
        for (int r = 0; r < 9; ++r)
            lam[r] = fmax(M[sym_index(r, r)], 0.0);
        double* out = p.proj_M + 45 * static_cast<size_t>(idx);
        bool fin = true;
#pragma unroll
        for (int r = 0; r < 9; ++r)
#pragma unroll
            for (int c = r; c < 9; ++c) {
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < 9; ++q)
                    s = fma(V[r][q] * lam[q], V[c][q], s);
                out[sym_index(r, c)] = s;
                fin = fin && isfinite(s);
            }
        if (!converged || !fin)
            atomicOr(p.status, kStatusProjection);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Differentiation kernels.
// ---------------------------------------------------------------------------------------------------------------
struct ElemParams {
    const double* rec;
    const double* x;
    const double* theta;
    const double* w;      // adjoint vector (num_dofs) for the VJP
    double* out;          // T x 2
    unsigned* status;
    int num_tets;
};

// dE/dtheta_e = w (L^2 / 2, (|F|^2 - 3) / 2 - L).
__global__ __launch_bounds__(256) void k_dE_dtheta(ElemParams p) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= p.num_tets)
        return;
    TetIn t;
    load_tet(p.rec, p.theta, e, t);
    double u[4][3];
    load_u(p.x, t.ids, u);
    Kin k;
    kinematics(u, t.Di, k);
    unsigned flags = 0;
    if (!kin_ok(k))
        flags |= kStatusDetF;
    if (!theta_ok(t.lam, t.mu))
        flags |= kStatusBadTheta;
    double sq = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q)
        sq = fma(k.F[q], k.F[q], sq);
    double2 r;
    r.x = t.w * (0.5 * k.L * k.L);
    r.y = t.w * (0.5 * (sq - 3.0) - k.L);
    if (!(isfinite(r.x) && isfinite(r.y)))
        flags |= kStatusNonFiniteOut;
    reinterpret_cast<double2*>(p.out)[e] = r;
    if (flags)
        atomicOr(p.status, flags);
}

// Force VJP in theta: d g_e[3a+i] / d lambda = w L q_a,i and d / d mu = w ((F g_a)_i - q_a,i)
// (dP/dlambda = L F^-T, dP/dmu = F - F^-T; neo_hookean.cpp:88-101).
__global__ __launch_bounds__(256) void k_force_vjp_theta(ElemParams p) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= p.num_tets)
        return;
    TetIn t;
    load_tet(p.rec, p.theta, e, t);
    double u[4][3];
    load_u(p.x, t.ids, u);
    Kin k;
    kinematics(u, t.Di, k);
    unsigned flags = 0;
    if (!kin_ok(k))
        flags |= kStatusDetF;
    if (!theta_ok(t.lam, t.mu))
        flags |= kStatusBadTheta;
    double G[4][3];
    shape_gradients(t.Di, G);
    double sl = 0.0, sm = 0.0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const double* wb = p.w + 3 * static_cast<size_t>(t.ids[b]);
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
    double2 r;
    r.x = t.w * k.L * sl;
    r.y = t.w * sm;
    if (!(isfinite(r.x) && isfinite(r.y)))
        flags |= kStatusNonFiniteOut;
    reinterpret_cast<double2*>(p.out)[e] = r;
    if (flags)
        atomicOr(p.status, flags);
}

// Vertex-gather products: MODE 0 = Exact H v, MODE 1 = sum_e (d g_e / d theta_e) dtheta_e.
struct GatherParams {
    const double* rec;
    const int32_t* inc_ptr;
    const int32_t* inc_tet; // stride (bytes) given by inc_stride
    int inc_stride;
    const double* x;
    const double* theta;
    const double* v; // MODE 0: vector (num_dofs); MODE 1: dtheta (T x 2)
    double* out;     // num_dofs
    unsigned* status;
    int num_vertices;
};

template <int MODE> __global__ __launch_bounds__(256) void k_vertex_gather(GatherParams p) {
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= p.num_vertices)
        return;
    unsigned flags = 0;
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    const int kb = p.inc_ptr[v], ke = p.inc_ptr[v + 1];
    for (int kk = kb; kk < ke; ++kk) {
        const int e = *reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(p.inc_tet) +
                                                         static_cast<size_t>(kk) * p.inc_stride);
        TetIn t;
        load_tet(p.rec, p.theta, e, t);
        double u[4][3];
        load_u(p.x, t.ids, u);
        const int a = (t.ids[0] == v) ? 0 : (t.ids[1] == v) ? 1 : (t.ids[2] == v) ? 2 : 3;
        Kin k;
        kinematics(u, t.Di, k);
        if (!kin_ok(k))
            flags |= kStatusDetF;
        if (!theta_ok(t.lam, t.mu))
            flags |= kStatusBadTheta;
        double G[4][3];
        shape_gradients(t.Di, G);
        double ga[3];
        select_row(G, a, ga);
        double qa[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            qa[i] = dot3(k.a[3 * i], k.a[3 * i + 1], k.a[3 * i + 2], ga[0], ga[1], ga[2]);
        if (MODE == 0) {
            const double pr = fma(t.lam, k.L, -t.mu);
            // sum_b H_ab v_b = w sum_b [mu (g_a.g_b) v_b + lambda q_a (q_b.v_b) - p q_b (q_a.v_b)]
            double sv[3] = {0.0, 0.0, 0.0}, sq[3] = {0.0, 0.0, 0.0};
            double sqv = 0.0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double* vb = p.v + 3 * static_cast<size_t>(t.ids[b]);
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
            y0 = fma(t.w, fma(t.mu, sv[0], fma(t.lam * qa[0], sqv, -(pr * sq[0]))), y0);
            y1 = fma(t.w, fma(t.mu, sv[1], fma(t.lam * qa[1], sqv, -(pr * sq[1]))), y1);
            y2 = fma(t.w, fma(t.mu, sv[2], fma(t.lam * qa[2], sqv, -(pr * sq[2]))), y2);
        } else {
            const double2 dt = reinterpret_cast<const double2*>(p.v)[e];
            double Fg[3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                Fg[i] = dot3(k.F[3 * i], k.F[3 * i + 1], k.F[3 * i + 2], ga[0], ga[1], ga[2]);
            const double dl = dt.x * k.L;
            y0 = fma(t.w, fma(dl, qa[0], dt.y * (Fg[0] - qa[0])), y0);
            y1 = fma(t.w, fma(dl, qa[1], dt.y * (Fg[1] - qa[1])), y1);
            y2 = fma(t.w, fma(dl, qa[2], dt.y * (Fg[2] - qa[2])), y2);
        }
    }
    double* o = p.out + 3 * static_cast<size_t>(v);
    o[0] = y0;
    o[1] = y1;
    o[2] = y2;
    if (!(isfinite(y0) && isfinite(y1) && isfinite(y2)))
        flags |= kStatusNonFiniteOut;
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

template <int LPV, int WAVES, bool WANT_G, int KIND, bool WIDE>
void launch_assemble_t(const AssembleParams& p, int num_tiles, int lds_doubles, bool global_acc,
                       int persistent_grid, hipStream_t stream) {
    const dim3 block(kWave * WAVES);
    const size_t bytes = global_acc ? 0 : static_cast<size_t>(lds_doubles) * sizeof(double);
    if (persistent_grid > 0) {
        const dim3 grid(std::min(persistent_grid, num_tiles));
        if (global_acc) {
            hipLaunchKernelGGL((k_assemble_persistent<LPV, WAVES, WANT_G, KIND, WIDE, true>), grid, block, 0,
                               stream, p, num_tiles);
        } else {
            auto kern = k_assemble_persistent<LPV, WAVES, WANT_G, KIND, WIDE, false>;
            static bool configured = false;
            if (!configured) {
                allow_large_lds(kern, WAVES);
                configured = true;
            }
            hipLaunchKernelGGL(kern, grid, block, bytes, stream, p, num_tiles);
        }
    } else {
        const dim3 grid(num_tiles);
        if (global_acc) {
            hipLaunchKernelGGL((k_assemble<LPV, WAVES, WANT_G, KIND, WIDE, true>), grid, block, 0, stream, p);
        } else {
            auto kern = k_assemble<LPV, WAVES, WANT_G, KIND, WIDE, false>;
            static bool configured = false;
            if (!configured) {
                allow_large_lds(kern, WAVES);
                configured = true;
            }
            hipLaunchKernelGGL(kern, grid, block, bytes, stream, p);
        }
    }
    check_launch("k_assemble");
}

template <int LPV, int WAVES>
void dispatch_kind(const AssembleParams& p, int want_g, int kind, bool wide, int num_tiles, int lds, bool gacc,
                   int persistent_grid, hipStream_t s) {
#define LPG_CASE(G, K, W)                                                                                            \
    if (!!want_g == G && kind == K && wide == W)                                                                      \
        return launch_assemble_t<LPV, WAVES, G, K, W>(p, num_tiles, lds, gacc, persistent_grid, s);
    LPG_CASE(false, -1, false)
    LPG_CASE(true, -1, false)
    LPG_CASE(false, 0, false)
    LPG_CASE(true, 0, false)
    LPG_CASE(false, 1, false)
    LPG_CASE(true, 1, false)
    LPG_CASE(false, -1, true)
    LPG_CASE(true, -1, true)
    LPG_CASE(false, 0, true)
    LPG_CASE(true, 0, true)
    LPG_CASE(false, 1, true)
    LPG_CASE(true, 1, true)
#undef LPG_CASE
    fail(kLogic, "k_assemble: unsupported configuration");
}

} // namespace

void launch_clear(unsigned* status, unsigned* counter, hipStream_t stream) {
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, stream, status, counter);
    check_launch("k_clear");
}

void launch_assemble(const DeviceMesh& m, const EvalArgs& a, const LaunchConfig& lc, const int32_t* tile_vbeg,
                     const int32_t* tile_ids, int num_tiles, int lds_doubles, bool global_acc, int persistent_grid,
                     hipStream_t stream) {
    if (num_tiles <= 0)
        return;
    AssembleParams p{};
    p.rec = m.rec;
    p.inc_ptr = m.inc_ptr;
    p.inc = m.inc;
    p.row_ptr = m.row_ptr;
    p.tile_vbeg = tile_vbeg;
    p.tile_ids = tile_ids;
    p.x = a.x;
    p.theta = a.theta;
    p.g = a.g;
    p.H = a.H;
    p.vertex_E = a.vertex_E;
    p.status = a.status;
    p.proj_slot = a.proj_slot;
    p.proj_M = a.proj_M;
#define LPG_CFG(L, W)                                                                                                \
    if (lc.lanes_per_vertex == L && lc.waves == W)                                                                    \
        return dispatch_kind<L, W>(p, a.want_g, a.kind, m.wide, num_tiles, lds_doubles, global_acc,           \
                                   persistent_grid, stream);
    LPG_CFG(1, 1)
    LPG_CFG(2, 1)
    LPG_CFG(4, 1)
    LPG_CFG(1, 2)
    LPG_CFG(2, 2)
    LPG_CFG(4, 2)
    LPG_CFG(1, 4)
    LPG_CFG(2, 4)
    LPG_CFG(4, 4)
#undef LPG_CFG
    fail(kInvalid, "launch_assemble: unsupported launch configuration");
}

void launch_energy_reduce(const double* vertex_E, const int32_t* chunk_begin, int num_chunks, double* chunk_sum,
                          const int32_t* seg_chunk, int num_segments, double* out, unsigned* status,
                          hipStream_t stream) {
    if (num_chunks > 0) {
        hipLaunchKernelGGL(k_energy_chunks, dim3(num_chunks), dim3(256), 0, stream, vertex_E, chunk_begin, chunk_sum);
        check_launch("k_energy_chunks");
    }
    if (num_segments > 0) {
        hipLaunchKernelGGL(k_energy_segments, dim3(num_segments), dim3(256), 0, stream, chunk_sum, seg_chunk, out,
                           status);
        check_launch("k_energy_segments");
    }
}

void launch_project_prepass(const DeviceMesh& m, const EvalArgs& a, hipStream_t stream) {
    if (m.num_tets == 0)
        return;
    ProjectParams p{m.rec, a.x, a.theta, m.num_tets, a.status, a.counter, a.proj_slot, a.proj_list, a.proj_M};
    hipLaunchKernelGGL(k_project_prepass, dim3((m.num_tets + 255) / 256), dim3(256), 0, stream, p);
    check_launch("k_project_prepass");
}

void launch_project_jacobi(const DeviceMesh& m, const EvalArgs& a, hipStream_t stream) {
    if (m.num_tets == 0)
        return;
    ProjectParams p{m.rec, a.x, a.theta, m.num_tets, a.status, a.counter, a.proj_slot, a.proj_list, a.proj_M};
    const int blocks = std::min(4096, (m.num_tets + 63) / 64);
    hipLaunchKernelGGL(k_project_jacobi, dim3(blocks), dim3(64), 0, stream, p);
    check_launch("k_project_jacobi");
}

void launch_dE_dtheta(const DeviceMesh& m, const double* x, const double* theta, double* out, unsigned* status,
                      hipStream_t stream) {
    if (m.num_tets == 0)
        return;
    ElemParams p{m.rec, x, theta, nullptr, out, status, m.num_tets};
    hipLaunchKernelGGL(k_dE_dtheta, dim3((m.num_tets + 255) / 256), dim3(256), 0, stream, p);
    check_launch("k_dE_dtheta");
}

void launch_force_vjp_theta(const DeviceMesh& m, const double* x, const double* theta, const double* w,
                            double* out, unsigned* status, hipStream_t stream) {
    if (m.num_tets == 0)
        return;
    ElemParams p{m.rec, x, theta, w, out, status, m.num_tets};
    hipLaunchKernelGGL(k_force_vjp_theta, dim3((m.num_tets + 255) / 256), dim3(256), 0, stream, p);
    check_launch("k_force_vjp_theta");
}

static void launch_gather(int mode, const DeviceMesh& m, const double* x, const double* theta, const double* v,
                          double* out, unsigned* status, hipStream_t stream) {
    if (m.num_vertices == 0)
        return;
    GatherParams p{m.rec,
                   m.inc_ptr,
                   static_cast<const int32_t*>(m.inc),
                   m.wide ? static_cast<int>(sizeof(IncWide)) : static_cast<int>(sizeof(IncNarrow)),
                   x,
                   theta,
                   v,
                   out,
                   status,
                   m.num_vertices};
    const dim3 grid((m.num_vertices + 255) / 256), block(256);
    if (mode == 0)
        hipLaunchKernelGGL(k_vertex_gather<0>, grid, block, 0, stream, p);
    else
        hipLaunchKernelGGL(k_vertex_gather<1>, grid, block, 0, stream, p);
    check_launch("k_vertex_gather");
}

void launch_force_jvp_theta(const DeviceMesh& m, const double* x, const double* theta, const double* dtheta,
                            double* out, unsigned* status, hipStream_t stream) {
    launch_gather(1, m, x, theta, dtheta, out, status, stream);
}

void launch_hvp(const DeviceMesh& m, const double* x, const double* theta, const double* v, double* out,
                unsigned* status, hipStream_t stream) {
    launch_gather(0, m, x, theta, v, out, status, stream);
}

} // namespace lpg
