// Device kernels of the equilibrium solver: the Hessian-vector product on the vertex-ordered CSR values, fixed-order
// dot products, and preconditioned CG fused into two kernels per iteration with a 3x3 block-Jacobi preconditioner.
// Every reduction has a fixed order, so a solve is reproducible bitwise run to run.
#include "lpg_solver.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace lpg {
namespace {

constexpr int kWave = 64;
constexpr int kLanesPerVertex = 16; // SpMV: one 16-lane group per vertex (3 rows, 3 deg(v) columns each)
constexpr int kDotBlocks = 1024;    // partial sums of a dot product (fixed, so the summation order is fixed)

// y = F (H p) + shift F p, with F the free-DOF mask (fixed rows are zero).  p must vanish on fixed DOFs, so the
// product is the free block of H applied to p's free part.  Rows 3v..3v+2 hold, for each neighbour w of v (sorted)
// and component k, the value at column 3w+k: the row i entries of v start at 9 nbr_ptr[v] + 3 deg(v) i.
__global__ __launch_bounds__(256) void k_spmv(const double* __restrict__ H, const int32_t* __restrict__ nbr_ptr,
                                              const int32_t* __restrict__ nbr, const uint8_t* __restrict__ fixed,
                                              const double* __restrict__ p, const double* __restrict__ shift,
                                              double* __restrict__ y, int num_vertices) {
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int v = gid / kLanesPerVertex;
    const int sub = gid % kLanesPerVertex;
    if (v >= num_vertices)
        return; // whole 16-lane groups leave together
    const int nb = nbr_ptr[v], deg = nbr_ptr[v + 1] - nb;
    const int cols = 3 * deg;
    const double* seg = H + 9 * static_cast<size_t>(nb);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int t = sub; t < cols; t += kLanesPerVertex) {
        const int w = nbr[nb + t / 3];
        const double pv = p[3 * static_cast<size_t>(w) + t % 3];
        a0 = fma(seg[t], pv, a0);
        a1 = fma(seg[cols + t], pv, a1);
        a2 = fma(seg[2 * cols + t], pv, a2);
    }
#pragma unroll
    for (int m = kLanesPerVertex / 2; m > 0; m >>= 1) {
        a0 += __shfl_xor(a0, m, kLanesPerVertex);
        a1 += __shfl_xor(a1, m, kLanesPerVertex);
        a2 += __shfl_xor(a2, m, kLanesPerVertex);
    }
    if (sub == 0) {
        const double s = shift ? *shift : 0.0;
        const size_t r = 3 * static_cast<size_t>(v);
        y[r] = fixed[r] ? 0.0 : fma(s, p[r], a0);
        y[r + 1] = fixed[r + 1] ? 0.0 : fma(s, p[r + 1], a1);
        y[r + 2] = fixed[r + 2] ? 0.0 : fma(s, p[r + 2], a2);
    }
}

// Partial sums of a.b (and of c.d when c is given) over a fixed grid-stride split.
__global__ __launch_bounds__(256) void k_dot_partial(const double* __restrict__ a, const double* __restrict__ b,
                                                     const double* __restrict__ c, const double* __restrict__ d,
                                                     int n, double* __restrict__ part) {
    __shared__ double red[2][256];
    double s = 0.0, t = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += 256 * gridDim.x) {
        s = fma(a[i], b[i], s);
        if (c)
            t = fma(c[i], d[i], t);
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = red[0][0];
        part[gridDim.x + blockIdx.x] = red[1][0];
    }
}

__global__ __launch_bounds__(256) void k_dot_final(const double* __restrict__ part, int nb, double* __restrict__ out,
                                                   int two) {
    __shared__ double red[2][256];
    double s = 0.0, t = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) {
        s += part[i];
        if (two)
            t += part[nb + i];
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0][0];
        if (two)
            out[1] = red[1][0];
    }
}

// Maximum |a_i| over free DOFs (non-negative doubles order as integers), one atomic per workgroup: atomics on a
// single address serialize, so the waves' maxima are combined in LDS first.
__global__ __launch_bounds__(256) void k_free_maxabs(const double* __restrict__ a, const uint8_t* __restrict__ fixed,
                                                     int n, unsigned long long* out_bits) {
    __shared__ double wave_max[256 / kWave];
    double m = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += 256 * gridDim.x)
        if (!fixed[i])
            m = fmax(m, fabs(a[i]));
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
        m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x % kWave) == 0)
        wave_max[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 1; w < 256 / kWave; ++w)
            m = fmax(m, wave_max[w]);
        atomicMax(out_bits, static_cast<unsigned long long>(__double_as_longlong(m)));
    }
}

__global__ void k_set_bits(unsigned long long* p, double v) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        p[0] = static_cast<unsigned long long>(__double_as_longlong(v));
}

// y = x + alpha v (alpha from the host).
__global__ __launch_bounds__(256) void k_axpy(const double* __restrict__ x, double alpha, const double* __restrict__ v,
                                              double* __restrict__ y, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        y[i] = fma(alpha, v[i], x[i]);
}

// y = a x + b w on free DOFs, 0 on fixed ones.
__global__ __launch_bounds__(256) void k_free_combine(const double* __restrict__ x, double a,
                                                      const double* __restrict__ w, double b,
                                                      const uint8_t* __restrict__ fixed, double* __restrict__ y,
                                                      int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        y[i] = fixed[i] ? 0.0 : fma(b, w ? w[i] : 0.0, a * (x ? x[i] : 0.0));
}

// diag[i] = H_ii on free DOFs, 0 on fixed ones.
__global__ __launch_bounds__(256) void k_diag(const double* __restrict__ H, const int32_t* __restrict__ nbr_ptr,
                                              const int32_t* __restrict__ self_pos, const uint8_t* __restrict__ fixed,
                                              double* __restrict__ diag, int num_vertices) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 3 * num_vertices)
        return;
    const int v = i / 3, c = i % 3;
    const int nb = nbr_ptr[v], deg = nbr_ptr[v + 1] - nb;
    diag[i] = fixed[i] ? 0.0 : H[9 * static_cast<size_t>(nb) + 3 * deg * c + 3 * self_pos[v] + c];
}

__global__ __launch_bounds__(256) void k_scale(double* __restrict__ a, double f, long long n) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += 256ll * gridDim.x)
        a[i] *= f;
}

// ---------------------------------------------------------------------------------------------------------------
// Fused preconditioned CG: two kernels per iteration.
//   k_pcg_apply:  p = z + beta p_old (beta = r.z / previous r.z), q = F (H p) + shift F p, and partials of p.q;
//                 the neighbours' p values are formed on the fly from z and p_old, so p needs no separate pass.
//   k_pcg_update: alpha = r.z / p.q, d += alpha p, r -= alpha q, z = B^-1 r (3x3 block-Jacobi), partials of r.z
//                 and r.r.
// Each kernel writes one partial sum per workgroup; every workgroup of the next kernel adds the previous kernel's
// partials itself, in index order with a fixed tree, so all of them hold the same bits and no grid-wide
// synchronisation (device-scope fences, which write back and invalidate the per-XCD L2 caches) is needed.  Device
// scalars: [kRz0, kRz0 + 1] r.z of the last two iterations (a ring indexed by parity), kRr r.r (written by
// k_pcg_finish), kFlag breakdown (p.q not positive: the system is not positive definite on the free DOFs).
// ---------------------------------------------------------------------------------------------------------------
constexpr int kRz0 = 0, kRr = 3, kFlag = 4;
constexpr int kThreads = 256;
constexpr int kApplyBlocks = 2048;  // workgroups of k_pcg_apply (at most): p.q partials
constexpr int kUpdateBlocks = 256;  // workgroups of k_pcg_begin / k_pcg_update (at most): r.z, r.r partials

// Workgroup sum with the same bits in every thread: an xor butterfly within each wave (partners add the same two
// values, so all 64 lanes end with identical totals), then the wave totals in wave order.  Two barriers instead of a
// shared-memory tree's nine.
template <int K> __device__ __forceinline__ void block_reduce(double (&v)[K], double (*red)[kThreads / kWave]) {
    const int tid = static_cast<int>(threadIdx.x);
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1)
            v[k] += __shfl_xor(v[k], off);
    __syncthreads(); // earlier readers of red are done
    if (tid % kWave == 0)
#pragma unroll
        for (int k = 0; k < K; ++k)
            red[k][tid / kWave] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double t = red[k][0];
#pragma unroll
        for (int w = 1; w < kThreads / kWave; ++w)
            t += red[k][w];
        v[k] = t;
    }
}

// This workgroup's sums of K per-thread values into part[k G + block] (G = the grid size).
template <int K> __device__ __forceinline__ void block_partials(double (&v)[K], double* __restrict__ part) {
    __shared__ double red[K][kThreads / kWave];
    block_reduce<K>(v, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k)
            part[k * gridDim.x + blockIdx.x] = v[k];
}

// The total of n partials, the same bits in every thread and workgroup (strided per-thread sums, fixed tree).
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int n) {
    __shared__ double red[1][kThreads / kWave];
    double v[1] = {0.0};
    for (int i = static_cast<int>(threadIdx.x); i < n; i += kThreads)
        v[0] += part[i];
    block_reduce<1>(v, red);
    return v[0];
}

// z_v = B_v^-1 r_v with the 6 stored entries (00, 01, 02, 11, 12, 22) of the symmetric inverse.
__device__ __forceinline__ void block_solve(const double* __restrict__ bi, const double r[3], double z[3]) {
    z[0] = fma(bi[0], r[0], fma(bi[1], r[1], bi[2] * r[2]));
    z[1] = fma(bi[1], r[0], fma(bi[3], r[1], bi[4] * r[2]));
    z[2] = fma(bi[2], r[0], fma(bi[4], r[1], bi[5] * r[2]));
}

// Block-Jacobi preconditioner: per vertex the inverse of its 3x3 diagonal block restricted to the free components
// (plus the shift), by Cholesky; fixed components get zero rows and columns.  A block that is not positive definite
// falls back to the inverted positive free diagonal entries (0 elsewhere), as scalar Jacobi.
__global__ __launch_bounds__(kThreads) void k_block_jacobi(const double* __restrict__ H,
                                                           const int32_t* __restrict__ nbr_ptr,
                                                           const int32_t* __restrict__ self_pos,
                                                           const uint8_t* __restrict__ fixed,
                                                           const double* __restrict__ shift,
                                                           double* __restrict__ binv, int num_vertices) {
    const int v = blockIdx.x * kThreads + threadIdx.x;
    if (v >= num_vertices)
        return;
    const int nb = nbr_ptr[v], deg = nbr_ptr[v + 1] - nb;
    const double* blk = H + 9 * static_cast<size_t>(nb) + 3 * self_pos[v];
    const double sh = shift ? *shift : 0.0;
    bool fr[3];
    double a[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        fr[i] = !fixed[3 * static_cast<size_t>(v) + i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double h = 0.5 * (blk[3 * deg * i + k] + blk[3 * deg * k + i]);
            a[i][k] = (fr[i] && fr[k]) ? h + (i == k ? sh : 0.0) : (i == k ? 1.0 : 0.0);
        }
    // Cholesky a = L L^T, then a^-1 = L^-T L^-1.
    const double l00 = a[0][0] > 0.0 ? sqrt(a[0][0]) : 0.0;
    const double l10 = l00 > 0.0 ? a[1][0] / l00 : 0.0, l20 = l00 > 0.0 ? a[2][0] / l00 : 0.0;
    const double d1 = a[1][1] - l10 * l10;
    const double l11 = d1 > 0.0 ? sqrt(d1) : 0.0;
    const double l21 = l11 > 0.0 ? (a[2][1] - l20 * l10) / l11 : 0.0;
    const double d2 = a[2][2] - l20 * l20 - l21 * l21;
    const double l22 = d2 > 0.0 ? sqrt(d2) : 0.0;
    double out[6];
    const bool pd = l00 > 0.0 && l11 > 0.0 && l22 > 0.0 && isfinite(l00 * l11 * l22);
    if (pd) {
        // L^-1 (lower): m00 = 1/l00, m11 = 1/l11, m22 = 1/l22, m10 = -l10 m00 m11, m21 = -l21 m11 m22,
        // m20 = -(l20 m00 + l21 m10) m22.
        const double m00 = 1.0 / l00, m11 = 1.0 / l11, m22 = 1.0 / l22;
        const double m10 = -l10 * m00 * m11, m21 = -l21 * m11 * m22, m20 = -(l20 * m00 + l21 * m10) * m22;
        out[0] = m00 * m00 + m10 * m10 + m20 * m20;
        out[1] = m11 * m10 + m21 * m20;
        out[2] = m22 * m20;
        out[3] = m11 * m11 + m21 * m21;
        out[4] = m22 * m21;
        out[5] = m22 * m22;
    } else {
        out[1] = out[2] = out[4] = 0.0;
        out[0] = a[0][0] > 0.0 ? 1.0 / a[0][0] : 0.0;
        out[3] = a[1][1] > 0.0 ? 1.0 / a[1][1] : 0.0;
        out[5] = a[2][2] > 0.0 ? 1.0 / a[2][2] : 0.0;
    }
    // Zero rows and columns of fixed components.
    const bool f0 = fr[0], f1 = fr[1], f2 = fr[2];
    out[0] = f0 ? out[0] : 0.0;
    out[1] = (f0 && f1) ? out[1] : 0.0;
    out[2] = (f0 && f2) ? out[2] : 0.0;
    out[3] = f1 ? out[3] : 0.0;
    out[4] = (f1 && f2) ? out[4] : 0.0;
    out[5] = f2 ? out[5] : 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q)
        binv[6 * static_cast<size_t>(v) + q] = out[q];
}

// PCG start: r = -F g, d = 0, z = B^-1 r; partials of r.z and r.r into part (row 0, row 1); s[kFlag] = 0.
__global__ __launch_bounds__(kThreads) void k_pcg_begin(const double* __restrict__ g, const uint8_t* __restrict__ fixed,
                                                        const double* __restrict__ binv, double* __restrict__ r,
                                                        double* __restrict__ d, double* __restrict__ z,
                                                        double* __restrict__ s, double* __restrict__ part,
                                                        int num_vertices) {
    double acc[2] = {0.0, 0.0};
    for (int v = blockIdx.x * kThreads + threadIdx.x; v < num_vertices; v += kThreads * gridDim.x) {
        const size_t o = 3 * static_cast<size_t>(v);
        double rv[3], zv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rv[c] = fixed[o + c] ? 0.0 : -g[o + c];
            r[o + c] = rv[c];
            d[o + c] = 0.0;
        }
        block_solve(binv + 6 * static_cast<size_t>(v), rv, zv);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            z[o + c] = zv[c];
            acc[0] = fma(rv[c], zv[c], acc[0]);
            acc[1] = fma(rv[c], rv[c], acc[1]);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        s[kFlag] = 0.0;
    block_partials<2>(acc, part);
}

// s[kRr] = r.r from the partials of the last k_pcg_begin / k_pcg_update (`nu` workgroups).  With rb (pinned host
// memory), also rb[0] = r.r, rb[1] = the breakdown flag and then, after a system-scope fence, the 32-bit word at
// rb + 2 = seq: a host that sees seq there reads both values without a copy or a stream synchronisation.
__global__ __launch_bounds__(kThreads) void k_pcg_finish(const double* __restrict__ part_u, int nu,
                                                         double* __restrict__ s, double* rb, unsigned seq) {
    const double rr = sum_partials(part_u + nu, nu);
    if (threadIdx.x == 0) {
        s[kRr] = rr;
        if (rb) {
            rb[0] = rr;
            rb[1] = s[kFlag];
            __threadfence_system();
            __hip_atomic_store(reinterpret_cast<unsigned*>(rb + 2), seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Iteration `it` (from 1): r.z of the current residual from the `nu` partials of part_u (stored by workgroup 0 in
// ring slot it % 2), beta = that / s[ring (it - 1) % 2] (it > 1; p = z for it = 1), p = z + beta p_old and
// q = F (H p) + shift F p on the vertex-ordered CSR (16 lanes per vertex), partials of p.q into part_a.
template <bool FIRST>
__global__ __launch_bounds__(kThreads) void k_pcg_apply(const double* __restrict__ H,
                                                        const int32_t* __restrict__ nbr_ptr,
                                                        const int32_t* __restrict__ nbr,
                                                        const uint8_t* __restrict__ fixed,
                                                        const double* __restrict__ z,
                                                        const double* __restrict__ p_old, double* __restrict__ p,
                                                        double* __restrict__ q, const double* __restrict__ shift,
                                                        double* __restrict__ s, int it, const double* __restrict__ part_u,
                                                        int nu, double* __restrict__ part_a, int num_vertices) {
    // A group's vertices v, v + G, ... (G groups).  Each vertex's first 3 kLanesPerVertex columns are staged in
    // registers (its row pointers, then the neighbour ids and values, then z and p_old at those columns); the first
    // vertex's loads are issued before the r.z partials are summed, which only the combination p = z + beta p_old
    // waits for.  Wider rows (more than 16 neighbours) take their remaining columns in a plain loop.
    constexpr int kStaged = 3;
    const int sub = static_cast<int>(threadIdx.x) % kLanesPerVertex;
    const int groups = kThreads / kLanesPerVertex * static_cast<int>(gridDim.x);
    const int first = (blockIdx.x * kThreads + threadIdx.x) / kLanesPerVertex;
    int nb = 0, cols = 0;
    int col_j[kStaged];
    double hv[kStaged][3], zv[kStaged], pov[kStaged], zs[3], ps[3];
    const auto load = [&](int v) {
        nb = nbr_ptr[v];
        cols = 3 * (nbr_ptr[v + 1] - nb);
        const size_t o = 3 * static_cast<size_t>(v);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            zs[c] = z[o + c];
            ps[c] = FIRST ? 0.0 : p_old[o + c];
        }
        const double* seg = H + 9 * static_cast<size_t>(nb);
#pragma unroll
        for (int m = 0; m < kStaged; ++m) {
            col_j[m] = 0;
            hv[m][0] = hv[m][1] = hv[m][2] = 0.0;
            if (cols > 0) { // a vertex in no tet has an empty row
                int t = sub + kLanesPerVertex * m;
                t = t < cols ? t : cols - 1; // lanes past the row's end reload its last column, unused
                col_j[m] = 3 * nbr[nb + t / 3] + t % 3;
                hv[m][0] = seg[t];
                hv[m][1] = seg[cols + t];
                hv[m][2] = seg[2 * cols + t];
            }
        }
#pragma unroll
        for (int m = 0; m < kStaged; ++m) {
            zv[m] = z[col_j[m]];
            pov[m] = FIRST ? 0.0 : p_old[col_j[m]];
        }
    };
    if (first < num_vertices)
        load(first);
    const double rz = sum_partials(part_u, nu);
    double beta = 0.0;
    if (!FIRST) {
        const double rz_prev = s[kRz0 + ((it - 1) & 1)];
        beta = rz_prev != 0.0 ? rz / rz_prev : 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        s[kRz0 + (it & 1)] = rz;
    const double sh = shift ? *shift : 0.0;
    double acc[1] = {0.0};
    for (int v = first; v < num_vertices; v += groups) {
        if (v != first)
            load(v);
        // Columns t = sub, sub + 16, ... in increasing order per lane, then the 16 lanes' sums by a fixed tree.
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
        for (int m = 0; m < kStaged; ++m) {
            const bool in = sub + kLanesPerVertex * m < cols;
            const double pv = FIRST ? zv[m] : fma(beta, pov[m], zv[m]);
            a0 = in ? fma(hv[m][0], pv, a0) : a0;
            a1 = in ? fma(hv[m][1], pv, a1) : a1;
            a2 = in ? fma(hv[m][2], pv, a2) : a2;
        }
        if (cols > kLanesPerVertex * kStaged) {
            const double* seg = H + 9 * static_cast<size_t>(nb);
            for (int t = sub + kLanesPerVertex * kStaged; t < cols; t += kLanesPerVertex) {
                const size_t j = 3 * static_cast<size_t>(nbr[nb + t / 3]) + t % 3;
                const double pv = FIRST ? z[j] : fma(beta, p_old[j], z[j]);
                a0 = fma(seg[t], pv, a0);
                a1 = fma(seg[cols + t], pv, a1);
                a2 = fma(seg[2 * cols + t], pv, a2);
            }
        }
#pragma unroll
        for (int m = kLanesPerVertex / 2; m > 0; m >>= 1) {
            a0 += __shfl_xor(a0, m, kLanesPerVertex);
            a1 += __shfl_xor(a1, m, kLanesPerVertex);
            a2 += __shfl_xor(a2, m, kLanesPerVertex);
        }
        if (sub == 0) {
            const size_t o = 3 * static_cast<size_t>(v);
            const double av[3] = {a0, a1, a2};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double pv = FIRST ? zs[c] : fma(beta, ps[c], zs[c]);
                const double qv = fixed[o + c] ? 0.0 : fma(sh, pv, av[c]);
                p[o + c] = pv;
                q[o + c] = qv;
                acc[0] = fma(pv, qv, acc[0]);
            }
        }
    }
    block_partials<1>(acc, part_a);
}

// Iteration `it`: p.q from the `na` partials of part_a, alpha = r.z / p.q with r.z = s[ring it % 2] (0 and the
// breakdown flag when p.q is not positive), d += alpha p, r -= alpha q, z = B^-1 r; partials of the new r.z and
// r.r into part_u.
__global__ __launch_bounds__(kThreads) void k_pcg_update(const double* __restrict__ p, const double* __restrict__ q,
                                                         double* __restrict__ d, double* __restrict__ r,
                                                         double* __restrict__ z, const double* __restrict__ binv,
                                                         double* __restrict__ s, int it,
                                                         const double* __restrict__ part_a, int na,
                                                         double* __restrict__ part_u, int num_vertices) {
    const double pq = sum_partials(part_a, na);
    const bool ok = pq > 0.0 && isfinite(pq);
    const double alpha = ok ? s[kRz0 + (it & 1)] / pq : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0 && !ok)
        s[kFlag] = 1.0;
    double acc[2] = {0.0, 0.0};
    for (int v = blockIdx.x * kThreads + threadIdx.x; v < num_vertices; v += kThreads * gridDim.x) {
        const size_t o = 3 * static_cast<size_t>(v);
        double rv[3], zv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            d[o + c] = fma(alpha, p[o + c], d[o + c]);
            rv[c] = fma(-alpha, q[o + c], r[o + c]);
            r[o + c] = rv[c];
        }
        block_solve(binv + 6 * static_cast<size_t>(v), rv, zv);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            z[o + c] = zv[c];
            acc[0] = fma(rv[c], zv[c], acc[0]);
            acc[1] = fma(rv[c], rv[c], acc[1]);
        }
    }
    block_partials<2>(acc, part_u);
}

unsigned update_blocks(int num_vertices) {
    return static_cast<unsigned>(std::max(1, std::min(kUpdateBlocks, (num_vertices + kThreads - 1) / kThreads)));
}

unsigned apply_blocks(int num_vertices) {
    const long long threads = static_cast<long long>(num_vertices) * kLanesPerVertex;
    return static_cast<unsigned>(std::max(1ll, std::min<long long>(kApplyBlocks, (threads + kThreads - 1) / kThreads)));
}

unsigned blocks_for(int n) {
    return static_cast<unsigned>(std::max(1, (n + 255) / 256));
}

} // namespace

void check_solver_launch(const char* what) {
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess)
        fail(kDevice, std::string(what) + ": " + hipGetErrorString(err));
}

void launch_spmv(const DeviceMesh& m, const double* H, const uint8_t* fixed, const double* p, const double* shift,
                 double* y, hipStream_t s) {
    if (m.num_vertices == 0)
        return;
    const long long threads = static_cast<long long>(m.num_vertices) * kLanesPerVertex;
    hipLaunchKernelGGL(k_spmv, dim3(static_cast<unsigned>((threads + 255) / 256)), dim3(256), 0, s, H, m.nbr_ptr,
                       m.nbr, fixed, p, shift, y, m.num_vertices);
    check_solver_launch("k_spmv");
}

void launch_diag(const DeviceMesh& m, const double* H, const uint8_t* fixed, double* diag, hipStream_t s) {
    if (m.num_vertices == 0)
        return;
    hipLaunchKernelGGL(k_diag, dim3(blocks_for(3 * m.num_vertices)), dim3(256), 0, s, H, m.nbr_ptr, m.self_pos, fixed,
                       diag, m.num_vertices);
    check_solver_launch("k_diag");
}

void launch_scale(double* a, double f, int64_t n, hipStream_t s) {
    if (n == 0)
        return;
    const unsigned blocks = static_cast<unsigned>(std::min<int64_t>(8192, (n + 255) / 256));
    hipLaunchKernelGGL(k_scale, dim3(blocks), dim3(256), 0, s, a, f, static_cast<long long>(n));
    check_solver_launch("k_scale");
}

void launch_dot(const double* a, const double* b, const double* c, const double* d, int n, double* part,
                double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(256), 0, s, a, b, c, d, n, part);
    check_solver_launch("k_dot_partial");
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, part, kDotBlocks, out, c ? 1 : 0);
    check_solver_launch("k_dot_final");
}

int dot_partials() {
    return 2 * kDotBlocks;
}

void launch_free_maxabs(const double* a, const uint8_t* fixed, int n, double* out, hipStream_t s) {
    auto* bits = reinterpret_cast<unsigned long long*>(out);
    hipLaunchKernelGGL(k_set_bits, dim3(1), dim3(64), 0, s, bits, 0.0);
    check_solver_launch("k_set_bits");
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_free_maxabs, dim3(std::min(256u, blocks_for(n))), dim3(256), 0, s, a, fixed, n, bits);
    check_solver_launch("k_free_maxabs");
}

void launch_block_jacobi(const DeviceMesh& m, const double* H, const uint8_t* fixed, const double* shift, double* binv,
                         hipStream_t s) {
    if (m.num_vertices == 0)
        return;
    hipLaunchKernelGGL(k_block_jacobi, dim3(blocks_for(m.num_vertices)), dim3(kThreads), 0, s, H, m.nbr_ptr,
                       m.self_pos, fixed, shift, binv, m.num_vertices);
    check_solver_launch("k_block_jacobi");
}

void launch_pcg_begin(const DeviceMesh& m, const double* g, const uint8_t* fixed, const double* binv, double* r,
                      double* d, double* z, double* scalars, double* part, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_begin, dim3(update_blocks(m.num_vertices)), dim3(kThreads), 0, s, g, fixed, binv, r, d, z,
                       scalars, part + kApplyBlocks, m.num_vertices);
    check_solver_launch("k_pcg_begin");
}

void launch_pcg_finish(const DeviceMesh& m, double* scalars, const double* part, hipStream_t s, double* rb,
                       unsigned seq) {
    hipLaunchKernelGGL(k_pcg_finish, dim3(1), dim3(kThreads), 0, s, part + kApplyBlocks,
                       static_cast<int>(update_blocks(m.num_vertices)), scalars, rb, seq);
    check_solver_launch("k_pcg_finish");
}

void launch_pcg_apply(const DeviceMesh& m, const double* H, const uint8_t* fixed, const double* z, const double* p_old,
                      double* p, double* q, const double* shift, int it, double* scalars, double* part,
                      hipStream_t s) {
    const dim3 grid(apply_blocks(m.num_vertices)), block(kThreads);
    const int nu = static_cast<int>(update_blocks(m.num_vertices));
    if (it == 1)
        hipLaunchKernelGGL(k_pcg_apply<true>, grid, block, 0, s, H, m.nbr_ptr, m.nbr, fixed, z, p_old, p, q, shift,
                           scalars, it, part + kApplyBlocks, nu, part, m.num_vertices);
    else
        hipLaunchKernelGGL(k_pcg_apply<false>, grid, block, 0, s, H, m.nbr_ptr, m.nbr, fixed, z, p_old, p, q, shift,
                           scalars, it, part + kApplyBlocks, nu, part, m.num_vertices);
    check_solver_launch("k_pcg_apply");
}

void launch_pcg_update(const DeviceMesh& m, const double* p, const double* q, double* d, double* r, double* z,
                       const double* binv, int it, double* scalars, double* part, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_update, dim3(update_blocks(m.num_vertices)), dim3(kThreads), 0, s, p, q, d, r, z, binv,
                       scalars, it, part, static_cast<int>(apply_blocks(m.num_vertices)), part + kApplyBlocks,
                       m.num_vertices);
    check_solver_launch("k_pcg_update");
}

int pcg_partials() {
    return kApplyBlocks + 2 * kUpdateBlocks;
}

void launch_axpy(const double* x, double alpha, const double* v, double* y, int n, hipStream_t s) {
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_axpy, dim3(blocks_for(n)), dim3(256), 0, s, x, alpha, v, y, n);
    check_solver_launch("k_axpy");
}

void launch_free_combine(const double* x, double a, const double* w, double b, const uint8_t* fixed, double* y, int n,
                         hipStream_t s) {
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_free_combine, dim3(blocks_for(n)), dim3(256), 0, s, x, a, w, b, fixed, y, n);
    check_solver_launch("k_free_combine");
}

} // namespace lpg
