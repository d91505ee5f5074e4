// Device kernels of the equilibrium solver: the Hessian-vector product on the vertex-ordered CSR values, the
// Jacobi preconditioner, fixed-order dot products and the fused vector updates of preconditioned CG.  Every
// reduction has a fixed order, so a solve is reproducible bitwise run to run.
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

// Inverse of the free diagonal (plus the shift); 0 on fixed DOFs and where the diagonal is not positive.
__global__ __launch_bounds__(256) void k_jacobi(const double* __restrict__ H, const int32_t* __restrict__ nbr_ptr,
                                                const int32_t* __restrict__ self_pos, const uint8_t* __restrict__ fixed,
                                                const double* __restrict__ shift, double* __restrict__ minv,
                                                int num_vertices) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 3 * num_vertices)
        return;
    const int v = i / 3, c = i % 3;
    const int nb = nbr_ptr[v], deg = nbr_ptr[v + 1] - nb;
    const double dgl = H[9 * static_cast<size_t>(nb) + 3 * deg * c + 3 * self_pos[v] + c] + (shift ? *shift : 0.0);
    minv[i] = (fixed[i] || !(dgl > 0.0)) ? 0.0 : 1.0 / dgl;
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

// Maximum |a_i| over free DOFs (non-negative doubles order as integers), one atomic per wave.
__global__ __launch_bounds__(256) void k_free_maxabs(const double* __restrict__ a, const uint8_t* __restrict__ fixed,
                                                     int n, unsigned long long* out_bits) {
    double m = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += 256 * gridDim.x)
        if (!fixed[i])
            m = fmax(m, fabs(a[i]));
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
        m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x % kWave) == 0)
        atomicMax(out_bits, static_cast<unsigned long long>(__double_as_longlong(m)));
}

__global__ void k_set_bits(unsigned long long* p, double v) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        p[0] = static_cast<unsigned long long>(__double_as_longlong(v));
}

// PCG start: r = -F g, d = 0, z = M^-1 r, p = z.
__global__ __launch_bounds__(256) void k_pcg_start(const double* __restrict__ g, const uint8_t* __restrict__ fixed,
                                                   const double* __restrict__ minv, double* __restrict__ r,
                                                   double* __restrict__ d, double* __restrict__ z,
                                                   double* __restrict__ p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const double ri = fixed[i] ? 0.0 : -g[i];
    r[i] = ri;
    d[i] = 0.0;
    const double zi = minv[i] * ri;
    z[i] = zi;
    p[i] = zi;
}

// Scalars of a PCG iteration (device): s[0] = r.z, s[1] = p.q, s[2] = new r.z, s[3] = r.r, s[4] = breakdown flag.
// alpha = r.z / p.q; a non-positive p.q (an indefinite system) sets the flag and a zero step.
__global__ __launch_bounds__(256) void k_pcg_step(const double* __restrict__ q, double* __restrict__ r,
                                                  double* __restrict__ d, double* __restrict__ z,
                                                  const double* __restrict__ p, const double* __restrict__ minv,
                                                  double* __restrict__ s, int n) {
    const double pq = s[1];
    const bool ok = pq > 0.0 && isfinite(pq);
    const double alpha = ok ? s[0] / pq : 0.0;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0 && !ok)
        s[4] = 1.0;
    if (i >= n)
        return;
    d[i] = fma(alpha, p[i], d[i]);
    const double ri = fma(-alpha, q[i], r[i]);
    r[i] = ri;
    z[i] = minv[i] * ri;
}

// p = z + beta p with beta = (new r.z) / (r.z); then r.z <- new r.z (by workgroup 0's first lane after all reads of
// the old value: the kernel that reads s[0] next runs after this one).
__global__ __launch_bounds__(256) void k_pcg_direction(const double* __restrict__ z, double* __restrict__ p,
                                                       const double* __restrict__ s, int n) {
    const double beta = s[0] != 0.0 ? s[2] / s[0] : 0.0;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        p[i] = fma(beta, p[i], z[i]);
}

__global__ void k_pcg_roll(double* s) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        s[0] = s[2];
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

void launch_jacobi(const DeviceMesh& m, const double* H, const uint8_t* fixed, const double* shift, double* minv,
                   hipStream_t s) {
    if (m.num_vertices == 0)
        return;
    hipLaunchKernelGGL(k_jacobi, dim3(blocks_for(3 * m.num_vertices)), dim3(256), 0, s, H, m.nbr_ptr, m.self_pos,
                       fixed, shift, minv, m.num_vertices);
    check_solver_launch("k_jacobi");
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
    hipLaunchKernelGGL(k_free_maxabs, dim3(std::min(1024u, blocks_for(n))), dim3(256), 0, s, a, fixed, n, bits);
    check_solver_launch("k_free_maxabs");
}

void launch_pcg_start(const double* g, const uint8_t* fixed, const double* minv, double* r, double* d, double* z,
                      double* p, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_start, dim3(blocks_for(n)), dim3(256), 0, s, g, fixed, minv, r, d, z, p, n);
    check_solver_launch("k_pcg_start");
}

void launch_pcg_step(const double* q, double* r, double* d, double* z, const double* p, const double* minv,
                     double* scalars, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_step, dim3(blocks_for(n)), dim3(256), 0, s, q, r, d, z, p, minv, scalars, n);
    check_solver_launch("k_pcg_step");
}

void launch_pcg_direction(const double* z, double* p, double* scalars, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_direction, dim3(blocks_for(n)), dim3(256), 0, s, z, p, scalars, n);
    check_solver_launch("k_pcg_direction");
    hipLaunchKernelGGL(k_pcg_roll, dim3(1), dim3(64), 0, s, scalars);
    check_solver_launch("k_pcg_roll");
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
