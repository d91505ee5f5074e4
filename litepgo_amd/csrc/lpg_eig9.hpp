// Symmetric eigen-decomposition of a 9x9 matrix in registers, for the PSD projection of element Hessians.
//
// Method: Householder reduction to tridiagonal form with the transformations accumulated, then the implicit QL
// algorithm with Wilkinson-type shifts on the tridiagonal matrix, applying each plane rotation to the accumulated
// basis (the scheme of Bowdler, Martin, Reinsch and Wilkinson, "The QR and QL algorithms for symmetric matrices",
// Numer. Math. 11 (1968), written here from the mathematics).  Every loop has compile-time bounds and every array
// index is a compile-time constant, so the matrices stay in registers: data-dependent steps (the deflation point m
// of a QL sweep) are handled by predication instead of run-time indices.
//
// Cost per matrix: ~1.2k FP64 operations for the reduction and the basis, ~65 per plane rotation in the QL sweeps
// (8 - l rotations per sweep for eigenvalue l, 1-3 sweeps each), ~0.4k for V max(Lambda, 0) V^T.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LPG_EIG_FN __host__ __device__ __forceinline__
#else
// Plain C++ build (the CPU test of this routine).
#include <cmath>
#define LPG_EIG_FN inline
#endif

namespace lpg {
namespace eig9 {

constexpr int N = 9;

// True if the predicate holds on any lane of the wave (the single lane on the host).
LPG_EIG_FN bool any_lane(bool b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __any(b);
#else
    return b;
#endif
}

LPG_EIG_FN constexpr int sym(int r, int c) {
    return r <= c ? r * N - r * (r - 1) / 2 + (c - r) : c * N - c * (c - 1) / 2 + (r - c);
}

// a: packed upper triangle (45) of a symmetric 9x9 matrix, destroyed.  On return d[i] are the eigenvalues and
// column i of V the corresponding unit eigenvector; returns false if a QL sweep did not converge within the limit.
// rsqrt/recip are the caller's accurate reciprocal (square root) routines.
template <class Rsqrt, class Recip>
LPG_EIG_FN bool sym_eig(double a[45], double d[N], double V[N][N], double scale, Rsqrt rsqrt,
                                        Recip recip) {
    double e[N]; // e[i] couples d[i] and d[i+1]; e[N-1] = 0
    // ---- Householder reduction: step k zeroes row/column k beyond its first off-diagonal entry.  The reflector
    // H_k = I - beta v v^T acts on indices k+1..N-1; the basis V = H_0 H_1 ... accumulates them.
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c)
            V[r][c] = r == c ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < N - 2; ++k) {
        const int m = N - 1 - k; // length of x = a(k, k+1..N-1)
        double v[N];
        double sigma = 0.0;
#pragma unroll
        for (int i = 1; i < N; ++i)
            if (i < m) {
                v[i] = a[sym(k, k + 1 + i)];
                sigma = fma(v[i], v[i], sigma);
            }
        const double x0 = a[sym(k, k + 1)];
        const double nx = sigma > 0.0 ? fma(x0, x0, sigma) * rsqrt(fma(x0, x0, sigma)) : fabs(x0);
        const double alpha = x0 > 0.0 ? -nx : nx;
        v[0] = x0 - alpha;
        const double vtv = fma(v[0], v[0], sigma);
        // sigma == 0: the column is already reduced; beta = 0 makes H_k the identity.
        const double beta = sigma > 0.0 ? 2.0 * recip(vtv) : 0.0;
        e[k] = sigma > 0.0 ? alpha : x0;
        // p = beta A22 v, K = beta/2 p.v, w = p - K v, A22 -= v w^T + w v^T.
        double pw[N];
        double pv = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (i < m) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < N; ++j)
                    if (j < m)
                        s = fma(a[sym(k + 1 + i, k + 1 + j)], v[j], s);
                pw[i] = beta * s;
                pv = fma(pw[i], v[i], pv);
            }
        const double Kc = 0.5 * beta * pv;
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (i < m)
                pw[i] = fma(-Kc, v[i], pw[i]);
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j)
                if (i < m && j < m && j >= i)
                    a[sym(k + 1 + i, k + 1 + j)] =
                        fma(-v[i], pw[j], fma(-pw[i], v[j], a[sym(k + 1 + i, k + 1 + j)]));
        // V <- V H_k on columns k+1..N-1.
#pragma unroll
        for (int r = 0; r < N; ++r) {
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (i < m)
                    t = fma(V[r][k + 1 + i], v[i], t);
            t *= beta;
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (i < m)
                    V[r][k + 1 + i] = fma(-t, v[i], V[r][k + 1 + i]);
        }
    }
    e[N - 2] = a[sym(N - 2, N - 1)];
    e[N - 1] = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
        d[i] = a[sym(i, i)];

    // ---- Implicit QL.  For each l, sweeps run until e[l] is negligible; a sweep from the deflation point m
    // (the first i >= l with negligible e[i]) down to l applies plane rotations i = m-1, ..., l.  The rotations
    // are unrolled over i = N-2 .. l and predicated on i < m.
    const double tiny = 4e-16 * scale;
    bool ok = true;
#pragma unroll
    for (int l = 0; l < N - 1; ++l) {
        for (int it = 0; it < 40; ++it) {
            int m = N - 1;
#pragma unroll
            for (int i = N - 2; i >= l; --i)
                if (fabs(e[i]) <= tiny)
                    m = i;
            const bool active = m != l;
            if (!any_lane(active))
                break;
            if (it == 39)
                ok = ok && !active;
            // Shift from the leading 2x2 block (d[l], e[l], d[l+1]).
            double g = (d[l + 1] - d[l]) * (0.5 * recip(e[l]));
            double r = fma(g, g, 1.0) * rsqrt(fma(g, g, 1.0));
            // d[m] is selected by predication (m is per lane).
            double dm = d[N - 1];
#pragma unroll
            for (int i = N - 2; i >= l; --i)
                if (i == m)
                    dm = d[i];
            g = dm - d[l] + e[l] * recip(g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            bool alive = active; // becomes false after an exact zero rotation (splits the matrix)
#pragma unroll
            for (int i = N - 2; i >= l; --i) {
                if (alive && i < m) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    const double rr2 = fma(f, f, g * g);
                    if (rr2 == 0.0) {
                        // Exact split: finish this sweep here (e[m] is cleared below) and test again.
                        d[i + 1] -= p;
                        alive = false;
                    } else {
                        const double ir = rsqrt(rr2);
                        e[i + 1] = rr2 * ir;
                        s = f * ir;
                        c = g * ir;
                        g = d[i + 1] - p;
                        const double rq = fma(d[i] - g, s, 2.0 * c * b);
                        p = s * rq;
                        d[i + 1] = g + p;
                        g = fma(c, rq, -b);
#pragma unroll
                        for (int k = 0; k < N; ++k) {
                            f = V[k][i + 1];
                            V[k][i + 1] = fma(s, V[k][i], c * f);
                            V[k][i] = fma(c, V[k][i], -(s * f));
                        }
                    }
                }
            }
            if (alive) {
                d[l] -= p;
                e[l] = g;
            }
            // e[m] = 0 for the active lanes (predicated over the possible m; m = N-1 restores the sentinel).
#pragma unroll
            for (int i = N - 1; i >= l; --i)
                if (active && i == m)
                    e[i] = 0.0;
        }
    }
    return ok;
}

} // namespace eig9
} // namespace lpg
