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

// ---------------------------------------------------------------------------------------------------------------
// PSD part without the full eigenbasis.
//
// proj(M) = V max(Lambda, 0) V^T = M + sum_{lambda_q < 0} |lambda_q| v_q v_q^T = sum_{lambda_q > 0} lambda_q v_q v_q^T.
// The eigenvalues come from the QL iteration on the tridiagonal form alone (no basis rotations), and eigenvectors
// are computed only for the smaller of the negative and positive sets (at most 4 of 9): inverse iteration on the
// tridiagonal matrix T - lambda_q I (Gaussian elimination with partial pivoting, two solves), orthogonalised
// against the vectors already found (modified Gram-Schmidt, so clustered eigenvalues give an orthonormal basis of
// their eigenspace), then mapped back through the Householder reflectors.  Errors: a vector's angle error
// ~ eps |T| / gap only mixes it with its neighbours; within the selected set the mixing cancels in the sum up to
// the cluster width, and a neighbour across zero has |lambda| <= gap, so proj(M) is accurate to a few eps |M|.
// ---------------------------------------------------------------------------------------------------------------

// Start vectors of the inverse iterations (fixed, generic values in [0.5, 1.5], one row per eigenvalue index).
constexpr double kStartVec[N][N] = {
    {1.06, 0.71, 1.33, 0.58, 1.21, 0.93, 1.47, 0.66, 1.12}, {0.84, 1.29, 0.62, 1.41, 0.77, 1.18, 0.53, 1.36, 0.98},
    {1.24, 0.57, 1.09, 0.88, 1.45, 0.69, 1.02, 1.31, 0.61}, {0.73, 1.38, 0.95, 0.52, 1.16, 1.27, 0.81, 0.64, 1.43},
    {1.35, 0.92, 0.68, 1.22, 0.59, 1.04, 1.39, 0.86, 0.74}, {0.55, 1.11, 1.42, 0.97, 1.30, 0.63, 0.76, 1.19, 1.05},
    {1.18, 0.83, 0.56, 1.34, 1.01, 1.46, 0.67, 0.91, 1.26}, {0.96, 1.49, 1.14, 0.72, 0.85, 0.54, 1.23, 1.07, 0.79},
    {1.44, 0.65, 0.89, 1.08, 0.74, 1.37, 1.15, 0.51, 1.28}};

// Values psd_part keeps in its spill store for the back-transformation: the Householder vectors, normalised so that
// each reflector is I - 2 v v^T (35 values; 17.5 KB of LDS per 64-lane wave).
constexpr int kSpill = (N - 1) * N / 2 - 1;

// a: packed upper triangle of the symmetric matrix M (destroyed).  On return
//   proj(M) = (add_M ? M : 0) + sum_{j < ns} wts[j] X[j] X[j]^T,
// with ns <= 4 unit vectors X[j] and weights wts[j] >= 0 (|lambda| of the negative eigenvalues when add_M, else the
// positive eigenvalues); X[j] and wts[j] for j >= ns are zero.  `spill` provides put(i, value) / get(i) for kSpill
// doubles (LDS on the device), which hold the reflectors while the inverse iterations run.  Returns false if a QL
// sweep did not converge.
template <class Rsqrt, class Recip, class Spill>
LPG_EIG_FN bool psd_part(double a[45], double X[4][N], double wts[4], int& ns, bool& add_M, double scale,
                         Rsqrt rsqrt, Recip recip, Spill& spill) {
    // No implicit FMA contraction: the QL sweep has two code paths, taken per wave, that must round alike, so every
    // lane's result depends on its own matrix only (not on which lanes share its wave).
#pragma clang fp contract(off)
    double e[N], d[N];
    // ---- Householder reduction; reflector k (unit vector on indices k+1..N-1) goes to the spill store at offset
    // off(k) = sum_{k' < k} (N - 1 - k').
#pragma unroll
    for (int k = 0; k < N - 2; ++k) {
        const int m = N - 1 - k;
        const int off = k * (N - 1) - k * (k - 1) / 2;
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
        const double bk = sigma > 0.0 ? 2.0 * recip(vtv) : 0.0;
        e[k] = sigma > 0.0 ? alpha : x0;
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
                pw[i] = bk * s;
                pv = fma(pw[i], v[i], pv);
            }
        const double Kc = 0.5 * bk * pv;
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
        // Stored as v / |v| (H_k = I - 2 v v^T); a zero vector when the step is the identity (sigma == 0).
        const double vs = sigma > 0.0 ? rsqrt(vtv) : 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (i < m)
                spill.put(off + i, v[i] * vs);
    }
    e[N - 2] = a[sym(N - 2, N - 1)];
    e[N - 1] = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
        d[i] = a[sym(i, i)];
    double d0[N], e0[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        d0[i] = d[i];
        e0[i] = e[i];
    }
    // The QL iteration runs on the reversed matrix (d[N-1-i], e[N-2-i]: the same eigenvalues).  On the tridiagonal
    // forms of reduced neo-Hookean Hessians it then converges in fewer sweeps: an emulation of the wave schedule
    // (each sweep repeated until the slowest of 64 lanes has converged) on listed tets counts 30% fewer plane
    // rotations per wave on a stretched 64^3 grid and 16% fewer on randomly perturbed small meshes.
#pragma unroll
    for (int i = 0; i < N; ++i) {
        d[i] = d0[N - 1 - i];
        e[i] = i < N - 1 ? e0[N - 2 - i] : 0.0;
    }

    // ---- Eigenvalues: implicit QL on (d, e), as in sym_eig but without the basis.
    const double tiny = 4e-16 * scale;
    bool ok = true;
#pragma unroll
    for (int l = 0; l < N - 1; ++l) {
        for (int it = 0; it < 40; ++it) {
            // A lane is active while e[l] is not negligible, and split when some e[i], l < i < N - 1, is: its
            // sweep then starts at the first such i (the deflation point m) instead of N - 1.
            const bool active = !(fabs(e[l]) <= tiny);
            if (!any_lane(active))
                break;
            if (it == 39)
                ok = ok && !active;
            bool split = false;
#pragma unroll
            for (int i = N - 2; i > l; --i)
                split = split || fabs(e[i]) <= tiny;
            double g = (d[l + 1] - d[l]) * (0.5 * recip(e[l]));
            double r = fma(g, g, 1.0) * rsqrt(fma(g, g, 1.0));
            const double shift_term = e[l] * recip(g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            if (!any_lane(active && split)) {
                // No active lane is split below l (the usual case): every e[i], i >= l, is above the deflation
                // threshold, so f = s e[i] never vanishes and no rotation meets f = g = 0.  The sweep then needs
                // neither the per-lane start m nor the exact-split branch, and does the same arithmetic as the
                // general sweep below.
                if (active) {
                    g = d[N - 1] - d[l] + shift_term;
#pragma unroll
                    for (int i = N - 2; i >= l; --i) {
                        const double f = s * e[i];
                        const double b = c * e[i];
                        const double rr2 = fma(f, f, g * g);
                        const double ir = rsqrt(rr2);
                        e[i + 1] = rr2 * ir;
                        s = f * ir;
                        c = g * ir;
                        g = d[i + 1] - p;
                        const double rq = fma(d[i] - g, s, 2.0 * c * b);
                        p = s * rq;
                        d[i + 1] = g + p;
                        g = fma(c, rq, -b);
                    }
                    d[l] -= p;
                    e[l] = g;
                    e[N - 1] = 0.0;
                }
                continue;
            }
            int m = N - 1;
#pragma unroll
            for (int i = N - 2; i >= l; --i)
                if (fabs(e[i]) <= tiny)
                    m = i;
            // d[m] is selected by predication (m is per lane).
            double dm = d[N - 1];
#pragma unroll
            for (int i = N - 2; i >= l; --i)
                if (i == m)
                    dm = d[i];
            g = dm - d[l] + shift_term;
            bool alive = active;
#pragma unroll
            for (int i = N - 2; i >= l; --i) {
                if (alive && i < m) {
                    const double f = s * e[i];
                    const double b = c * e[i];
                    const double rr2 = fma(f, f, g * g);
                    if (rr2 == 0.0) {
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
                    }
                }
            }
            if (alive) {
                d[l] -= p;
                e[l] = g;
            }
#pragma unroll
            for (int i = N - 1; i >= l; --i)
                if (active && i == m)
                    e[i] = 0.0;
        }
    }

    // ---- The smaller eigen-set: at most 4 eigenvalues (negative ones when there are at most 4, else the
    // positive ones, of which there are then at most 4), compacted in index order.
    int nneg = 0;
#pragma unroll
    for (int q = 0; q < N; ++q)
        nneg += d[q] < 0.0 ? 1 : 0;
    const bool use_neg = nneg <= 4;
    add_M = use_neg;
    double sl[4] = {0.0, 0.0, 0.0, 0.0};
    ns = 0;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const bool sel = use_neg ? d[q] < 0.0 : d[q] > 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (sel && j == ns)
                sl[j] = d[q];
        ns += sel ? 1 : 0;
    }

    // ---- Their vectors in the tridiagonal basis by inverse iteration, orthogonalised against the earlier ones.
    const double piv_floor = 2.220446049250313e-16 * scale;
    double Y[4][N];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int i = 0; i < N; ++i)
            Y[j][i] = 0.0;
        if (!any_lane(j < ns))
            continue;
        const double lq = sl[j];
        // T - lq I = P L U by Gaussian elimination with partial pivoting (U has two superdiagonals): per step the
        // pivot row, the multiplier and whether rows were swapped.
        double u1[N], u2[N], iu[N], mult[N - 1];
        bool swp[N - 1];
        double c0 = d0[0] - lq, c1 = e0[0];
#pragma unroll
        for (int i = 0; i < N - 1; ++i) {
            const double n0 = e0[i], n1 = d0[i + 1] - lq, n2 = i + 2 < N ? e0[i + 1] : 0.0;
            const bool sw = fabs(n0) > fabs(c0);
            double p0 = sw ? n0 : c0;
            const double p1 = sw ? n1 : c1, p2 = sw ? n2 : 0.0;
            const double o0 = sw ? c0 : n0, o1 = sw ? c1 : n1, o2 = sw ? 0.0 : n2;
            if (fabs(p0) < piv_floor)
                p0 = p0 < 0.0 ? -piv_floor : piv_floor;
            iu[i] = recip(p0);
            const double mu = o0 * iu[i];
            u1[i] = p1;
            u2[i] = p2;
            mult[i] = mu;
            swp[i] = sw;
            c0 = fma(-mu, p1, o1);
            c1 = fma(-mu, p2, o2);
        }
        if (fabs(c0) < piv_floor)
            c0 = c0 < 0.0 ? -piv_floor : piv_floor;
        iu[N - 1] = recip(c0);
        double y[N];
#pragma unroll
        for (int i = 0; i < N; ++i)
            y[i] = kStartVec[j][i];
        // Two solves.  With an eigenvalue accurate to rounding, one solve from a generic start vector already lies
        // in the eigenspace up to rounding, and the second settles clusters; each further solve re-amplifies the
        // directions of the vectors already found, which the Gram-Schmidt pass must then cancel again.  On the
        // 185,549 matrices of the CPU test at 30 times its case counts (random, clustered, repeated and zero
        // spectra, neo-Hookean M) two solves leave the projection within 6.1e-14 of max|M|, three within 2.1e-13,
        // one only within 3.1e-11.
        constexpr int kSolves = 2;
#pragma unroll
        for (int it = 0; it <= kSolves; ++it) {
            // Modified Gram-Schmidt against the vectors already found, then normalisation.  The last pass runs
            // twice: in a repeated eigenvalue the last vector's remaining component can be small, and a single
            // pass then leaves it measurably non-orthogonal to the others.
#pragma unroll
            for (int pass = 0; pass < (it == kSolves ? 2 : 1); ++pass)
#pragma unroll
            for (int jj = 0; jj < j; ++jj) {
                double dot = 0.0;
#pragma unroll
                for (int i = 0; i < N; ++i)
                    dot = fma(y[i], Y[jj][i], dot);
#pragma unroll
                for (int i = 0; i < N; ++i)
                    y[i] = fma(-dot, Y[jj][i], y[i]);
            }
            double nrm = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i)
                nrm = fma(y[i], y[i], nrm);
            const double inv = rsqrt(nrm);
#pragma unroll
            for (int i = 0; i < N; ++i)
                y[i] *= inv;
            if (it == kSolves)
                break;
            // y <- (T - lq I)^-1 y: forward elimination with the recorded swaps, then back substitution.
#pragma unroll
            for (int i = 0; i < N - 1; ++i) {
                const double zi = y[i], zj = y[i + 1];
                y[i] = swp[i] ? zj : zi;
                y[i + 1] = fma(-mult[i], y[i], swp[i] ? zi : zj);
            }
            y[N - 1] = y[N - 1] * iu[N - 1];
            y[N - 2] = fma(-u1[N - 2], y[N - 1], y[N - 2]) * iu[N - 2];
#pragma unroll
            for (int i = N - 3; i >= 0; --i)
                y[i] = fma(-u2[i], y[i + 2], fma(-u1[i], y[i + 1], y[i])) * iu[i];
        }
#pragma unroll
        for (int i = 0; i < N; ++i)
            Y[j][i] = y[i];
    }

    // ---- Back to the basis of M: x = H_0 H_1 ... H_{N-3} y.
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        wts[j] = j < ns ? fabs(sl[j]) : 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i)
            X[j][i] = 0.0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (!any_lane(j < ns))
            break;
        double x[N];
#pragma unroll
        for (int i = 0; i < N; ++i)
            x[i] = j < ns ? Y[j][i] : 0.0; // lanes with fewer vectors add nothing (and no NaN)
#pragma unroll
        for (int k = N - 3; k >= 0; --k) {
            const int m = N - 1 - k;
            const int off = k * (N - 1) - k * (k - 1) / 2;
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (i < m)
                    t = fma(spill.get(off + i), x[k + 1 + i], t);
            t *= 2.0;
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (i < m)
                    x[k + 1 + i] = fma(-t, spill.get(off + i), x[k + 1 + i]);
        }
#pragma unroll
        for (int i = 0; i < N; ++i)
            X[j][i] = x[i];
    }
    return ok;
}

// ---------------------------------------------------------------------------------------------------------------
// Definiteness certificate of the neo-Hookean tangent A(i+3j, m+3l) = mu d_im d_jl + lambda a_ij a_ml - p a_il a_mj
// (a = F^-T, p = lambda log J - mu; neo_hookean.cpp:43-60) without forming A.  In the singular-value frame
// F = U diag(sigma) V^T the 9x9 form splits into
//   * three 2x2 blocks [[mu, -p s_ij], [-p s_ij, mu]] with s_ij = 1 / (sigma_i sigma_j), i < j: positive definite
//     iff mu sigma_i sigma_j > |p| for every pair, i.e. (min pair product = J / sigma_max)
//     mu^2 J^2 > p^2 sigma_max^2, i.e. the 3x3 matrix X = mu^2 J^2 I - p^2 C (C = F^T F) is positive definite;
//   * the 3x3 block B = diag(mu - p / c_i) + lambda s s^T (c_i = sigma_i^2, s_i = 1 / sigma_i).  For p <= 0 it is
//     positive definite.  For p > 0 and the pair condition strict, c_mid >= sigma_min sigma_mid > p / mu, so at most
//     one diagonal entry is negative and, by interlacing, B > 0 iff det B > 0; det B = N / I3 with
//     N = mu^3 I3 - mu^2 p I2 + mu p^2 I1 - p^3 + lambda (mu^2 I2 - 2 mu p I1 + 3 p^2)
//     (I1, I2, I3: the invariants of C; the determinant lemma det(D + lambda s s^T) = det D + lambda s^T adj(D) s).
// Returns true only when both hold with a relative margin of kTangentMargin, so a certified tangent is positive
// definite beyond rounding; false means indefinite or too close to singular to tell (the caller then runs the
// full eigen-analysis).  F row-major (F[3i+j] = F_ij), J = det F.
constexpr double kTangentMargin = 1e-10;

LPG_EIG_FN bool tangent_certified_pd(const double F[9], double J, double lam, double mu, double p) {
    double C[9]; // C = F^T F, C[3j+k] = sum_i F_ij F_ik
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = j; k < 3; ++k) {
            const double v = fma(F[j], F[k], fma(F[3 + j], F[3 + k], F[6 + j] * F[6 + k]));
            C[3 * j + k] = v;
            C[3 * k + j] = v;
        }
    const double p2 = p * p;
    const double t = (1.0 - kTangentMargin) * (mu * J) * (mu * J);
    const double x00 = fma(-p2, C[0], t), x11 = fma(-p2, C[4], t), x22 = fma(-p2, C[8], t);
    const double x01 = -p2 * C[1], x02 = -p2 * C[2], x12 = -p2 * C[5];
    const double m2 = fma(x00, x11, -(x01 * x01));
    const double m3 = fma(x00, fma(x11, x22, -(x12 * x12)), fma(-x01, fma(x01, x22, -(x12 * x02)),
                                                                 x02 * fma(x01, x12, -(x11 * x02))));
    if (!(x00 > 0.0 && m2 > 0.0 && m3 > 0.0))
        return false;
    if (p <= 0.0)
        return true;
    const double I1 = C[0] + C[4] + C[8];
    const double I2 = fma(C[0], C[4], fma(C[0], C[8], fma(C[4], C[8], -fma(C[1], C[1], fma(C[2], C[2], C[5] * C[5])))));
    const double I3 = J * J;
    const double mu2 = mu * mu, mu3 = mu2 * mu, p3 = p2 * p;
    const double t1 = mu3 * I3, t2 = mu2 * p * I2, t3 = mu * p2 * I1;
    const double t4 = lam * mu2 * I2, t5 = 2.0 * lam * mu * p * I1, t6 = 3.0 * lam * p2;
    const double N3 = ((t1 - t2) + (t3 - p3)) + ((t4 - t5) + t6);
    const double mag = ((fabs(t1) + fabs(t2)) + (fabs(t3) + p3)) + ((fabs(t4) + fabs(t5)) + t6);
    return N3 > kTangentMargin * mag;
}

} // namespace eig9
} // namespace lpg
