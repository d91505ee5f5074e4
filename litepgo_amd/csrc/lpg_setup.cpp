// Host-side construction: validation, the device arrays, the sparsity pattern, vertex incidence lists and tiles.
//
// The pattern is built at vertex granularity.  The reference forms every (row, col) pair of every element's 12
// DOFs and deduplicates them with setFromTriplets (src/core/dof/energy_assembler.cpp:38-46); because a Tet P1
// element couples all 3 DOFs of its 4 vertices, that set is exactly {(3v+i, 3w+k) : v, w share a tet}, so each
// vertex's sorted neighbour list (itself included) expands to its three rows.  Row 3v+i of the CSR matrix is
// therefore [3w, 3w+1, 3w+2 for w in nbr(v)], and rows 3v, 3v+1, 3v+2 are contiguous in the value array: a vertex
// owns one contiguous segment of 9 |nbr(v)| values, which is what the device kernels exploit.
#include "lpg_internal.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <thread>

// Explicit fma only (the geometry check mirrors the kernels bit for bit).
#pragma clang fp contract(off)

namespace lpg {
namespace {

template <class F> void parallel_for(int64_t count, int nthreads, F&& body) {
    nthreads = std::max(1, std::min<int>(nthreads, static_cast<int>(std::min<int64_t>(count, 64))));
    if (nthreads <= 1) {
        for (int64_t i = 0; i < count; ++i)
            body(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    const int64_t chunk = std::max<int64_t>(1, count / (nthreads * 16));
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([&] {
            for (;;) {
                const int64_t b = next.fetch_add(chunk);
                if (b >= count)
                    return;
                const int64_t e = std::min(count, b + chunk);
                for (int64_t i = b; i < e; ++i)
                    body(i);
            }
        });
    for (auto& th : pool)
        th.join();
}

bool finite9(const double* p) {
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(p[i]))
            return false;
    return true;
}

} // namespace

void build_mesh_setup(MeshSetup& s, int32_t nv, int32_t nt, const int32_t* tets, const double* rest,
                      const double* theta, const double* Fp, int nthreads) {
    // MeshTopology (src/core/mesh/mesh_topology.hpp:25-43) and VertexDofMap (vertex_dof_map.hpp:20-24) checks.
    if (nv < 0 || nt < 0)
        fail(kInvalid, "MeshTopology requires a non-negative vertex count");
    if (static_cast<int64_t>(nv) * 3 > std::numeric_limits<int32_t>::max())
        fail(kInvalid, "VertexDofMap requires positive width and total DOFs fitting in int");
    if (nt > 0 && (!tets || !rest || !theta))
        fail(kInvalid, "VolumetricFemEnergy: incompatible map or response counts");
    for (int64_t e = 0; e < nt; ++e) {
        int32_t v[4] = {tets[4 * e], tets[4 * e + 1], tets[4 * e + 2], tets[4 * e + 3]};
        std::sort(v, v + 4);
        if (v[0] < 0 || v[3] >= nv)
            fail(kInvalid, "MeshTopology cell " + std::to_string(e) + " has vertex indices outside [0, num_vertices)");
        if (v[0] == v[1] || v[1] == v[2] || v[2] == v[3])
            fail(kInvalid, "MeshTopology cell " + std::to_string(e) + " must contain distinct vertices");
    }
    s.fp.clear();
    s.fp_error = 0;
    s.fp_error_msg.clear();
    bool identity = true;
    if (Fp)
        for (int64_t e = 0; e < nt && identity; ++e)
            for (int k = 0; k < 9; ++k)
                if (Fp[9 * e + k] != (k % 4 == 0 ? 1.0 : 0.0))
                    identity = false;
    if (!identity) {
        // Fp column-major (Fp(i, j) = Fp[9 e + i + 3 j]); Fp^-1 = adj(Fp) / det Fp, stored row-major.
        s.fp.assign(10 * static_cast<size_t>(nt), 0.0);
        for (int64_t e = 0; e < nt; ++e) {
            double A[9]; // row-major
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    A[3 * i + j] = Fp[9 * e + i + 3 * j];
            bool finite = true;
            for (double a : A)
                finite = finite && std::isfinite(a);
            if (!finite) {
                if (!s.fp_error) {
                    s.fp_error = kInvalid;
                    s.fp_error_msg = "MultiplicativeKinematics requires finite F and Fp";
                }
                continue;
            }
            double C[9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                    C[3 * i + j] = A[3 * i1 + j1] * A[3 * i2 + j2] - A[3 * i1 + j2] * A[3 * i2 + j1];
                }
            const double det = A[0] * C[0] + A[1] * C[1] + A[2] * C[2];
            double* out = &s.fp[10 * e];
            bool ok = std::isfinite(det) && det > 0.0;
            for (int i = 0; i < 3 && ok; ++i)
                for (int j = 0; j < 3; ++j) {
                    out[3 * i + j] = C[3 * j + i] / det;
                    ok = ok && std::isfinite(out[3 * i + j]);
                }
            out[9] = det;
            if (!ok && !s.fp_error) {
                s.fp_error = kDomain;
                s.fp_error_msg = (std::isfinite(det) && det > 0.0)
                                     ? "MultiplicativeKinematics: non-finite inverse or elastic gradient"
                                     : "MultiplicativeKinematics requires finite det(Fp) > 0";
            }
        }
    }
    s.num_vertices = nv;
    s.num_tets = nt;
    s.num_dofs = 3 * nv;

    // Rest geometry checks of VolumetricElement3D (volumetric_element.hpp:23-39) and the deformation-gradient
    // sampling (deformation_gradient_sampling_response.hpp:24-41): Dm = [X1-X0 | X2-X0 | X3-X0] must have a finite,
    // positive determinant, w = det(Dm)/6 and Dm^-1 must be finite.  The kernels recompute Dm^-1 and w from the rest
    // positions, so nothing per tet is stored beyond connectivity and theta.
    std::atomic<int64_t> bad_element{std::numeric_limits<int64_t>::max()};
    std::atomic<int> bad_kind{0};
    parallel_for(nt, nthreads, [&](int64_t e) {
        const int32_t* c = tets + 4 * e;
        double Dm[9]; // row-major Dm(r, k) = (X_{k+1} - X_0)_r
        for (int k = 0; k < 3; ++k)
            for (int r = 0; r < 3; ++r)
                Dm[3 * r + k] = rest[3 * static_cast<int64_t>(c[k + 1]) + r] - rest[3 * static_cast<int64_t>(c[0]) + r];
        auto note = [&](int kind) {
            int64_t cur = bad_element.load();
            while (e < cur && !bad_element.compare_exchange_weak(cur, e)) {
            }
            if (bad_element.load() == e)
                bad_kind.store(kind);
        };
        if (!finite9(Dm)) {
            note(1);
            return;
        }
        // Same operation sequence as the kernels' rest_geometry (explicit fma, 1/det then products), so the
        // host and the device agree bitwise on det, w and Dm^-1.
        double C[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                C[3 * i + j] = std::fma(Dm[3 * i1 + j1], Dm[3 * i2 + j2], -(Dm[3 * i1 + j2] * Dm[3 * i2 + j1]));
            }
        const double det = std::fma(Dm[2], C[2], std::fma(Dm[1], C[1], Dm[0] * C[0]));
        const double w = det * (1.0 / 6.0);
        if (!std::isfinite(det) || det <= 0 || !std::isfinite(w)) {
            note(2);
            return;
        }
        const double inv = 1.0 / det;
        double Dinv[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                Dinv[3 * i + j] = C[3 * j + i] * inv;
        if (!finite9(Dinv))
            note(3);
    });
    if (bad_element.load() != std::numeric_limits<int64_t>::max()) {
        const int kind = bad_kind.load();
        fail(kInvalid, kind == 1   ? "VolumetricElement3D: invalid reference Jacobian"
                       : kind == 2 ? "VolumetricElement3D: invalid reference determinant or volume weight"
                                   : "Deformation gradient sampling: non-finite physical gradient");
    }
    s.conn.assign(tets, tets + 4 * static_cast<int64_t>(nt));
    s.theta.assign(theta, theta + 2 * static_cast<int64_t>(nt));
    s.Xr.assign(4 * static_cast<size_t>(nv), 0.0);
    for (int64_t v = 0; v < nv; ++v)
        for (int c = 0; c < 3; ++c)
            s.Xr[4 * v + c] = rest[3 * v + c];
    // NeoHookean checks theta when it evaluates (neo_hookean.cpp:15-16); remember the outcome for eval.
    s.theta_valid = true;
    for (int64_t e = 0; e < nt; ++e) {
        const double l = theta[2 * e], m = theta[2 * e + 1];
        if (!std::isfinite(l) || !std::isfinite(m) || l < 0.0 || m <= 0.0) {
            s.theta_valid = false;
            break;
        }
    }

    // Incidences, element order per vertex (the reference's dof_contributions_ order, energy_assembler.cpp:48-54).
    s.inc_ptr.assign(static_cast<size_t>(nv) + 1, 0);
    for (int64_t k = 0; k < 4 * static_cast<int64_t>(nt); ++k)
        ++s.inc_ptr[tets[k] + 1];
    for (int32_t v = 0; v < nv; ++v)
        s.inc_ptr[v + 1] += s.inc_ptr[v];
    s.inc_tet.resize(4 * static_cast<size_t>(nt));
    {
        std::vector<int32_t> fill(s.inc_ptr.begin(), s.inc_ptr.end() - 1);
        for (int64_t e = 0; e < nt; ++e)
            for (int b = 0; b < 4; ++b)
                s.inc_tet[fill[tets[4 * e + b]]++] = static_cast<int32_t>(e);
    }

    // Sorted neighbour lists (self included).
    std::vector<std::vector<int32_t>> nb(nv);
    parallel_for(nv, nthreads, [&](int64_t v) {
        auto& list = nb[v];
        list.reserve(4 * (s.inc_ptr[v + 1] - s.inc_ptr[v]) + 1);
        list.push_back(static_cast<int32_t>(v));
        for (int32_t k = s.inc_ptr[v]; k < s.inc_ptr[v + 1]; ++k)
            for (int b = 0; b < 4; ++b)
                list.push_back(tets[4 * static_cast<int64_t>(s.inc_tet[k]) + b]);
        std::sort(list.begin(), list.end());
        list.erase(std::unique(list.begin(), list.end()), list.end());
    });
    s.nbr_ptr.assign(static_cast<size_t>(nv) + 1, 0);
    s.max_degree = 0;
    for (int32_t v = 0; v < nv; ++v) {
        // A vertex without elements has an empty pattern row (no self entry), as in the reference.
        const int32_t deg = s.inc_ptr[v + 1] > s.inc_ptr[v] ? static_cast<int32_t>(nb[v].size()) : 0;
        s.nbr_ptr[v + 1] = s.nbr_ptr[v] + deg;
        s.max_degree = std::max(s.max_degree, deg);
    }
    s.wide = s.max_degree > 255;
    s.nbr.resize(s.nbr_ptr[nv]);
    parallel_for(nv, nthreads, [&](int64_t v) {
        if (s.nbr_ptr[v + 1] > s.nbr_ptr[v])
            std::copy(nb[v].begin(), nb[v].end(), s.nbr.begin() + s.nbr_ptr[v]);
    });
    std::vector<std::vector<int32_t>>().swap(nb);

    // CSR pattern: rows 3v+i all hold [3w + k for w in nbr(v), k = 0..2].
    s.row_ptr.assign(static_cast<size_t>(s.num_dofs) + 1, 0);
    int64_t total = 0;
    for (int32_t v = 0; v < nv; ++v) {
        const int64_t len = 3 * static_cast<int64_t>(s.nbr_ptr[v + 1] - s.nbr_ptr[v]);
        for (int i = 0; i < 3; ++i) {
            total += len;
            if (total > std::numeric_limits<int32_t>::max())
                fail(kInvalid, "EnergyAssembler: local Hessian entry count must fit in int");
            s.row_ptr[3 * static_cast<int64_t>(v) + i + 1] = static_cast<int32_t>(total);
        }
    }
    s.nnz = static_cast<int32_t>(total);
    s.col_idx.resize(static_cast<size_t>(s.nnz));
    parallel_for(nv, nthreads, [&](int64_t v) {
        const int32_t b = s.nbr_ptr[v], d = s.nbr_ptr[v + 1] - b;
        for (int i = 0; i < 3; ++i) {
            int32_t* out = s.col_idx.data() + s.row_ptr[3 * v + i];
            for (int32_t q = 0; q < d; ++q)
                for (int k = 0; k < 3; ++k)
                    out[3 * q + k] = 3 * s.nbr[b + q] + k;
        }
    });

    // Column-block positions per incidence, and each vertex's own position in its neighbour list.
    s.inc_colpos.resize(16 * static_cast<size_t>(nt));
    s.self_pos.assign(static_cast<size_t>(nv), 0);
    parallel_for(nv, nthreads, [&](int64_t v) {
        const int32_t* nb_begin = s.nbr.data() + s.nbr_ptr[v];
        const int32_t* nb_end = s.nbr.data() + s.nbr_ptr[v + 1];
        s.self_pos[v] = static_cast<int32_t>(std::lower_bound(nb_begin, nb_end, static_cast<int32_t>(v)) - nb_begin);
        for (int32_t k = s.inc_ptr[v]; k < s.inc_ptr[v + 1]; ++k) {
            const int64_t e = s.inc_tet[k];
            for (int b = 0; b < 4; ++b)
                s.inc_colpos[4 * static_cast<int64_t>(k) + b] =
                    static_cast<uint16_t>(std::lower_bound(nb_begin, nb_end, tets[4 * e + b]) - nb_begin);
        }
    });
}

void build_slot_map(const MeshSetup& s, int32_t* offsets) {
    // offsets[144 e + i + 12 j] = CSC position of (row idx[i], col idx[j]) = CSR position of (row idx[j], col idx[i])
    // (the pattern is symmetric, and CSC column c is CSR row c).
    parallel_for(s.num_tets, 8, [&](int64_t e) {
        const int32_t* ids = s.conn.data() + 4 * e;
        for (int j = 0; j < 12; ++j) {
            const int32_t vj = ids[j / 3], cj = j % 3;
            const int32_t* nb_begin = s.nbr.data() + s.nbr_ptr[vj];
            const int32_t* nb_end = s.nbr.data() + s.nbr_ptr[vj + 1];
            for (int i = 0; i < 12; ++i) {
                const int32_t vi = ids[i / 3], ci = i % 3;
                const int32_t pos = static_cast<int32_t>(std::lower_bound(nb_begin, nb_end, vi) - nb_begin);
                offsets[144 * e + i + 12 * j] = s.row_ptr[3 * vj + cj] + 3 * pos + ci;
            }
        }
    });
}

void build_tiles(const MeshSetup& s, const TileConfig& cfg, std::vector<int32_t>& tile_vbeg, TileCaps& caps,
                 std::vector<int32_t>* oversized, const std::vector<int32_t>* breaks) {
    tile_vbeg.clear();
    caps = TileCaps{};
    if (oversized)
        oversized->clear();
    const int32_t nv = s.num_vertices;
    auto segment = [&](int32_t vb, int32_t ve) { return int64_t{9} * (s.nbr_ptr[ve] - s.nbr_ptr[vb]); };
    int32_t v = 0;
    size_t next_break = 0;
    tile_vbeg.push_back(0);
    while (v < nv) {
        const int32_t start = v;
        while (breaks && next_break < breaks->size() && (*breaks)[next_break] <= start)
            ++next_break;
        const int32_t limit = (breaks && next_break < breaks->size()) ? std::min(nv, (*breaks)[next_break]) : nv;
        ++v; // a tile holds at least one vertex
        while (v < limit && v - start < cfg.vertices_per_tile && segment(start, v + 1) <= cfg.lds_budget_doubles)
            ++v;
        const int64_t seg = segment(start, v);
        if (seg > cfg.lds_budget_doubles && oversized)
            oversized->push_back(static_cast<int32_t>(tile_vbeg.size() - 1));
        else
            caps.segment = std::max<int32_t>(caps.segment, static_cast<int32_t>(seg));
        tile_vbeg.push_back(v);
    }
}

void build_energy_chunks(const std::vector<int32_t>& seg_vbeg, std::vector<int32_t>& chunk_begin,
                         std::vector<int32_t>& seg_chunk) {
    constexpr int32_t kChunk = 1024;
    chunk_begin.assign(1, seg_vbeg.empty() ? 0 : seg_vbeg.front());
    seg_chunk.assign(1, 0);
    for (size_t m = 0; m + 1 < seg_vbeg.size(); ++m) {
        for (int32_t v = seg_vbeg[m]; v < seg_vbeg[m + 1]; v += kChunk)
            chunk_begin.push_back(std::min(seg_vbeg[m + 1], v + kChunk));
        seg_chunk.push_back(static_cast<int32_t>(chunk_begin.size()) - 1);
    }
}

} // namespace lpg
