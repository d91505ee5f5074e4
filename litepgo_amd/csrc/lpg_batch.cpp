// C ABI of the multi-mesh batch, graph mode and the launch autotuner (include/litepgo_fem_batch.h).
#include "litepgo_fem_batch.h"
#include "lpg_handle.hpp"

#include <algorithm>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using lpg::fail;
using lpg::guarded;

namespace {

// Candidate launch shapes of the autotuner: (lanes per vertex, waves, vertices per tile, LDS budget).
struct Shape {
    int lanes, waves, vpt, budget;
};
constexpr Shape kShapes[] = {{4, 1, 16, 4096}, {4, 1, 12, 4096}, {4, 1, 8, 4096}, {4, 2, 32, 6144},
                             {4, 4, 64, 12288}};
// Persistent grid sizes tried for batches (0 = one workgroup per tile).
constexpr int kGrids[] = {0, 1024, 2048, 4096};

double time_assembly(lpg_tet_energy* h, const double* x, int kind, double* g, double* H, double* energies,
                     int repeats) {
    // One eager warm-up, then `repeats` timed evaluations; the assembly phase time from the handle's events.
    if (!h->ev[0])
        for (hipEvent_t& e : h->ev)
            lpg::hip_check(hipEventCreate(&e), "hipEventCreate");
    const bool timing_was = h->timing;
    h->timing = false;
    (void)lpg::run_eval(h, x, nullptr, 1, kind, g, H, energies, h->own_stream);
    h->timing = true;
    for (double& t : h->time_ms)
        t = 0;
    h->timed_evals = 0;
    for (int r = 0; r < repeats; ++r)
        (void)lpg::run_eval(h, x, nullptr, 1, kind, g, H, energies, h->own_stream);
    h->timing = timing_was;
    return h->timed_evals ? h->time_ms[1] / h->timed_evals : std::numeric_limits<double>::infinity();
}

// Autotuner choices cached per process.  Key: device, Hessian kind, mesh count, the histogram of tets per mesh in
// power-of-two bins, wide records and plastic parts; value: {lanes, waves, vertices per tile, LDS budget, grid} and
// the timed assembly ms.
struct TuneChoice {
    int32_t pick[5];
    double ms;
};
std::mutex g_tune_mutex;
std::map<std::vector<int64_t>, TuneChoice> g_tune_cache;

std::vector<int64_t> tune_key(const lpg_tet_energy* h, int kind) {
    std::vector<int64_t> key{h->device, kind, static_cast<int64_t>(h->seg_tbeg.size()) - 1,
                             h->setup.wide ? 1 : 0, h->setup.fp.empty() ? 0 : 1};
    std::vector<int64_t> bins(33, 0);
    for (size_t m = 0; m + 1 < h->seg_tbeg.size(); ++m) {
        const int64_t t = h->seg_tbeg[m + 1] - h->seg_tbeg[m];
        int b = 0;
        while (b < 32 && (int64_t{1} << (b + 1)) <= t)
            ++b;
        ++bins[t > 0 ? b + 1 : 0];
    }
    key.insert(key.end(), bins.begin(), bins.end());
    return key;
}

} // namespace

extern "C" {

lpg_status lpg_tet_batch_create(int32_t num_meshes, const int32_t* num_vertices, const int32_t* num_tets,
                                const int32_t* tets, const double* rest_positions, const double* theta,
                                lpg_tet_energy** out) {
    return guarded([&] {
        if (!out)
            fail(lpg::kInvalid, "lpg_tet_batch_create: out is null");
        *out = nullptr;
        if (num_meshes <= 0 || !num_vertices || !num_tets)
            fail(lpg::kInvalid, "lpg_tet_batch_create: needs at least one mesh");
        std::vector<int32_t> vbeg(1, 0), tbeg(1, 0);
        for (int32_t m = 0; m < num_meshes; ++m) {
            if (num_vertices[m] < 0 || num_tets[m] < 0)
                fail(lpg::kInvalid, "MeshTopology requires a non-negative vertex count");
            const int64_t nv = static_cast<int64_t>(vbeg.back()) + num_vertices[m];
            const int64_t nt = static_cast<int64_t>(tbeg.back()) + num_tets[m];
            if (3 * nv > std::numeric_limits<int32_t>::max() || nt > std::numeric_limits<int32_t>::max() / 4)
                fail(lpg::kInvalid, "lpg_tet_batch_create: the batch exceeds int32 indexing");
            vbeg.push_back(static_cast<int32_t>(nv));
            tbeg.push_back(static_cast<int32_t>(nt));
        }
        const int32_t V = vbeg.back(), T = tbeg.back();
        if (T > 0 && (!tets || !rest_positions || !theta))
            fail(lpg::kInvalid, "VolumetricFemEnergy: incompatible map or response counts");
        // Union connectivity: mesh-local ids shifted by the mesh's vertex offset, checked per mesh.
        std::vector<int32_t> conn(4 * static_cast<size_t>(T));
        for (int32_t m = 0; m < num_meshes; ++m)
            for (int64_t e = tbeg[m]; e < tbeg[m + 1]; ++e)
                for (int b = 0; b < 4; ++b) {
                    const int32_t local = tets[4 * e + b];
                    if (local < 0 || local >= num_vertices[m])
                        fail(lpg::kInvalid, "MeshTopology cell " + std::to_string(e - tbeg[m]) + " of mesh " +
                                                std::to_string(m) +
                                                " has vertex indices outside [0, num_vertices)");
                    conn[4 * e + b] = local + vbeg[m];
                }
        auto h = std::make_unique<lpg_tet_energy>();
        const int nthreads = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        lpg::build_mesh_setup(h->setup, V, T, conn.data(), rest_positions, theta, nullptr, nthreads);
        h->seg_vbeg = std::move(vbeg);
        h->seg_tbeg = std::move(tbeg);
        *out = h.release();
    });
}

int32_t lpg_tet_batch_num_meshes(const lpg_tet_energy* h) {
    if (!h)
        return 0;
    return h->seg_vbeg.empty() ? 1 : static_cast<int32_t>(h->seg_vbeg.size()) - 1;
}

lpg_status lpg_tet_batch_offsets(const lpg_tet_energy* h, int32_t* vertex_offsets, int32_t* tet_offsets,
                                 int32_t* nnz_offsets) {
    return guarded([&] {
        if (!h || !vertex_offsets || !tet_offsets || !nnz_offsets)
            fail(lpg::kInvalid, "lpg_tet_batch_offsets: null argument");
        const auto& s = h->setup;
        const std::vector<int32_t> vb = h->seg_vbeg.empty() ? std::vector<int32_t>{0, s.num_vertices} : h->seg_vbeg;
        const std::vector<int32_t> tb = h->seg_tbeg.empty() ? std::vector<int32_t>{0, s.num_tets} : h->seg_tbeg;
        for (size_t m = 0; m < vb.size(); ++m) {
            vertex_offsets[m] = vb[m];
            tet_offsets[m] = tb[m];
            // Rows 3v..3v+2 hold 9 |nbr(v)| values, so a mesh's values start at 9 nbr_ptr[first vertex].
            nnz_offsets[m] = 9 * s.nbr_ptr[vb[m]];
        }
    });
}

lpg_status lpg_tet_batch_eval_device(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                     int32_t want_gradient, int32_t hessian_kind, double* energies_dev,
                                     double* gradient_dev, double* hessian_values_dev, void* stream) {
    return guarded([&] {
        if (!h)
            fail(lpg::kInvalid, "lpg_tet_batch_eval_device: null handle");
        if (!energies_dev)
            fail(lpg::kInvalid, "lpg_tet_batch_eval_device: energies pointer is null");
        lpg::BusyScope scope(h->busy);
        (void)lpg::run_eval(h, x_dev, theta_dev, want_gradient, hessian_kind, gradient_dev, hessian_values_dev,
                            energies_dev, static_cast<hipStream_t>(stream));
    });
}

lpg_status lpg_tet_batch_max_feasible_step_device(lpg_tet_energy* h, const double* x_dev, const double* dx_dev,
                                                  double* alphas_dev, void* stream) {
    return guarded([&] {
        if (!h || !alphas_dev)
            fail(lpg::kInvalid, "lpg_tet_batch_max_feasible_step_device: null argument");
        lpg::BusyScope scope(h->busy);
        (void)lpg::feasible_step_call(h, x_dev, dx_dev, true, alphas_dev, static_cast<hipStream_t>(stream));
    });
}

lpg_status lpg_tet_energy_set_graph(lpg_tet_energy* h, int32_t enable) {
    return guarded([&] {
        if (!h)
            fail(lpg::kInvalid, "null handle");
        lpg::BusyScope scope(h->busy);
        h->graph.reset();
        h->graph.enabled = enable != 0;
    });
}

lpg_status lpg_tet_energy_autotune(lpg_tet_energy* h, const double* x_dev, int32_t hessian_kind, int32_t repeats,
                                   int32_t* chosen, double* best_ms) {
    return guarded([&] {
        if (!h || !chosen)
            fail(lpg::kInvalid, "lpg_tet_energy_autotune: null argument");
        if (hessian_kind < 0 || hessian_kind > 1)
            fail(lpg::kInvalid, "Energy: invalid Hessian kind");
        lpg::BusyScope scope(h->busy);
        lpg::ensure_device(h);
        const auto& s = h->setup;
        if (!x_dev && s.num_dofs > 0)
            fail(lpg::kInvalid, "Energy::eval: x must have num_dofs finite entries");
        const bool refresh = repeats < 0;
        repeats = std::max(1, repeats < 0 ? -repeats : repeats);
        const std::vector<int64_t> key = tune_key(h, hessian_kind);
        if (!refresh) {
            TuneChoice c;
            bool hit = false;
            {
                std::lock_guard<std::mutex> lock(g_tune_mutex);
                const auto it = g_tune_cache.find(key);
                if (it != g_tune_cache.end()) {
                    c = it->second;
                    hit = true;
                }
            }
            if (hit) {
                lpg::LaunchConfig lc;
                lc.lanes_per_vertex = c.pick[0];
                lc.waves = c.pick[1];
                lpg::set_plan(h, lc, c.pick[2], c.pick[3], c.pick[4]);
                std::copy(c.pick, c.pick + 5, chosen);
                if (best_ms)
                    *best_ms = c.ms;
                return;
            }
        }
        // Scratch outputs on the device for the timing runs.
        lpg::DeviceBuffer g, H, E;
        g.ensure(std::max<size_t>(1, s.num_dofs) * sizeof(double));
        H.ensure(std::max<size_t>(1, s.nnz) * sizeof(double));
        const int num_meshes = lpg_tet_batch_num_meshes(h);
        E.ensure(std::max(1, num_meshes) * sizeof(double));
        const bool graph_was = h->graph.enabled;
        h->graph.enabled = false;
        double best = std::numeric_limits<double>::infinity();
        int32_t pick[5] = {h->plan.lc.lanes_per_vertex, h->plan.lc.waves, h->plan.vertices_per_tile,
                           h->plan.lds_budget, h->plan.persistent_grid};
        try {
            for (const Shape& sh : kShapes) {
                lpg::LaunchConfig lc;
                lc.lanes_per_vertex = sh.lanes;
                lc.waves = sh.waves;
                if (!lpg::launch_config_ok(h, lc, sh.vpt, sh.budget))
                    continue;
                for (int grid : kGrids) {
                    if (grid > 0 && num_meshes < 2)
                        break; // a single mesh uses one workgroup per tile
                    lpg::set_plan(h, lc, sh.vpt, sh.budget, grid);
                    const double ms = time_assembly(h, x_dev, hessian_kind, g.as<double>(), H.as<double>(),
                                                    num_meshes > 1 ? E.as<double>() : nullptr, repeats);
                    if (ms < best) {
                        best = ms;
                        pick[0] = sh.lanes;
                        pick[1] = sh.waves;
                        pick[2] = sh.vpt;
                        pick[3] = sh.budget;
                        pick[4] = grid;
                    }
                }
            }
        } catch (...) {
            h->graph.enabled = graph_was;
            g.release();
            H.release();
            E.release();
            throw;
        }
        lpg::LaunchConfig lc;
        lc.lanes_per_vertex = pick[0];
        lc.waves = pick[1];
        lpg::set_plan(h, lc, pick[2], pick[3], pick[4]);
        h->graph.enabled = graph_was;
        g.release();
        H.release();
        E.release();
        std::copy(pick, pick + 5, chosen);
        if (best_ms)
            *best_ms = best;
        TuneChoice c;
        std::copy(pick, pick + 5, c.pick);
        c.ms = best;
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tune_cache[key] = c;
    });
}

} // extern "C"
