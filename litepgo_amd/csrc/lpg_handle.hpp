// Host-side state of an energy handle and the helpers shared by the plain and batched C ABI entry points.
#pragma once

#include "litepgo_fem.h"
#include "lpg_internal.hpp"

#include <atomic>
#include <string>
#include <vector>

namespace lpg {

void hip_check(hipError_t err, const char* what);

// Owning device allocation that only grows.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    void ensure(size_t want);
    void release();
    template <class T> T* as() const {
        return static_cast<T*>(ptr);
    }
};

void upload(DeviceBuffer& buf, const void* src, size_t bytes);

// Tiles of the assembly launch for one launch configuration.
struct TilePlan {
    LaunchConfig lc;
    int vertices_per_tile = 0;
    int lds_budget = 0;
    int persistent_grid = 0; // 0: one workgroup per tile
    std::vector<int32_t> tile_vbeg;
    std::vector<int32_t> normal_ids, oversized_ids;
    TileCaps caps;
    DeviceBuffer d_tile_vbeg, d_normal_ids, d_oversized_ids;
    int num_tiles() const {
        return static_cast<int>(tile_vbeg.size()) - 1;
    }
};

// A captured evaluation sequence (hipGraph) and the call it was captured for.  Kernel nodes hold the buffer
// pointers and the launch plan, so the graph is replayed only for the same pointers, kinds and plan version.
struct GraphKey {
    const void* x = nullptr;
    const void* theta = nullptr;
    void* g = nullptr;
    void* H = nullptr;
    void* energies = nullptr;
    int want_g = 0;
    int kind = -2;
    int plan_version = -1;
    bool operator==(const GraphKey& o) const {
        return x == o.x && theta == o.theta && g == o.g && H == o.H && energies == o.energies &&
               want_g == o.want_g && kind == o.kind && plan_version == o.plan_version;
    }
};
struct GraphCache {
    bool enabled = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    GraphKey key;     // what `exec` was captured for
    GraphKey pending; // the last eager call: a repeat of it is captured
    void reset();
};

// Overlapping calls on one handle are rejected, as MeshEnergy's BusyScope does (mesh_energy.cpp:49).
struct BusyScope {
    std::atomic<bool>& flag;
    explicit BusyScope(std::atomic<bool>& f);
    ~BusyScope();
};

// Stores the message of a failed call for lpg_last_error and returns its status.
template <class F> lpg_status guarded(F&& body);

} // namespace lpg

struct lpg_tet_energy {
    lpg::MeshSetup setup;
    // Vertex ranges of the meshes whose energies are reported separately: {0, V} for a plain energy, the mesh
    // starts plus V for a batch.  Tiles never cross these boundaries.
    std::vector<int32_t> seg_vbeg;
    std::vector<int32_t> seg_tbeg; // tet offsets of the meshes of a batch ({0, T} for a plain energy)
    std::vector<int32_t> chunk_begin, seg_chunk; // energy reduction chunks (build_energy_chunks)
    int device = 0;
    bool device_ready = false;
    hipStream_t own_stream = nullptr;
    lpg::DeviceBuffer d_Xr, d_conn, d_theta, d_inc_ptr, d_inc, d_nbr_ptr, d_nbr, d_self_pos, d_row_ptr;
    lpg::DeviceBuffer d_scalars; // energy (double) + status + counter
    lpg::DeviceBuffer d_vertex_E, d_chunk_begin, d_chunk_sum, d_seg_chunk;
    lpg::DeviceBuffer d_seg_energy; // per-mesh energies of a batch evaluated as one energy
    lpg::DeviceBuffer d_proj_slot, d_proj_list, d_proj_H;
    lpg::DeviceBuffer d_x, d_g, d_H; // staging for the host-pointer API
    lpg::DeviceBuffer d_dx;          // staging of the step direction (host-pointer feasible step)
    lpg::DeviceBuffer d_seg_tbeg;    // seg_tbeg on the device (per-mesh feasible steps)
    lpg::DeviceBuffer d_seg_all;     // {0, T}: the whole handle as one segment
    lpg::DeviceBuffer d_fp;          // plastic parts (MeshSetup::fp) when some Fp differs from the identity
    lpg::DeviceBuffer d_fp_elem;     // per-incidence values of the plastic force response (3 doubles each)
    lpg::DeviceBuffer d_tet_inc;     // incidence index of (tet, slot), 4 per tet (built on first use)
    lpg::TilePlan plan;
    int plan_version = 0;
    lpg::GraphCache graph;
    std::atomic<bool> busy{false};
    int32_t last_indefinite = 0;
    int32_t proj_capacity = 0; // list entries d_proj_H holds (grown when an evaluation lists more tets)
    // An evaluation enqueued by lpg_tet_energy_eval_enqueue and not yet waited for (the handle stays busy).
    struct PendingEval {
        bool active = false;
        const double* x = nullptr;
        const double* theta = nullptr;
        int want_g = 0, kind = -1;
        double *g = nullptr, *H = nullptr, *energies = nullptr;
        hipStream_t stream = nullptr;
    } pending_eval;
    // The device status word and list counter are known to be zero (the last operation was an evaluation whose last
    // kernel read and zeroed them): the next single-mesh evaluation then needs no clearing kernel.
    bool status_clean = false;
    void* pinned = nullptr;     // energy + status + counter read-back
    void* pinned_dev = nullptr; // its device address (the last reduction kernel writes it directly), or null
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double time_ms[4] = {0, 0, 0, 0};
    int32_t timed_evals = 0;

    lpg::DeviceMesh mesh() const;
    double* d_energy() const {
        return d_scalars.as<double>();
    }
    unsigned* d_status() const {
        return reinterpret_cast<unsigned*>(d_scalars.as<char>() + 8);
    }
    unsigned* d_counter() const {
        return reinterpret_cast<unsigned*>(d_scalars.as<char>() + 12);
    }
    ~lpg_tet_energy();
};

namespace lpg {

// Uploads the mesh and creates the stream and read-back buffer on first use.  A plastic part rejected at
// construction fails the call unless the call brings its own plastic parts (snapshot_fp false).
void ensure_device(lpg_tet_energy* h, bool snapshot_fp = true);
// Launch shapes accepted by lpg_tet_energy_set_launch (and tried by the autotuner).
bool launch_config_ok(const lpg_tet_energy* h, const LaunchConfig& lc, int vertices_per_tile, int lds_budget);
// (Re)builds the tiles for a launch configuration.
void set_plan(lpg_tet_energy* h, const LaunchConfig& lc, int vertices_per_tile, int lds_budget,
              int persistent_grid);
// Device memory an evaluation uses besides its inputs and outputs: by default the handle's own buffers (allocated
// or grown on demand), or memory the caller lays out (the allocation-free entry point).
struct EvalScratch {
    void* scalars = nullptr;      // energy (double) + status + counter: 16 bytes
    double* vertex_E = nullptr;   // num_vertices
    double* chunk_sum = nullptr;  // energy reduction chunks
    double* seg_energy = nullptr; // per-segment energies of a batch evaluated as one energy
    int32_t* proj_slot = nullptr; // num_tets (Projected only)
    int32_t* proj_list = nullptr; // num_tets (Projected only)
    double* proj_H = nullptr;     // proj_capacity * kProjPerTet (Projected only)
    int32_t proj_capacity = 0;
};
// Checks arguments and enqueues one evaluation (no synchronisation).  energies (device) receives one energy per
// vertex segment (seg_vbeg); null means the scalar slot (plain energies).  ws: the scratch to use instead of the
// handle's buffers (then nothing is allocated, and the handle must already be on the device).
// readback: pinned memory (device address) that the last kernel fills with the energy, status and counter, in place
// of a copy; only for single-mesh handles (null: the caller copies the scalars itself).
// skip_clear: the status word and counter are already zero (readback only; the last kernel zeroes them again).
void enqueue_eval(lpg_tet_energy* h, const double* x, const double* theta, int want_g, int kind, double* g, double* H,
                  double* energies, hipStream_t stream, const EvalScratch* ws = nullptr, void* readback = nullptr,
                  bool skip_clear = false);
// Enqueues the read-back of status, counter and the first energy from `scalars` (null: the handle's own slot),
// synchronises, maps the status to an error.
double finish(lpg_tet_energy* h, hipStream_t stream, bool want_counter, const void* scalars = nullptr);
// One evaluation (eager, or replayed from the handle's captured graph when graph mode is on and the call repeats);
// returns the first energy (the plain energy; per-mesh energies go to `energies` when given).
double run_eval(lpg_tet_energy* h, const double* x, const double* theta, int want_g, int kind, double* g, double* H,
                double* energies, hipStream_t stream);
// Feasible step of x along dx (device arrays): one bound per segment into `alphas` (device; null: the handle's
// read-back slot), whole handle or per mesh; synchronises once and returns the first bound.
double feasible_step_call(lpg_tet_energy* h, const double* x, const double* dx, bool per_mesh, double* alphas,
                          hipStream_t s);
// Accumulates the per-phase event times of the last eager evaluation (timing mode).
void collect_timing(lpg_tet_energy* h);
// The status mapping alone (status word as read back).
void raise_status(unsigned status);
// The value, counter and status already read back into the handle's pinned slot (after a synchronisation).
double read_back_status(lpg_tet_energy* h, bool want_counter);
// Raised by finish()/read-back when the projection store was too small; run_eval grows it and evaluates again.
struct ProjectionOverflow : std::runtime_error {
    int32_t listed;
    explicit ProjectionOverflow(int32_t n)
        : std::runtime_error("MeshEnergy: projection store overflow"), listed(n) {}
};

extern thread_local std::string g_last_error;

template <class F> lpg_status guarded(F&& body) {
    try {
        body();
        return LPG_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return static_cast<lpg_status>(e.code);
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return LPG_DEVICE_ERROR;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return LPG_LOGIC_ERROR;
    }
}

} // namespace lpg
