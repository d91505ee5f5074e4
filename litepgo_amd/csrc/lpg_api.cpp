// C ABI of the plain energy (include/litepgo_fem.h): handle lifetime, validation, device buffers, launch
// sequencing and status mapping.
#include "litepgo_fem.h"
#include "litepgo_fem_ext.h"
#include "lpg_handle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>

namespace lpg {

thread_local std::string g_last_error;

void hip_check(hipError_t err, const char* what) {
    if (err != hipSuccess)
        fail(kDevice, std::string(what) + ": " + hipGetErrorString(err));
}

void DeviceBuffer::ensure(size_t want) {
    if (want <= bytes && ptr)
        return;
    release();
    if (want == 0)
        return;
    if (hipMalloc(&ptr, want) != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        fail(kDevice, "hipMalloc of " + std::to_string(want) + " bytes failed");
    }
    bytes = want;
}

void DeviceBuffer::release() {
    if (ptr)
        (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void upload(DeviceBuffer& buf, const void* src, size_t bytes) {
    buf.ensure(bytes);
    if (bytes && hipMemcpy(buf.ptr, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
        fail(kDevice, "hipMemcpy to device failed");
}

BusyScope::BusyScope(std::atomic<bool>& f) : flag(f) {
    bool expected = false;
    if (!flag.compare_exchange_strong(expected, true))
        fail(kBusy, "MeshEnergy is already in use by another call");
}
BusyScope::~BusyScope() {
    flag.store(false);
}

void GraphCache::reset() {
    if (exec)
        (void)hipGraphExecDestroy(exec);
    if (graph)
        (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key = GraphKey{};
    pending = GraphKey{};
}

} // namespace lpg

lpg::DeviceMesh lpg_tet_energy::mesh() const {
    lpg::DeviceMesh m;
    m.Xr = d_Xr.as<double>();
    m.conn = d_conn.as<int32_t>();
    m.theta = d_theta.as<double>();
    m.inc_ptr = d_inc_ptr.as<int32_t>();
    m.inc = d_inc.ptr;
    m.nbr_ptr = d_nbr_ptr.as<int32_t>();
    m.nbr = d_nbr.as<int32_t>();
    m.self_pos = d_self_pos.as<int32_t>();
    m.row_ptr = d_row_ptr.as<int32_t>();
    m.fp = setup.fp.empty() ? nullptr : d_fp.as<double>();
    m.num_vertices = setup.num_vertices;
    m.num_tets = setup.num_tets;
    m.wide = setup.wide;
    return m;
}

lpg_tet_energy::~lpg_tet_energy() {
    graph.reset();
    for (lpg::DeviceBuffer* b : {&d_Xr, &d_conn, &d_theta, &d_inc_ptr, &d_inc, &d_nbr_ptr, &d_nbr, &d_self_pos, &d_row_ptr, &d_scalars, &d_vertex_E, &d_chunk_begin,
                                 &d_chunk_sum, &d_seg_chunk, &d_proj_slot, &d_proj_list, &d_proj_H, &d_x, &d_g, &d_H,
                                 &d_dx, &d_seg_tbeg, &d_seg_all, &d_seg_energy, &d_fp, &d_fp_elem, &d_tet_inc,
                                 &plan.d_tile_vbeg, &plan.d_normal_ids, &plan.d_oversized_ids})
        b->release();
    if (pinned)
        (void)hipHostFree(pinned);
    for (hipEvent_t e : ev)
        if (e)
            (void)hipEventDestroy(e);
    if (own_stream)
        (void)hipStreamDestroy(own_stream);
}

namespace lpg {

bool launch_config_ok(const lpg_tet_energy* h, const LaunchConfig& lc, int vertices_per_tile, int lds_budget) {
    if (!launch_shape_supported(lc.lanes_per_vertex, lc.waves, h->setup.wide) || vertices_per_tile <= 0 ||
        vertices_per_tile > lc.waves * (64 / lc.lanes_per_vertex) || lds_budget <= 0)
        return false;
    // The budget plus the gradient and energy slots must fit one workgroup's LDS.
    return assemble_lds_bytes(lc, TileCaps{lds_budget}, false, 0) <= assemble_lds_limit(lc);
}

void set_plan(lpg_tet_energy* h, const LaunchConfig& lc, int vertices_per_tile, int lds_budget,
              int persistent_grid) {
    TilePlan& p = h->plan;
    ++h->plan_version;
    h->graph.reset();
    p.lc = lc;
    p.vertices_per_tile = vertices_per_tile;
    p.lds_budget = lds_budget;
    p.persistent_grid = persistent_grid;
    std::vector<int32_t> oversized;
    build_tiles(h->setup, TileConfig{vertices_per_tile, lds_budget}, p.tile_vbeg, p.caps, &oversized,
                &h->seg_vbeg);
    p.oversized_ids = oversized;
    p.normal_ids.clear();
    if (!oversized.empty()) {
        std::vector<char> over(p.num_tiles(), 0);
        for (int32_t t : oversized)
            over[t] = 1;
        for (int32_t t = 0; t < p.num_tiles(); ++t)
            if (!over[t])
                p.normal_ids.push_back(t);
    }
    upload(p.d_tile_vbeg, p.tile_vbeg.data(), p.tile_vbeg.size() * sizeof(int32_t));
    upload(p.d_normal_ids, p.normal_ids.data(), p.normal_ids.size() * sizeof(int32_t));
    upload(p.d_oversized_ids, p.oversized_ids.data(), p.oversized_ids.size() * sizeof(int32_t));
}

// Device state is created on the first call that needs the GPU, so construction, the pattern and the slot map
// work on hosts without one.
void ensure_device(lpg_tet_energy* h, bool snapshot_fp) {
    // A rejected plastic part fails every evaluation that uses it (the reference raises it per evaluated element).
    if (snapshot_fp && h->setup.fp_error)
        fail(h->setup.fp_error, h->setup.fp_error_msg);
    if (h->device_ready) {
        hip_check(hipSetDevice(h->device), "hipSetDevice");
        return;
    }
    hip_check(hipGetDevice(&h->device), "hipGetDevice");
    if (!h->own_stream)
        hip_check(hipStreamCreate(&h->own_stream), "hipStreamCreate");
    if (!h->pinned) {
        hip_check(hipHostMalloc(&h->pinned, 64, hipHostMallocDefault), "hipHostMalloc");
        if (hipHostGetDevicePointer(&h->pinned_dev, h->pinned, 0) != hipSuccess)
            h->pinned_dev = nullptr;
    }
    const auto& s = h->setup;
    upload(h->d_Xr, s.Xr.data(), s.Xr.size() * sizeof(double));
    upload(h->d_conn, s.conn.data(), s.conn.size() * sizeof(int32_t));
    upload(h->d_theta, s.theta.data(), s.theta.size() * sizeof(double));
    upload(h->d_inc_ptr, s.inc_ptr.data(), s.inc_ptr.size() * sizeof(int32_t));
    upload(h->d_nbr_ptr, s.nbr_ptr.data(), s.nbr_ptr.size() * sizeof(int32_t));
    upload(h->d_nbr, s.nbr.data(), s.nbr.size() * sizeof(int32_t));
    upload(h->d_row_ptr, s.row_ptr.data(), s.row_ptr.size() * sizeof(int32_t));
    upload(h->d_self_pos, s.self_pos.data(), s.self_pos.size() * sizeof(int32_t));
    if (!s.fp.empty())
        upload(h->d_fp, s.fp.data(), s.fp.size() * sizeof(double));
    // Device incidence records (IncNarrow / IncWide): slot a of v in the tet, and the column positions of the tet's
    // vertices at slots a^1, a^2, a^3.
    const size_t ninc = s.inc_tet.size();
    std::vector<uint32_t> words(ninc * (s.wide ? 4 : 2));
    for (int32_t v = 0; v < s.num_vertices; ++v)
        for (int32_t k = s.inc_ptr[v]; k < s.inc_ptr[v + 1]; ++k) {
            const int32_t* ids = s.conn.data() + 4 * static_cast<int64_t>(s.inc_tet[k]);
            const uint32_t a = ids[0] == v ? 0u : ids[1] == v ? 1u : ids[2] == v ? 2u : 3u;
            const uint16_t* cp = s.inc_colpos.data() + 4 * static_cast<int64_t>(k);
            const uint32_t c1 = cp[a ^ 1u], c2 = cp[a ^ 2u], c3 = cp[a ^ 3u];
            if (!s.wide) {
                words[2 * k] = static_cast<uint32_t>(s.inc_tet[k]);
                words[2 * k + 1] = c1 | (c2 << 8) | (c3 << 16) | (a << 24);
            } else {
                words[4 * k] = static_cast<uint32_t>(s.inc_tet[k]);
                words[4 * k + 1] = c1 | (c2 << 16);
                words[4 * k + 2] = c3 | (a << 16);
                words[4 * k + 3] = 0;
            }
        }
    upload(h->d_inc, words.data(), words.size() * sizeof(uint32_t));
    h->d_scalars.ensure(64);
    if (h->seg_vbeg.empty()) {
        h->seg_vbeg = {0, s.num_vertices};
        h->seg_tbeg = {0, s.num_tets};
    }
    build_energy_chunks(h->seg_vbeg, h->chunk_begin, h->seg_chunk);
    upload(h->d_seg_tbeg, h->seg_tbeg.data(), h->seg_tbeg.size() * sizeof(int32_t));
    {
        const int32_t all[2] = {0, s.num_tets};
        upload(h->d_seg_all, all, sizeof all);
    }
    upload(h->d_chunk_begin, h->chunk_begin.data(), h->chunk_begin.size() * sizeof(int32_t));
    upload(h->d_seg_chunk, h->seg_chunk.data(), h->seg_chunk.size() * sizeof(int32_t));
    h->d_chunk_sum.ensure(std::max<size_t>(1, h->chunk_begin.size()) * sizeof(double));
    h->d_seg_energy.ensure(std::max<size_t>(1, h->seg_chunk.size()) * sizeof(double));
    h->d_vertex_E.ensure(std::max<size_t>(1, s.num_vertices) * sizeof(double));
    if (h->plan.tile_vbeg.empty()) {
        // Default launch: 4 lanes per vertex, one wave per workgroup, 16 vertices per tile, segments up to
        // 4096 doubles (32 KiB) staged in LDS; one workgroup per tile for a single mesh, a persistent grid of
        // 2048 workgroups walking the tiles of a batch.
        LaunchConfig lc;
        lc.lanes_per_vertex = 4;
        lc.waves = 1;
        set_plan(h, lc, 16, 4096, h->seg_vbeg.size() > 2 ? 2048 : 0);
    }
    h->device_ready = true;
}

void raise_status(unsigned status) {
    if (status & kStatusNonFiniteX)
        fail(kInvalid, "Energy::eval: x must have num_dofs finite entries");
    if (status & kStatusBadTheta)
        fail(kInvalid, "NeoHookean requires finite lambda >= 0 and mu > 0");
    if (status & kStatusNonFiniteFp)
        fail(kInvalid, "MultiplicativeKinematics requires finite F and Fp");
    if (status & kStatusBadDetFp)
        fail(kDomain, "MultiplicativeKinematics requires finite det(Fp) > 0");
    if (status & kStatusDetF)
        fail(kDomain, "NeoHookean requires finite det(Fe) > 0");
    if (status & kStatusProjection)
        fail(kDomain, "MeshEnergy: non-finite projected Hessian");
    if (status & kStatusNonFiniteOut)
        fail(kDomain, "VolumetricElement3D: non-finite integrated outputs");
    if (status & kStatusNonFiniteStep)
        fail(kDomain, "VolumetricFemEnergy: non-finite input increment");
}

double finish(lpg_tet_energy* h, hipStream_t stream, bool want_counter, const void* scalars) {
    hip_check(hipMemcpyAsync(h->pinned, scalars ? scalars : h->d_scalars.ptr, 16, hipMemcpyDeviceToHost, stream),
              "read-back");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    double value;
    unsigned status, counter;
    std::memcpy(&value, h->pinned, 8);
    std::memcpy(&status, static_cast<char*>(h->pinned) + 8, 4);
    std::memcpy(&counter, static_cast<char*>(h->pinned) + 12, 4);
    if (want_counter)
        h->last_indefinite = static_cast<int32_t>(counter);
    if (status & kStatusProjOverflow)
        throw ProjectionOverflow(static_cast<int32_t>(counter));
    raise_status(status);
    return value;
}

void enqueue_eval(lpg_tet_energy* h, const double* x, const double* theta, int want_g, int kind, double* g, double* H,
                  double* energies, hipStream_t stream, const EvalScratch* ws, void* readback, bool skip_clear) {
    if (kind < -1 || kind > 1)
        fail(kInvalid, "Energy: invalid Hessian kind");
    if ((want_g && !g && h->setup.num_dofs > 0) || (kind >= 0 && !H && h->setup.nnz > 0) ||
        (!x && h->setup.num_dofs > 0))
        fail(kInvalid, "lpg_tet_energy_eval: missing input or output buffer");
    if (!theta && !h->setup.theta_valid && h->setup.num_tets > 0)
        fail(kInvalid, "NeoHookean requires finite lambda >= 0 and mu > 0");
    if (ws && !h->device_ready)
        fail(kLogic, "the mesh of this handle is not on the device yet: evaluate it once through an allocating "
                     "entry point first");
    ensure_device(h); // (only selects the device once the handle is there)
    EvalScratch own;
    if (!ws) {
        own.scalars = h->d_scalars.ptr;
        own.vertex_E = h->d_vertex_E.as<double>();
        own.chunk_sum = h->d_chunk_sum.as<double>();
        own.seg_energy = h->d_seg_energy.as<double>();
        if (kind == 1) {
            const size_t T = static_cast<size_t>(h->setup.num_tets);
            h->d_proj_slot.ensure(std::max<size_t>(1, T) * sizeof(int32_t));
            h->d_proj_list.ensure(std::max<size_t>(1, T) * sizeof(int32_t));
            // The low-rank store holds the listed (not certified positive definite) tets only: 1/16 of the tets to
            // start with, grown to the listed count when an evaluation lists more (run_eval repeats it then).
            if (h->proj_capacity == 0)
                h->proj_capacity = static_cast<int32_t>(std::min<size_t>(T, std::max<size_t>(4096, T / 16)));
            h->d_proj_H.ensure(std::max<size_t>(1, h->proj_capacity) * kProjPerTet * sizeof(double));
            own.proj_slot = h->d_proj_slot.as<int32_t>();
            own.proj_list = h->d_proj_list.as<int32_t>();
            own.proj_H = h->d_proj_H.as<double>();
            own.proj_capacity = h->proj_capacity;
        }
        ws = &own;
    }
    EvalArgs a;
    a.x = x;
    a.theta = theta;
    a.g = want_g ? g : nullptr;
    a.H = kind >= 0 ? H : nullptr;
    a.want_g = want_g;
    a.kind = kind;
    a.vertex_E = ws->vertex_E;
    a.energy = static_cast<double*>(ws->scalars);
    a.status = reinterpret_cast<unsigned*>(static_cast<char*>(ws->scalars) + 8);
    a.counter = reinterpret_cast<unsigned*>(static_cast<char*>(ws->scalars) + 12);
    const DeviceMesh m = h->mesh();
    if (kind == 1) {
        a.proj_slot = ws->proj_slot;
        a.proj_list = ws->proj_list;
        a.proj_H = ws->proj_H;
        a.proj_capacity = ws->proj_capacity;
    }
    if (h->timing)
        hip_check(hipEventRecord(h->ev[0], stream), "hipEventRecord");
    if (!readback)
        h->status_clean = false;
    if (!(readback && skip_clear))
        launch_clear(a.status, a.counter, stream);
    if (kind == 1) {
        launch_project_prepass(m, a, stream);
        launch_project_jacobi(m, a, stream);
    }
    if (h->timing)
        hip_check(hipEventRecord(h->ev[1], stream), "hipEventRecord");
    const TilePlan& p = h->plan;
    if (p.oversized_ids.empty()) {
        launch_assemble(m, a, p.lc, p.d_tile_vbeg.as<int32_t>(), nullptr, p.num_tiles(), p.caps, false,
                        p.persistent_grid, stream);
    } else {
        launch_assemble(m, a, p.lc, p.d_tile_vbeg.as<int32_t>(), p.d_normal_ids.as<int32_t>(),
                        static_cast<int>(p.normal_ids.size()), p.caps, false, p.persistent_grid, stream);
        launch_assemble(m, a, p.lc, p.d_tile_vbeg.as<int32_t>(), p.d_oversized_ids.as<int32_t>(),
                        static_cast<int>(p.oversized_ids.size()), TileCaps{}, true, 0, stream);
    }
    if (h->timing)
        hip_check(hipEventRecord(h->ev[2], stream), "hipEventRecord");
    const int segments = static_cast<int>(h->seg_chunk.size()) - 1;
    // Per-segment energies go to `energies`; without it, a batch's segment energies go to scratch and their total
    // to the handle's energy slot (the single-energy contract of the union).
    double* seg_out = energies ? energies : (segments > 1 ? ws->seg_energy : a.energy);
    launch_energy_reduce(a.vertex_E, h->d_chunk_begin.as<int32_t>(), static_cast<int>(h->chunk_begin.size()) - 1,
                         ws->chunk_sum, h->d_seg_chunk.as<int32_t>(), segments, seg_out, a.status, stream, a.counter,
                         segments == 1 ? readback : nullptr);
    if (!energies && segments > 1)
        launch_energy_total(seg_out, segments, a.energy, stream);
    if (readback && segments != 1)
        launch_readback(a.energy, a.status, a.counter, readback, stream);
    if (h->timing)
        hip_check(hipEventRecord(h->ev[3], stream), "hipEventRecord");
}

void collect_timing(lpg_tet_energy* h) {
    if (!h->timing)
        return;
    float t01 = 0, t12 = 0, t23 = 0, t03 = 0;
    hip_check(hipEventElapsedTime(&t01, h->ev[0], h->ev[1]), "hipEventElapsedTime");
    hip_check(hipEventElapsedTime(&t12, h->ev[1], h->ev[2]), "hipEventElapsedTime");
    hip_check(hipEventElapsedTime(&t23, h->ev[2], h->ev[3]), "hipEventElapsedTime");
    hip_check(hipEventElapsedTime(&t03, h->ev[0], h->ev[3]), "hipEventElapsedTime");
    h->time_ms[0] += t01;
    h->time_ms[1] += t12;
    h->time_ms[2] += t23;
    h->time_ms[3] += t03;
    ++h->timed_evals;
}

double read_back_status(lpg_tet_energy* h, bool want_counter);

namespace {

double read_back(lpg_tet_energy* h, bool want_counter) {
    return read_back_status(h, want_counter);
}

} // namespace

double read_back_status(lpg_tet_energy* h, bool want_counter) {
    double value;
    unsigned status, counter;
    std::memcpy(&value, h->pinned, 8);
    std::memcpy(&status, static_cast<char*>(h->pinned) + 8, 4);
    std::memcpy(&counter, static_cast<char*>(h->pinned) + 12, 4);
    if (want_counter)
        h->last_indefinite = static_cast<int32_t>(counter);
    if (status & kStatusProjOverflow)
        throw ProjectionOverflow(static_cast<int32_t>(counter));
    raise_status(status);
    return value;
}

namespace {
double run_eval_once(lpg_tet_energy* h, const double* x, const double* theta, int want_g, int kind, double* g,
                     double* H, double* energies, hipStream_t stream);
} // namespace

double run_eval(lpg_tet_energy* h, const double* x, const double* theta, int want_g, int kind, double* g, double* H,
                double* energies, hipStream_t stream) {
    try {
        return run_eval_once(h, x, theta, want_g, kind, g, H, energies, stream);
    } catch (const ProjectionOverflow& o) {
        // More tets were listed for projection than the store holds: grow it to the listed count (plus an eighth)
        // and evaluate again; a captured graph holds the old store, so it is dropped.
        const int64_t T = h->setup.num_tets;
        h->proj_capacity = static_cast<int32_t>(std::min<int64_t>(T, int64_t{o.listed} + o.listed / 8 + 1024));
        h->graph.reset();
        return run_eval_once(h, x, theta, want_g, kind, g, H, energies, stream);
    }
}

namespace {

// The fused read-back's status word (pinned bytes 8..11) is set to all ones before a graph replay; the replay's
// last kernel writes the real status there after the energy and the counter.  Waiting for it to change returns as
// soon as that kernel has written them, without the stream synchronisation's completion round trip; every earlier
// kernel of the replay (the gradient and Hessian writers) has finished by then.  A stream that ends without the
// word changing (a device error) is reported through the stream's own status.
constexpr unsigned kReadbackArmed = 0xffffffffu;

void arm_readback(lpg_tet_energy* h) {
    __atomic_store_n(reinterpret_cast<unsigned*>(static_cast<char*>(h->pinned) + 8), kReadbackArmed, __ATOMIC_RELEASE);
}

void wait_readback(lpg_tet_energy* h, hipStream_t stream) {
    const unsigned* word = reinterpret_cast<const unsigned*>(static_cast<const char*>(h->pinned) + 8);
    for (unsigned spin = 1; __atomic_load_n(word, __ATOMIC_ACQUIRE) == kReadbackArmed; ++spin) {
        if ((spin & 4095u) != 0u)
            continue;
        const hipError_t q = hipStreamQuery(stream);
        if (q == hipErrorNotReady)
            continue;
        hip_check(q, "hipStreamQuery");
        if (__atomic_load_n(word, __ATOMIC_ACQUIRE) == kReadbackArmed)
            fail(kDevice, "evaluation: the stream finished without writing its read-back");
        break;
    }
}

double run_eval_once(lpg_tet_energy* h, const double* x, const double* theta, int want_g, int kind, double* g,
                     double* H, double* energies, hipStream_t stream) {
    ensure_device(h);
    // The evaluation hands its scalars to the host from its last kernel (no copy node): the energy reduction's for a
    // single mesh, a read-back kernel after it for several.
    void* fused = h->pinned_dev;
    GraphCache& gc = h->graph;
    if (gc.enabled && !h->timing) {
        GraphKey k;
        k.x = x;
        k.theta = theta;
        k.g = want_g ? g : nullptr;
        k.H = kind >= 0 ? H : nullptr;
        k.energies = energies;
        k.want_g = want_g;
        k.kind = kind;
        k.plan_version = h->plan_version;
        if (!(gc.exec && gc.key == k) && gc.pending == k) {
            // The call repeats: capture the sequence (kernels and the status read-back) on the handle's stream.
            // Allocations and one-time attribute calls already happened in the eager call before.
            gc.reset();
            hip_check(hipStreamBeginCapture(h->own_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                // With the fused read-back the graph holds no clearing kernel: its last kernel leaves the status
                // word and counter zero, and a replay after anything else clears them first (below).
                enqueue_eval(h, x, theta, want_g, kind, g, H, energies, h->own_stream, nullptr, fused, fused != nullptr);
                if (!fused)
                    hip_check(hipMemcpyAsync(h->pinned, h->d_scalars.ptr, 16, hipMemcpyDeviceToHost, h->own_stream),
                              "read-back");
            } catch (...) {
                hipGraph_t broken = nullptr;
                (void)hipStreamEndCapture(h->own_stream, &broken);
                if (broken)
                    (void)hipGraphDestroy(broken);
                throw;
            }
            hip_check(hipStreamEndCapture(h->own_stream, &gc.graph), "hipStreamEndCapture");
            hip_check(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            gc.key = k;
        }
        if (gc.exec && gc.key == k) {
            if (fused && !h->status_clean)
                launch_clear(h->d_status(), h->d_counter(), stream);
            h->status_clean = false;
            if (fused)
                arm_readback(h);
            hip_check(hipGraphLaunch(gc.exec, stream), "hipGraphLaunch");
            if (fused)
                wait_readback(h, stream);
            else
                hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            h->status_clean = fused != nullptr;
            return read_back(h, kind == 1);
        }
        gc.pending = k;
    }
    const bool skip_clear = fused && h->status_clean;
    h->status_clean = false;
    enqueue_eval(h, x, theta, want_g, kind, g, H, energies, stream, nullptr, fused, skip_clear);
    double value;
    if (fused) {
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        h->status_clean = true;
        value = read_back(h, kind == 1);
    } else {
        value = finish(h, stream, kind == 1);
    }
    collect_timing(h);
    return value;
}

void eval_device(lpg_tet_energy* h, const double* x, const double* theta, int want_g, int kind, double* value,
                 double* g, double* H, hipStream_t stream) {
    if (!value)
        fail(kInvalid, "lpg_tet_energy_eval: value pointer is null");
    *value = run_eval(h, x, theta, want_g, kind, g, H, nullptr, stream);
}

template <class Launch>
lpg_status derivative_call(lpg_tet_energy* h, const double* x, const double* theta, void* stream, Launch&& launch,
                           bool snapshot_fp = true) {
    return guarded([&] {
        if (!h)
            fail(kInvalid, "null handle");
        BusyScope scope(h->busy);
        if (!x && h->setup.num_dofs > 0)
            fail(kInvalid, "Energy::eval: x must have num_dofs finite entries");
        if (!theta && !h->setup.theta_valid && h->setup.num_tets > 0)
            fail(kInvalid, "NeoHookean requires finite lambda >= 0 and mu > 0");
        ensure_device(h, snapshot_fp);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        h->status_clean = false;
        launch_clear(h->d_status(), h->d_counter(), s);
        launch(h->mesh(), h->d_status(), s);
        (void)finish(h, s, false);
    });
}

} // namespace

// Energy::max_feasible_step (energy.cpp:35-45) over MeshEnergy::feasible_step (mesh_energy.cpp:88-107): one bound per
// segment (the whole handle, or each mesh of a batch), written to `alphas` on the device; returns the first.
double feasible_step_call(lpg_tet_energy* h, const double* x, const double* dx, bool per_mesh, double* alphas,
                          hipStream_t s) {
    const auto& st = h->setup;
    if ((!x || !dx) && st.num_dofs > 0)
        fail(kInvalid, "Energy::max_feasible_step: x and dx must have num_dofs finite entries");
    if (!st.theta_valid && st.num_tets > 0)
        fail(kInvalid, "NeoHookean requires finite lambda >= 0 and mu > 0");
    ensure_device(h);
    h->status_clean = false;
    const int segments = per_mesh ? static_cast<int>(h->seg_tbeg.size()) - 1 : 1;
    double* out = alphas ? alphas : h->d_energy();
    launch_feasible_step(h->mesh(), x, dx, per_mesh ? h->d_seg_tbeg.as<int32_t>() : h->d_seg_all.as<int32_t>(),
                         segments, out, h->d_status(), s);
    if (alphas) {
        // Copy the first bound into the read-back slot so finish() returns it.
        hip_check(hipMemcpyAsync(h->d_energy(), alphas, sizeof(double), hipMemcpyDeviceToDevice, s), "step copy");
    }
    return finish(h, s, false);
}

} // namespace lpg

using lpg::fail;
using lpg::guarded;

extern "C" {

size_t lpg_last_error(char* buf, size_t len) {
    const std::string& msg = lpg::g_last_error;
    if (buf && len > 0) {
        const size_t n = std::min(len - 1, msg.size());
        std::memcpy(buf, msg.data(), n);
        buf[n] = 0;
    }
    return msg.size();
}

lpg_status lpg_tet_energy_create(int32_t num_vertices, int32_t num_tets, const int32_t* tets,
                                 const double* rest_positions, const double* theta, const double* Fp,
                                 int32_t batch_size, lpg_tet_energy** out) {
    return guarded([&] {
        if (!out)
            fail(lpg::kInvalid, "lpg_tet_energy_create: out is null");
        *out = nullptr;
        if (batch_size <= 0)
            fail(lpg::kInvalid, "MeshEnergy: invalid batch size");
        auto h = std::make_unique<lpg_tet_energy>();
        const int nthreads = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        lpg::build_mesh_setup(h->setup, num_vertices, num_tets, tets, rest_positions, theta, Fp, nthreads);
        *out = h.release();
    });
}

void lpg_tet_energy_destroy(lpg_tet_energy* energy) {
    delete energy;
}

int32_t lpg_tet_energy_num_dofs(const lpg_tet_energy* h) {
    return h ? h->setup.num_dofs : 0;
}
int32_t lpg_tet_energy_num_tets(const lpg_tet_energy* h) {
    return h ? h->setup.num_tets : 0;
}
int32_t lpg_tet_energy_nnz(const lpg_tet_energy* h) {
    return h ? h->setup.nnz : 0;
}

lpg_status lpg_tet_energy_pattern(const lpg_tet_energy* h, int32_t* row_ptr, int32_t* col_idx) {
    return guarded([&] {
        if (!h || !row_ptr || (!col_idx && h->setup.nnz > 0))
            fail(lpg::kInvalid, "lpg_tet_energy_pattern: null argument");
        std::copy(h->setup.row_ptr.begin(), h->setup.row_ptr.end(), row_ptr);
        std::copy(h->setup.col_idx.begin(), h->setup.col_idx.end(), col_idx);
    });
}

lpg_status lpg_tet_energy_slot_map(const lpg_tet_energy* h, int32_t* offsets) {
    return guarded([&] {
        if (!h || (!offsets && h->setup.num_tets > 0))
            fail(lpg::kInvalid, "lpg_tet_energy_slot_map: null argument");
        lpg::build_slot_map(h->setup, offsets);
    });
}

lpg_status lpg_tet_energy_eval(lpg_tet_energy* h, const double* x, int32_t want_gradient, int32_t hessian_kind,
                               double* value, double* gradient, double* hessian_values) {
    return guarded([&] {
        if (!h)
            fail(lpg::kInvalid, "lpg_tet_energy_eval: null handle");
        lpg::BusyScope scope(h->busy);
        const auto& s = h->setup;
        if (!x && s.num_dofs > 0)
            fail(lpg::kInvalid, "Energy::eval: x must have num_dofs finite entries");
        for (int32_t i = 0; i < s.num_dofs; ++i)
            if (!std::isfinite(x[i]))
                fail(lpg::kInvalid, "Energy::eval: x must have num_dofs finite entries");
        if ((want_gradient && !gradient && s.num_dofs > 0) || (hessian_kind >= 0 && !hessian_values && s.nnz > 0))
            fail(lpg::kInvalid, "lpg_tet_energy_eval: missing output buffer");
        lpg::ensure_device(h);
        lpg::upload(h->d_x, x, s.num_dofs * sizeof(double));
        if (want_gradient)
            h->d_g.ensure(s.num_dofs * sizeof(double));
        if (hessian_kind >= 0)
            h->d_H.ensure(static_cast<size_t>(s.nnz) * sizeof(double));
        lpg::eval_device(h, h->d_x.as<double>(), nullptr, want_gradient, hessian_kind, value, h->d_g.as<double>(),
                         h->d_H.as<double>(), h->own_stream);
        if (want_gradient && s.num_dofs > 0)
            lpg::hip_check(hipMemcpy(gradient, h->d_g.ptr, s.num_dofs * sizeof(double), hipMemcpyDeviceToHost),
                           "gradient read-back");
        if (hessian_kind >= 0 && s.nnz > 0)
            lpg::hip_check(hipMemcpy(hessian_values, h->d_H.ptr, static_cast<size_t>(s.nnz) * sizeof(double),
                                     hipMemcpyDeviceToHost),
                           "Hessian read-back");
    });
}

lpg_status lpg_tet_energy_eval_device(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                      int32_t want_gradient, int32_t hessian_kind, double* value,
                                      double* gradient_dev, double* hessian_values_dev, void* stream) {
    return guarded([&] {
        if (!h)
            fail(lpg::kInvalid, "lpg_tet_energy_eval_device: null handle");
        lpg::BusyScope scope(h->busy);
        lpg::eval_device(h, x_dev, theta_dev, want_gradient, hessian_kind, value, gradient_dev, hessian_values_dev,
                         static_cast<hipStream_t>(stream));
    });
}

int32_t lpg_tet_energy_last_indefinite_count(const lpg_tet_energy* h) {
    return h ? h->last_indefinite : 0;
}

lpg_status lpg_tet_energy_dE_dtheta_device(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                           double* out_dev, void* stream) {
    return lpg::derivative_call(h, x_dev, theta_dev, stream,
                                [&](const lpg::DeviceMesh& m, unsigned* st, hipStream_t s) {
                                    lpg::launch_dE_dtheta(m, x_dev, theta_dev, out_dev, st, s);
                                });
}

lpg_status lpg_tet_energy_force_vjp_theta_device(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                                 const double* w_dev, double* out_dev, void* stream) {
    return lpg::derivative_call(h, x_dev, theta_dev, stream,
                                [&](const lpg::DeviceMesh& m, unsigned* st, hipStream_t s) {
                                    lpg::launch_force_vjp_theta(m, x_dev, theta_dev, w_dev, out_dev, st, s);
                                });
}

lpg_status lpg_tet_energy_force_jvp_theta_device(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                                 const double* dtheta_dev, double* out_dev, void* stream) {
    return lpg::derivative_call(h, x_dev, theta_dev, stream,
                                [&](const lpg::DeviceMesh& m, unsigned* st, hipStream_t s) {
                                    lpg::launch_force_jvp_theta(m, x_dev, theta_dev, dtheta_dev, out_dev, st, s);
                                });
}

namespace lpg {
namespace {
// Scratch of the plastic force response: one 3-vector per incidence, written by tet through the incidence index of
// each (tet, slot), then summed per vertex over its contiguous incidence range.
double* fp_scratch(lpg_tet_energy* h) {
    const auto& s = h->setup;
    const size_t T = static_cast<size_t>(s.num_tets);
    if (T > 0 && h->d_tet_inc.bytes < 4 * T * sizeof(int32_t)) {
        std::vector<int32_t> tet_inc(4 * T);
        for (int32_t v = 0; v < s.num_vertices; ++v)
            for (int32_t k = s.inc_ptr[v]; k < s.inc_ptr[v + 1]; ++k) {
                const int64_t e = s.inc_tet[k];
                const int32_t* ids = s.conn.data() + 4 * e;
                const int a = ids[0] == v ? 0 : ids[1] == v ? 1 : ids[2] == v ? 2 : 3;
                tet_inc[4 * e + a] = k;
            }
        upload(h->d_tet_inc, tet_inc.data(), tet_inc.size() * sizeof(int32_t));
    }
    h->d_fp_elem.ensure(12 * T * sizeof(double));
    return h->d_fp_elem.as<double>();
}
} // namespace
} // namespace lpg

lpg_status lpg_tet_energy_force_fp_device(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                          const double* fp_dev, double* out_dev, void* stream) {
    return lpg::derivative_call(h, x_dev, theta_dev, stream, [&](const lpg::DeviceMesh& m, unsigned* st, hipStream_t s) {
        double* scratch = lpg::fp_scratch(h);
        lpg::launch_force_fp(m, x_dev, theta_dev, fp_dev, nullptr, nullptr, h->d_tet_inc.as<int32_t>(), scratch,
                             out_dev, st, s);
    }, fp_dev == nullptr);
}

lpg_status lpg_tet_energy_force_vjp_fp_device(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                              const double* fp_dev, const double* w_dev, double* out_dev,
                                              void* stream) {
    return lpg::derivative_call(h, x_dev, theta_dev, stream, [&](const lpg::DeviceMesh& m, unsigned* st, hipStream_t s) {
        if (!w_dev && m.num_tets > 0)
            lpg::fail(lpg::kInvalid, "force_vjp_fp: w must have num_dofs entries");
        lpg::launch_force_fp(m, x_dev, theta_dev, fp_dev, w_dev, nullptr, nullptr, nullptr, out_dev, st, s);
    }, fp_dev == nullptr);
}

lpg_status lpg_tet_energy_force_jvp_fp_device(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                              const double* fp_dev, const double* dfp_dev, double* out_dev,
                                              void* stream) {
    return lpg::derivative_call(h, x_dev, theta_dev, stream, [&](const lpg::DeviceMesh& m, unsigned* st, hipStream_t s) {
        if (!dfp_dev && m.num_tets > 0)
            lpg::fail(lpg::kInvalid, "force_jvp_fp: dFp must have num_tets x 9 entries");
        double* scratch = lpg::fp_scratch(h);
        lpg::launch_force_fp(m, x_dev, theta_dev, fp_dev, nullptr, dfp_dev, h->d_tet_inc.as<int32_t>(), scratch,
                             out_dev, st, s);
    }, fp_dev == nullptr);
}

lpg_status lpg_tet_energy_hvp_device(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                     const double* v_dev, double* out_dev, void* stream) {
    return lpg::derivative_call(h, x_dev, theta_dev, stream,
                                [&](const lpg::DeviceMesh& m, unsigned* st, hipStream_t s) {
                                    lpg::launch_hvp(m, x_dev, theta_dev, v_dev, out_dev, st, s);
                                });
}

lpg_status lpg_tet_energy_vjp_device(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                     const double* gE_dev, const double* w_dev, double* gx_dev, double* gtheta_dev,
                                     void* stream) {
    return lpg::derivative_call(h, x_dev, theta_dev, stream,
                                [&](const lpg::DeviceMesh& m, unsigned* st, hipStream_t s) {
                                    lpg::launch_vjp(m, x_dev, theta_dev, gE_dev, w_dev, gx_dev, gtheta_dev, st, s);
                                });
}

lpg_status lpg_tet_energy_max_feasible_step(lpg_tet_energy* h, const double* x, const double* dx, double* alpha) {
    return guarded([&] {
        if (!h || !alpha)
            fail(lpg::kInvalid, "lpg_tet_energy_max_feasible_step: null argument");
        lpg::BusyScope scope(h->busy);
        const auto& s = h->setup;
        if ((!x || !dx) && s.num_dofs > 0)
            fail(lpg::kInvalid, "Energy::max_feasible_step: x and dx must have num_dofs finite entries");
        for (int32_t i = 0; i < s.num_dofs; ++i)
            if (!std::isfinite(x[i]))
                fail(lpg::kInvalid, "Energy::max_feasible_step: x must have num_dofs finite entries");
        for (int32_t i = 0; i < s.num_dofs; ++i)
            if (!std::isfinite(dx[i]))
                fail(lpg::kInvalid, "Energy::max_feasible_step: dx must have num_dofs finite entries");
        lpg::ensure_device(h);
        lpg::upload(h->d_x, x, s.num_dofs * sizeof(double));
        lpg::upload(h->d_dx, dx, s.num_dofs * sizeof(double));
        *alpha = lpg::feasible_step_call(h, h->d_x.as<double>(), h->d_dx.as<double>(), false, nullptr,
                                         h->own_stream);
    });
}

lpg_status lpg_tet_energy_max_feasible_step_device(lpg_tet_energy* h, const double* x_dev, const double* dx_dev,
                                                   double* alpha, void* stream) {
    return guarded([&] {
        if (!h || !alpha)
            fail(lpg::kInvalid, "lpg_tet_energy_max_feasible_step_device: null argument");
        lpg::BusyScope scope(h->busy);
        *alpha = lpg::feasible_step_call(h, x_dev, dx_dev, false, nullptr, static_cast<hipStream_t>(stream));
    });
}

// ---- extensions (litepgo_fem_ext.h) ----
lpg_status lpg_lame_device(const double* E_dev, const double* nu_dev, int64_t n, double* theta_dev, void* stream) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!E_dev || !nu_dev || !theta_dev)))
            fail(lpg::kInvalid, "lpg_lame_device: invalid argument");
        lpg::launch_lame(E_dev, nu_dev, n, theta_dev, static_cast<hipStream_t>(stream));
    });
}

lpg_status lpg_lame_vjp_device(const double* E_dev, const double* nu_dev, const double* gtheta_dev, int64_t n,
                               double* gE_dev, double* gnu_dev, void* stream) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!E_dev || !nu_dev || !gtheta_dev)))
            fail(lpg::kInvalid, "lpg_lame_vjp_device: invalid argument");
        lpg::launch_lame_vjp(E_dev, nu_dev, gtheta_dev, n, gE_dev, gnu_dev, static_cast<hipStream_t>(stream));
    });
}

lpg_status lpg_tet_energy_set_launch(lpg_tet_energy* h, int32_t lanes_per_vertex, int32_t waves,
                                     int32_t vertices_per_tile, int32_t lds_budget_doubles) {
    return guarded([&] {
        if (!h)
            fail(lpg::kInvalid, "null handle");
        lpg::BusyScope scope(h->busy);
        lpg::LaunchConfig lc;
        lc.lanes_per_vertex = lanes_per_vertex;
        lc.waves = waves;
        if (!lpg::launch_config_ok(h, lc, vertices_per_tile, lds_budget_doubles))
            fail(lpg::kInvalid, "lpg_tet_energy_set_launch: unsupported configuration");
        lpg::ensure_device(h);
        lpg::set_plan(h, lc, vertices_per_tile, lds_budget_doubles, 0);
    });
}

lpg_status lpg_tet_energy_get_launch(const lpg_tet_energy* h, int32_t* cfg4) {
    return guarded([&] {
        if (!h || !cfg4)
            fail(lpg::kInvalid, "null argument");
        cfg4[0] = h->plan.lc.lanes_per_vertex;
        cfg4[1] = h->plan.lc.waves;
        cfg4[2] = h->plan.vertices_per_tile;
        cfg4[3] = h->plan.lds_budget;
    });
}

lpg_status lpg_tet_energy_set_timing(lpg_tet_energy* h, int32_t enable) {
    return guarded([&] {
        if (!h)
            fail(lpg::kInvalid, "null handle");
        lpg::BusyScope scope(h->busy);
        lpg::ensure_device(h);
        if (enable && !h->ev[0])
            for (hipEvent_t& e : h->ev)
                lpg::hip_check(hipEventCreate(&e), "hipEventCreate");
        h->timing = enable != 0;
        for (double& t : h->time_ms)
            t = 0;
        h->timed_evals = 0;
    });
}

lpg_status lpg_tet_energy_kernel_times(const lpg_tet_energy* h, double* ms, int32_t* evals) {
    return guarded([&] {
        if (!h || !ms || !evals)
            fail(lpg::kInvalid, "null argument");
        for (int k = 0; k < 4; ++k)
            ms[k] = h->time_ms[k];
        *evals = h->timed_evals;
    });
}

int32_t lpg_tet_energy_max_degree(const lpg_tet_energy* h) {
    return h ? h->setup.max_degree : 0;
}

lpg_status lpg_tet_energy_eval_enqueue(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                       int32_t want_gradient, int32_t hessian_kind, double* energies_dev,
                                       double* gradient_dev, double* hessian_values_dev, void* stream) {
    return guarded([&] {
        if (!h)
            fail(lpg::kInvalid, "lpg_tet_energy_eval_enqueue: null handle");
        // The handle stays busy until lpg_tet_energy_eval_wait.
        bool expected = false;
        if (!h->busy.compare_exchange_strong(expected, true))
            fail(lpg::kBusy, "MeshEnergy is already in use by another call");
        try {
            const hipStream_t s = static_cast<hipStream_t>(stream);
            lpg::enqueue_eval(h, x_dev, theta_dev, want_gradient, hessian_kind, gradient_dev, hessian_values_dev,
                              energies_dev, s);
            lpg::hip_check(hipMemcpyAsync(h->pinned, h->d_scalars.ptr, 16, hipMemcpyDeviceToHost, s), "read-back");
            h->pending_eval = {true, x_dev, theta_dev, want_gradient, hessian_kind, gradient_dev, hessian_values_dev,
                               energies_dev, s};
        } catch (...) {
            h->busy.store(false);
            throw;
        }
    });
}

lpg_status lpg_tet_energy_eval_wait(lpg_tet_energy* h, double* value) {
    return guarded([&] {
        if (!h)
            fail(lpg::kInvalid, "lpg_tet_energy_eval_wait: null handle");
        if (!h->pending_eval.active)
            fail(lpg::kLogic, "lpg_tet_energy_eval_wait: no evaluation was enqueued on this handle");
        const auto e = h->pending_eval;
        h->pending_eval.active = false;
        // Release the handle however the wait ends.
        struct Release {
            std::atomic<bool>& f;
            ~Release() { f.store(false); }
        } release{h->busy};
        double v;
        try {
            lpg::hip_check(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
            v = lpg::read_back_status(h, e.kind == 1);
        } catch (const lpg::ProjectionOverflow&) {
            // The projection store was too small for this state: evaluate again with a grown store (synchronous).
            v = lpg::run_eval(h, e.x, e.theta, e.want_g, e.kind, e.g, e.H, e.energies, e.stream);
        }
        if (value && !e.energies)
            *value = v; // with an energies buffer the energies are there only
    });
}

} // extern "C"

// ---------------------------------------------------------------------------------------------------------------
// Allocation-free evaluation: the per-evaluation scratch (status and energy slot, vertex energies, reduction
// partials, projection slots, list and low-rank store) lives in one static device array per GPU instead of the
// handle's allocated buffers.
// ---------------------------------------------------------------------------------------------------------------
namespace {

constexpr size_t kNoAllocScratchBytes = size_t{64} << 20; // 64 MiB per device, held until the process ends
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {}; // device address of `scratch` on each device (null until looked up)

// The device address of `scratch` on `device`, looked up once per device with that device current.  The host-side
// name `scratch` is not a device address; only the pointer looked up here is passed to kernels and copies.
char* scratch_on(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device < 0 || device >= kMaxDevices)
        lpg::fail(lpg::kDevice, "device index out of range for the allocation-free scratch");
    if (!g_scratch_ptr[device]) {
        int current = 0;
        lpg::hip_check(hipGetDevice(&current), "hipGetDevice");
        if (current != device)
            lpg::hip_check(hipSetDevice(device), "hipSetDevice");
        void* ptr = nullptr;
        const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (current != device)
            (void)hipSetDevice(current);
        lpg::hip_check(err, "hipGetSymbolAddress(scratch)");
        if (!ptr)
            lpg::fail(lpg::kDevice, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[device] = ptr;
    }
    return static_cast<char*>(g_scratch_ptr[device]);
}

size_t round256(size_t n) {
    return (n + 255) & ~size_t{255};
}

} // namespace

extern "C" {

lpg_status lpg_tet_energy_eval_device_noalloc(lpg_tet_energy* h, const double* x_dev, const double* theta_dev,
                                              int32_t want_gradient, int32_t hessian_kind, double* value,
                                              double* gradient_dev, double* hessian_values_dev, void* stream) {
    return guarded([&] {
        if (!h)
            fail(lpg::kInvalid, "lpg_tet_energy_eval_device_noalloc: null handle");
        if (!value)
            fail(lpg::kInvalid, "lpg_tet_energy_eval: value pointer is null");
        lpg::BusyScope scope(h->busy);
        if (!h->device_ready)
            fail(lpg::kLogic, "lpg_tet_energy_eval_device_noalloc: the mesh of this handle is not on the device "
                              "yet: evaluate it once through an allocating entry point first");
        const size_t V = static_cast<size_t>(std::max(1, h->setup.num_vertices));
        const size_t T = static_cast<size_t>(std::max(1, h->setup.num_tets));
        const size_t chunks = std::max<size_t>(1, h->chunk_begin.size());
        const size_t segs = std::max<size_t>(1, h->seg_chunk.size());
        // Layout: scalars | vertex energies | chunk sums | segment energies | [slots | list | low-rank store].
        size_t off = 256;
        const size_t o_vE = off;
        off += round256(V * sizeof(double));
        const size_t o_chunk = off;
        off += round256(chunks * sizeof(double));
        const size_t o_seg = off;
        off += round256(segs * sizeof(double));
        size_t o_slot = 0, o_list = 0, o_H = 0;
        int32_t capacity = 0;
        if (hessian_kind == 1) {
            o_slot = off;
            off += round256(T * sizeof(int32_t));
            o_list = off;
            off += round256(T * sizeof(int32_t));
            o_H = off;
            const size_t per = lpg::kProjPerTet * sizeof(double);
            const size_t room = off < kNoAllocScratchBytes ? (kNoAllocScratchBytes - off) / per : 0;
            capacity = static_cast<int32_t>(std::min(T, room));
            if (capacity > 0)
                off += capacity * per;
            else
                off += per; // reported below as too large
        }
        if (off > kNoAllocScratchBytes)
            fail(lpg::kInvalid, "lpg_tet_energy_eval_device_noalloc: this mesh needs " + std::to_string(off) +
                                    " bytes of scratch; the allocation-free entry point has " +
                                    std::to_string(kNoAllocScratchBytes) + " per device");
        char* base = scratch_on(h->device);
        lpg::EvalScratch ws;
        ws.scalars = base;
        ws.vertex_E = reinterpret_cast<double*>(base + o_vE);
        ws.chunk_sum = reinterpret_cast<double*>(base + o_chunk);
        ws.seg_energy = reinterpret_cast<double*>(base + o_seg);
        if (hessian_kind == 1) {
            ws.proj_slot = reinterpret_cast<int32_t*>(base + o_slot);
            ws.proj_list = reinterpret_cast<int32_t*>(base + o_list);
            ws.proj_H = reinterpret_cast<double*>(base + o_H);
            ws.proj_capacity = capacity;
        }
        const hipStream_t s = static_cast<hipStream_t>(stream);
        lpg::enqueue_eval(h, x_dev, theta_dev, want_gradient, hessian_kind, gradient_dev, hessian_values_dev, nullptr,
                          s, &ws);
        try {
            *value = lpg::finish(h, s, hessian_kind == 1, base);
        } catch (const lpg::ProjectionOverflow& o) {
            fail(lpg::kInvalid, "lpg_tet_energy_eval_device_noalloc: " + std::to_string(o.listed) +
                                    " tets need the eigen-analysis, the scratch holds " + std::to_string(capacity));
        }
    });
}

} // extern "C"
