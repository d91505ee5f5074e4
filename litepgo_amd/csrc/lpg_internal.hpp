// Internal declarations shared by the host setup, the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace lpg {

// Error classes of the reference (std::invalid_argument, std::domain_error, std::logic_error).
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
enum Code { kOk = 0, kInvalid = 1, kDomain = 2, kLogic = 3, kBusy = 4, kDevice = 5 };

[[noreturn]] inline void fail(int code, const std::string& msg) {
    throw Error(code, msg);
}

// Device status bits, OR-ed by kernels and mapped to a Code by the host.
enum StatusBits : unsigned {
    kStatusNonFiniteX = 1u,     // x has a non-finite entry                      -> invalid_argument
    kStatusBadTheta = 2u,       // lambda < 0, mu <= 0 or non-finite             -> invalid_argument
    kStatusDetF = 4u,           // det F <= 0 or non-finite kinematics           -> domain_error
    kStatusNonFiniteOut = 8u,   // non-finite element or assembled outputs       -> domain_error
    kStatusProjection = 16u,    // projection failed to converge / non-finite    -> domain_error
    kStatusNonFiniteStep = 32u, // non-finite deformation-gradient increment     -> domain_error
    kStatusNonFiniteFp = 64u,   // a per-call plastic part with a non-finite entry -> invalid_argument
    kStatusBadDetFp = 128u,     // a per-call plastic part with det <= 0 or non-finite -> domain_error
    kStatusProjOverflow = 256u, // more listed tets than the projection store holds -> grow it and evaluate again
};

// Device layout (no per-tet geometry is stored; Dm^-1 and w are recomputed from the rest positions):
//   Xr[V]    : rest positions as double4 (x, y, z, 0), 32 bytes per vertex
//   conn[T]  : the 4 vertex ids of a tet as int4, 16 bytes per tet
//   theta[T] : (lambda, mu) as double2, 16 bytes per tet

// Incidence record of (vertex v, tet e), for v at slot a of e.  The kernels process the tet in the vertex order
// (a, a^1, a^2, a^3) -- an even permutation of (0, 1, 2, 3), so det Dm and w keep their sign -- which puts v at
// local slot 0.  The record holds the tet id, a, and for j = 1..3 the position of vertex e[a^j] in v's sorted
// neighbour list (its column block in v's rows).  Narrow form: 3 x 8-bit positions and a in bits 24..25.
struct IncNarrow {
    int32_t tet;
    uint32_t packed; // cp1 | cp2 << 8 | cp3 << 16 | a << 24
};
struct IncWide {
    int32_t tet;
    uint32_t cp12; // cp1 | cp2 << 16
    uint32_t cp3a; // cp3 | a << 16
    uint32_t pad;
};

// Host-side result of construction (everything the kernels read, plus the exported pattern).
struct MeshSetup {
    int32_t num_vertices = 0;
    int32_t num_tets = 0;
    int32_t num_dofs = 0;
    int32_t nnz = 0;
    int32_t max_degree = 0;          // max |neighbours incl. self| over vertices
    bool wide = false;               // max_degree > 255
    bool theta_valid = true;
    std::vector<int32_t> row_ptr;    // num_dofs + 1
    std::vector<int32_t> col_idx;    // nnz
    std::vector<int32_t> nbr_ptr;    // num_vertices + 1
    std::vector<int32_t> nbr;        // sorted vertex neighbours incl. self
    std::vector<int32_t> inc_ptr;    // num_vertices + 1
    std::vector<int32_t> inc_tet;    // 4 T, element order per vertex
    std::vector<uint16_t> inc_colpos; // 4 T x 4, position of e[b] in nbr(v) for slot b (original order)
    std::vector<int32_t> self_pos;   // num_vertices, position of v in nbr(v)
    std::vector<int32_t> conn;       // T x 4 vertex ids
    std::vector<double> Xr;          // V x 4 rest positions (x, y, z, 0)
    std::vector<double> theta;       // T x 2 (lambda, mu) snapshot
    // Plastic parts, empty when every Fp is the identity: per tet row-major Fp^-1 and det Fp (10 doubles).  An Fp
    // that MultiplicativeKinematics rejects (multiplicative_kinematics.cpp:9-21) is recorded here and raised by
    // every evaluation, as the reference raises it per evaluated element.
    std::vector<double> fp;
    int fp_error = 0;
    std::string fp_error_msg;
};

// Builds MeshSetup; throws Error(kInvalid, ...) with the reference's messages on invalid input.
void build_mesh_setup(MeshSetup& s, int32_t nv, int32_t nt, const int32_t* tets, const double* rest,
                      const double* theta, const double* Fp, int nthreads);

// Reference-convention slot map from the setup (offsets[144 e + i + 12 j]).
void build_slot_map(const MeshSetup& s, int32_t* offsets);

// Vertex tiles: contiguous vertex ranges handled by one workgroup.  tile_vbeg has tiles + 1 entries.
struct TileConfig {
    int vertices_per_tile;    // upper bound on vertices in a tile
    int lds_budget_doubles;   // upper bound on the H segment (9 |nbr(v)| doubles per vertex) staged in LDS
};
// Largest H segment over the tiles that stage it in LDS.
struct TileCaps {
    int32_t segment = 0;
};
// Tiles never cross a vertex listed in `breaks` (mesh boundaries of a batch); null means no breaks.  A tile of one
// vertex whose segment exceeds the budget is listed in `oversized` and accumulates in HBM instead.
void build_tiles(const MeshSetup& s, const TileConfig& cfg, std::vector<int32_t>& tile_vbeg, TileCaps& caps,
                 std::vector<int32_t>* oversized, const std::vector<int32_t>* breaks = nullptr);

// ------------------------------------------------------------------------------------------------------------
// Kernel parameters and launchers (lpg_kernels.hip).
// ------------------------------------------------------------------------------------------------------------
struct DeviceMesh {
    const double* Xr = nullptr;     // V x 4
    const int32_t* conn = nullptr;  // T x 4
    const double* theta = nullptr;  // T x 2 snapshot
    const int32_t* inc_ptr = nullptr;
    const void* inc = nullptr; // IncNarrow* or IncWide*
    const int32_t* nbr_ptr = nullptr;
    const int32_t* nbr = nullptr;
    const int32_t* self_pos = nullptr;
    const int32_t* row_ptr = nullptr;
    const double* fp = nullptr; // T x 10 plastic parts (row-major Fp^-1, det Fp), null when every Fp = I
    int32_t num_vertices = 0;
    int32_t num_tets = 0;
    bool wide = false;
};

struct EvalArgs {
    const double* x = nullptr;
    const double* theta = nullptr; // optional override, T x 2
    double* g = nullptr;
    double* H = nullptr;
    int want_g = 0;
    int kind = -1;
    // scratch
    double* vertex_E = nullptr;  // one per vertex
    double* energy = nullptr;    // 1 double (device)
    unsigned* status = nullptr;  // 1 word (device)
    unsigned* counter = nullptr; // indefinite count (device)
    int32_t* proj_slot = nullptr; // T
    int32_t* proj_list = nullptr; // T
    double* proj_H = nullptr;    // proj_capacity x kProjPerTet (list-indexed, first *counter used)
    int32_t proj_capacity = 0;   // list entries proj_H holds; a longer list sets kStatusProjOverflow
};

// Projection of a listed (not certified positive definite) tet as a low-rank term: proj(H_e) = (replace ? 0 : H_e)
// + sum_{v < count} u_v u_v^T with at most kProjVectors 12-vectors u_v (element DOF order 3a + i).  Stored per list
// entry as [count, 3 unused, u_0, u_1, u_2, u_3] (32-byte aligned); u_0 is written (zero) even when count is 0.
// The tet's proj_slot holds its list index s (H_e plus the term) or -2 - s (the term replaces H_e); -1: not listed.
constexpr int kProjVectors = 4;
constexpr int kProjPerTet = 4 + 12 * kProjVectors;

struct LaunchConfig {
    int lanes_per_vertex = 4; // lanes sharing one vertex's incidences
    int waves = 1;            // waves per workgroup
};
// Launch shapes with compiled assembly kernels: (lanes per vertex, waves).
bool launch_shape_supported(int lanes_per_vertex, int waves, bool wide);

void launch_clear(unsigned* status, unsigned* counter, hipStream_t stream);
// Processes tiles tile_ids[0..num_tiles) (or 0..num_tiles-1 when tile_ids is null).  global_acc accumulates in
// the H output itself instead of LDS (tiles whose segment exceeds the LDS budget); otherwise the H segment is
// staged in LDS, sized by caps.  persistent_grid > 0 launches that many workgroups, each walking the tile list with
// a grid stride.
void launch_assemble(const DeviceMesh& m, const EvalArgs& a, const LaunchConfig& lc, const int32_t* tile_vbeg,
                     const int32_t* tile_ids, int num_tiles, const TileCaps& caps, bool global_acc,
                     int persistent_grid, hipStream_t stream);
// Dynamic LDS bytes of one assembly workgroup for a launch shape and staged sizes.
size_t assemble_lds_bytes(const LaunchConfig& lc, const TileCaps& caps, bool global_acc, int kind);
// Largest dynamic LDS allocation an assembly workgroup of this shape may request.
size_t assemble_lds_limit(const LaunchConfig& lc);
// Energy per mesh from the per-vertex energies: chunk c sums vertex_E[chunk_begin[c] .. chunk_begin[c+1]),
// segment m sums chunk_sum[seg_chunk[m] .. seg_chunk[m+1]) into out[m].
// readback (one segment only; device-accessible pinned memory, or null): receives the energy, the status word and
// `counter` (16 bytes) from the last kernel of the evaluation, in place of a separate device-to-host copy.
void launch_energy_reduce(const double* vertex_E, const int32_t* chunk_begin, int num_chunks, double* chunk_sum,
                          const int32_t* seg_chunk, int num_segments, double* out, unsigned* status,
                          hipStream_t stream, const unsigned* counter = nullptr, void* readback = nullptr);
// Chunks of at most 1024 vertices, restarting at every segment start: chunk_begin (chunks + 1 entries) and
// seg_chunk (segments + 1 entries) for segment starts seg_vbeg (segments + 1 entries, the last = num_vertices).
void build_energy_chunks(const std::vector<int32_t>& seg_vbeg, std::vector<int32_t>& chunk_begin,
                         std::vector<int32_t>& seg_chunk);
void launch_project_prepass(const DeviceMesh& m, const EvalArgs& a, hipStream_t stream);
void launch_project_jacobi(const DeviceMesh& m, const EvalArgs& a, hipStream_t stream);
void launch_dE_dtheta(const DeviceMesh& m, const double* x, const double* theta, double* out, unsigned* status,
                      hipStream_t stream);
void launch_force_vjp_theta(const DeviceMesh& m, const double* x, const double* theta, const double* w,
                            double* out, unsigned* status, hipStream_t stream);
// Force response to the plastic part at the per-call Fp (9 per tet, column-major; null: the handle's snapshot):
// w given: out = (df/dFp)^T w (9 per tet); else out = f (dfp null) or (df/dFp) dfp (num_dofs), through `scratch`
// (3 doubles per incidence), written through `tet_inc` (the incidence index of each (tet, slot), 4 per tet).
void launch_force_fp(const DeviceMesh& m, const double* x, const double* theta, const double* fp, const double* w,
                     const double* dfp, const int32_t* tet_inc, double* scratch, double* out, unsigned* status,
                     hipStream_t stream);
void launch_force_jvp_theta(const DeviceMesh& m, const double* x, const double* theta, const double* dtheta,
                            double* out, unsigned* status, hipStream_t stream);
void launch_hvp(const DeviceMesh& m, const double* x, const double* theta, const double* v, double* out,
                unsigned* status, hipStream_t stream);
// gx = s f + H w and gtheta = s dE/dtheta + (df/dtheta)^T w with s = *gE (0 if null), w optional; either
// output may be null.
void launch_vjp(const DeviceMesh& m, const double* x, const double* theta, const double* gE, const double* w,
                double* gx, double* gtheta, unsigned* status, hipStream_t stream);
// out[0] = sum of seg[0..n) in a fixed order (one workgroup).
void launch_energy_total(const double* seg, int n, double* out, hipStream_t stream);
// Energy, status word and list counter to host-visible pinned memory (status word last), then the status word and
// the counter zeroed: an evaluation's last kernel when its energy reduction cannot carry the read-back itself.
void launch_readback(const double* energy, unsigned* status, const unsigned* counter, void* readback,
                     hipStream_t stream);
// Per-segment feasible steps (alphas: num_segments doubles on the device, min over the segment's tets, 1 when
// none limits); also clears the status word.
void launch_feasible_step(const DeviceMesh& m, const double* x, const double* dx, const int32_t* seg_tbeg,
                          int num_segments, double* alphas, unsigned* status, hipStream_t stream);
// theta_e = (lambda, mu)(E_e, nu_e) and its reverse product.
void launch_lame(const double* E, const double* nu, int64_t n, double* theta, hipStream_t stream);
void launch_lame_vjp(const double* E, const double* nu, const double* gtheta, int64_t n, double* gE, double* gnu,
                     hipStream_t stream);

} // namespace lpg
