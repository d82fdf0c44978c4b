// internal.h -- context and resident-graph layout shared by the library's
// translation units.
#pragma once
#include "common.h"
#include "sort.h"
#include "../../include/rbsc.h"
#include <vector>

// Edges per lane and per wave tile of the signature kernel.
constexpr int SIG_V = 8;
constexpr u64 SIG_TW = 64 * SIG_V;
constexpr u32 NO_ROW = 0xFFFFFFFFu;

// A graph resident in HBM, in compressed CSR form: only rows with at least
// one line are stored (row k is node row_node[k], its lines are
// dst[row_ptr[k] .. row_ptr[k+1])), in the order the lines were given.
// Nodes without lines keep signature 0 (rbsc.cu:110-111).
struct rbsc_dgraph {
    u32 N = 0;
    u64 E = 0;
    u32 U = 0;                  // rows with >= 1 line
    u64* row_ptr = nullptr;     // U + 1
    u32* row_node = nullptr;    // U
    u32* dst = nullptr;         // E
    u64* wt = nullptr;          // E or null (all 1)
    u32* init = nullptr;        // N or null
    u64 n_tiles = 0;            // wave tiles of SIG_TW lines
    u32* tile_row = nullptr;    // n_tiles + 1: row holding each tile's first line
    u32* carry_row = nullptr;   // n_tiles: row continuing past the tile, or NO_ROW
    u64* carry_val = nullptr;   // n_tiles: that row's partial signature
    u64 bytes = 0;
    bool owns = true;
    // Streaming: lines kept in pinned host memory (dst == nullptr on device).
    u32* dst_host = nullptr;    // E, pinned
    u64* wt_host = nullptr;     // E, pinned, or null
};

// Per-context device workspace for one refinement of up to cap_n nodes.
struct Workspace {
    u64 cap_n = 0;
    u64* keys[2] = {nullptr, nullptr};
    u64* P[2] = {nullptr, nullptr};   // (colour << 32) | node, sorted by colour
    u64* sig = nullptr;
    u64* hc = nullptr;                 // randomize(colour + salt) per node
    u64* rank_status = nullptr;
    u64 rank_tiles = 0;
    int pc = 0;                        // P[pc] holds the last refinement's order
};

// Inputs of one refinement beyond the graph itself.
struct RefineParams {
    const u32* d_init = nullptr;      // initial colours on the device (else graph's / default)
    const u64* d_sig_fill = nullptr;  // per-node signature of nodes without lines (else 0)
    u32 k0 = 1;                       // K before round 0 (rbsc.cu:339)
    u32 max_rounds = 0;               // 0 = to the fixpoint
    u64 slice_lines = 0;              // streaming slice size; 0 = lines resident
};
struct RefineOut {
    u32 K = 0;
    u32 rounds = 0;
    float device_ms = 0;
    double pcie_bytes = 0;            // bytes streamed host -> device
};

struct rbsc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;
    u64 budget = 0;
    SortWorkspace sort;
    Workspace ws;
    u32* d_err = nullptr;
    u32* d_scalar = nullptr;   // [0] = K of the last rank pass
    u32* h_scalar = nullptr;   // pinned mirror
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // streaming double buffer
    u32* sbuf[2] = {nullptr, nullptr};
    u64* swbuf[2] = {nullptr, nullptr};
    u64 sbuf_lines = 0;
    hipEvent_t ev_copy[2] = {nullptr, nullptr}, ev_done[2] = {nullptr, nullptr};
    // per-phase profiling (rbsc_ctx_profile)
    bool prof = false;
    static constexpr int NPHASE = 5;
    std::vector<hipEvent_t> prof_ev;   // pairs (start, stop) recorded this refine
    std::vector<int> prof_phase;       // phase of each pair
    double prof_ms[NPHASE] = {0, 0, 0, 0, 0};
    u64 prof_n[NPHASE] = {0, 0, 0, 0, 0};
    size_t prof_used = 0;
};

// Phase markers; no-ops unless profiling is on.
void prof_begin(rbsc_ctx* ctx, int phase);
void prof_end(rbsc_ctx* ctx);
int prof_collect(rbsc_ctx* ctx);

int ws_reserve(rbsc_ctx* ctx, u64 n);
int refine_core(rbsc_ctx* ctx, rbsc_dgraph* g, const RefineParams& p, RefineOut* out);
// Raw colours (dense ranks of the last round) of the last refine_core, to host.
int download_colors(rbsc_ctx* ctx, u64 N, u32* host);
int graph_upload_host(rbsc_ctx* ctx, const rbsc_graph* gr, bool lines_on_host, rbsc_dgraph** out);
int quotient_refine(rbsc_ctx* ctx, const rbsc_graph* gr, const rbsc_opts* opts, rbsc_result* res,
                    uint32_t* partition_out);
int check_kernel_errors(rbsc_ctx* ctx);
int dgraph_build_tiles(rbsc_ctx* ctx, rbsc_dgraph* g);
int dgraph_alloc(rbsc_dgraph* g, u32 N, u64 E, u32 U, bool weighted, bool has_init,
                 bool lines_on_device = true);
void dgraph_release(rbsc_dgraph* g);
// Canonical first-occurrence labels (rbsc.cu:431-445) of a colour vector.
u32 canonicalize_host(u32 N, const u32* colors, u32* out);
