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
constexpr u32 PAD_DST = 0xFFFFFFFFu;  // head of a padding line (contributes to a spare slot)

// A graph as it arrives, on the device: rows with >= 1 line in input order,
// line heads as original node ids.
struct RawGraph {
    u64 N = 0, E = 0, U = 0;
    u64* row_ptr = nullptr;     // U + 1
    u32* row_node = nullptr;    // U
    u32* dst = nullptr;         // E
    u64* wt = nullptr;          // E or null
    u32* init = nullptr;        // N or null
    // In-degrees (they only order the internal ids) from every line, or from
    // a sample of the lines of a large graph: the sample saves most of the
    // count's cost but orders heads of similar in-degree less well, which
    // costs the binned layout's gathers a little locality every round, so
    // only a graph refined once (the one-shot rbsc_refine) takes it.
    bool sample_indeg = false;
    // Lines copied from the host in chunks, so that the build's line kernels
    // run on each chunk while the next one crosses the link: chunk c holds
    // lines [chunk_end[c - 1], chunk_end[c]) and is on the device once
    // chunk_ev[c] (recorded on the copying stream) has completed; n_chunks 0
    // means every line is present.  dst_err (with chunks) receives the flag
    // of a line whose head is not below N, checked by the build before any
    // kernel indexes by head.
    int n_chunks = 0;
    u64 chunk_end[4] = {0, 0, 0, 0};
    hipEvent_t chunk_ev[4] = {nullptr, nullptr, nullptr, nullptr};
    u32* dst_err = nullptr;
};

// Binned signature layout (bins.hip): stored rows cut into bins of S rows
// whose signatures are summed in LDS; each bin lists its lines as slots, hot
// slots (head < H) first, sorted by head, then cold slots (H <= head < GL),
// sorted by head, whose colours a scatter pass writes every round.  Lines
// into heads >= GL (never gathered) are summed per row into dsum.
struct BinChunk {
    u64 lo, hi;    // slot range
    u32 bin;
    u32 flags;     // CHUNK_FIRST: adds the bin's dsum terms; CHUNK_ONLY: the
};                 // bin's only chunk (stores; otherwise adds atomically)
constexpr u32 CHUNK_FIRST = 1, CHUNK_ONLY = 2;
constexpr u64 BIN_GROUPS = 8;  // scatter groups, one per XCD
// Warm heads are cut into WARM_SEGS segments, WARM_PER_XCD per XCD.
constexpr int WARM_SEGS = 32;
constexpr int WARM_PER_XCD = WARM_SEGS / (int)BIN_GROUPS;
struct WarmMap {
    unsigned char seg[BIN_GROUPS][WARM_PER_XCD];
};
// A run of at most WARM_ITEM warm slots of one bin and segment: slots
// [lo, lo + n) receive their colours at buffer indexes [dst, dst + n).
struct WarmItem {
    u64 lo, dst;
    u32 n, pad;
};
constexpr u32 WARM_ITEM = 8192;
struct WarmFirst {
    u32 f[BIN_GROUPS + 1];  // items of XCD x: [f[x], f[x + 1])
};
struct rbsc_bins {
    u32 log_s = 0, H = 0, GL = 0;
    // Heads below D (dense) are gathered by the bin pass itself; heads in
    // [D, H) (warm) are gathered once per round by warm_gather_kernel into
    // the refinement's buffer, like cold slots receive their colours: the
    // warm heads are cut into WARM_SEGS segments of about equal lines, dealt
    // to the XCDs so that each XCD gathers from its own segments only and
    // their colours stay in its L2.  A bin's slots are its dense, warm and
    // cold slots, the first two in head order; bin_hot counts the dense ones,
    // and the buffer covers the warm and cold ones.  The warm slots of a bin
    // in one segment are one run, cut into work items of at most WARM_ITEM
    // slots, listed by XCD.
    u32 D = 0;
    u64 n_warm = 0;
    WarmItem* warm_items = nullptr;
    WarmFirst warm_first{};
    u64 bytes_warm = 0;
    bool remap = false;             // lines into repeated rows point at their stored rows
    // Hashed layout: refinements on it keep per-node hashed colours
    // randomize(c + r) (u64, written once per node by the rank pass), hot
    // slots gather those and cold slots receive them, so no line is hashed.
    // Otherwise (the default) raw u32 colours are gathered and every line
    // hashes its colour.
    bool hashed = false;
    u64 n_bins = 0, n_slots = 0, n_hot = 0, n_cold = 0, n_chunks = 0, n_zero = 0;
    u64* bin_slot = nullptr;        // n_bins + 1: first slot of each bin
    u64* bin_hot = nullptr;         // n_bins: dense slots of each bin
    // The colours of bin b's cold slots [c0, c1) (c0 = bin_slot + bin_hot)
    // change every round; they live in the refinement's own cold-colour
    // buffer (rbsc_ctx::cold, n_cold_pad + 512 entries of u32, or u64 on the
    // hashed layout) at cold_pos[b] + (s - c0 rounded down to 8), each bin's
    // region padded to whole aligned groups of 8.  The layout itself is only
    // read by refinements.
    u64* cold_pos = nullptr;        // n_bins
    u64 n_cold_pad = 0;
    // Slot arrays and scatter records carry 512 zeroed entries past the end.
    u32* slot_key = nullptr;        // n_slots: hot = head; cold = 0
    unsigned short* slot_row = nullptr;  // n_slots: row - first row of the bin
    u64* slot_wt = nullptr;         // n_slots or null
    u64* dsum = nullptr;            // R: weight of the row's lines into heads >= GL
    // Scatter tiles: tile t = 8c + g is chunk c (scatter_chunk lines in head
    // order) of the cold lines of the bins in group g.  With head_order
    // (staged scatter) csc_head lists a tile's lines in head order, csc_perm
    // gives each of them its rank in the tile's slot order and csc_slot lists
    // the targets in slot order; otherwise both arrays are in slot order.
    u64 scatter_chunk = 0;
    bool head_order = false;
    u64 n_tiles = 0;
    u32* csc_head = nullptr;        // n_cold
    u32* csc_slot = nullptr;        // n_cold: cold-buffer index receiving a line's colour
    unsigned short* csc_perm = nullptr;  // n_cold with head_order
    u64* tile_line = nullptr;       // n_tiles + 1: first record of each tile
    BinChunk* chunks = nullptr;     // n_chunks
    u32* zero_bins = nullptr;       // n_zero: bins with several chunks
    u64 bytes = 0;
};
void bins_release(rbsc_bins* b, hipStream_t s);
// Builds g->bins (hot heads below H, of them dense below D, 2^log_s rows per
// bin, chunks of at most chunk_slots slots, scatter tiles of scatter_chunk
// lines).
int dgraph_bin(rbsc_ctx* ctx, rbsc_dgraph* g, u32 H, u32 D, u32 log_s, u64 chunk_slots,
               u64 scatter_chunk, bool head_order, bool hashed);
// One round on the binned layout: cold colours into their slots, then the
// signatures of the stored rows into sig.  col: per-node colours by internal
// id, u32 raw colours on the raw layout, u64 hashed colours of the round
// (randomize(c + salt)) on the hashed layout.
// cold_buf: the refinement's cold-colour buffer, bins_cold_bytes(b) bytes.
// Device address of [p, p + bytes) when it lies inside one buffer from
// rbsc_host_alloc (coherent page-locked memory kernels may write), else null.
void* host_device_ptr(const void* p, u64 bytes);
u64 bins_cold_bytes(const rbsc_bins* b);
void bins_scatter(rbsc_ctx* ctx, const rbsc_dgraph* g, const void* col, void* cold_buf);
void bins_sig(rbsc_ctx* ctx, const rbsc_dgraph* g, const void* col, const void* cold_buf,
              u64 salt, u64 dang_h, u64* sig);

// A graph resident in HBM in the round's layout (graph.hip).  Refinement
// only reads it (all per-refinement state lives in the context), so several
// contexts may refine one resident graph at the same time.  Internal ids
// put nodes with lines first (in-degree descending) and nodes without lines
// last; row j is node j; lines are grouped by row.  Nodes without lines keep
// signature 0 (rbsc.cu:110-111).
struct rbsc_dgraph {
    u32 N = 0;
    u64 E = 0;                  // stored lines (the rows' lines below)
    u64 E_in = 0;               // input lines (E plus the lines of duplicate rows)
    u32 U = 0;                  // nodes with >= 1 line
    u32 R = 0;                  // stored rows: nodes [0, R); [R, U) repeat a stored row
    u32* dup_rep = nullptr;     // U - R: stored row whose signature node R + i copies
    u64* row_ptr = nullptr;     // R + 1
    u32* pi = nullptr;          // N: original id -> internal id
    u32* dst = nullptr;         // E, internal ids (null when streamed)
    u64* wt = nullptr;          // E or null (all 1)
    u32* init = nullptr;        // N (internal order) or null
    // Without initial colours: per stored row, the weight of its lines into
    // nodes without lines (internal ids >= U) and, with weights, the weight of
    // all its lines (without weights: its line count); the first round's
    // signatures follow from them (first_round_sig).
    u64* dang_w = nullptr;      // R or null
    u64* row_w = nullptr;       // R or null
    u32* end_bits = nullptr;    // bit p: line p is the last line of its row
    u64 end_words = 0;
    u64 n_tiles = 0;            // wave tiles of SIG_TW lines
    u32* tile_row = nullptr;    // n_tiles + 1: row holding each tile's first line
    u32* carry_row = nullptr;   // n_tiles: row continuing past the tile, or NO_ROW
    u64 bytes = 0;
    // Streaming: lines kept in pinned host memory (dst == nullptr on device).
    u32* dst_host = nullptr;    // E, pinned
    u64* wt_host = nullptr;     // E, pinned, or null
    // Destination-partitioned layout (rbsc_graph_relayout, n_parts > 1): lines
    // regrouped by (dst mod n_parts), each partition padded to whole tiles;
    // "segments" are the (partition, row) runs, seg_row[k] the row of segment k
    // (padding segments point at the spare signature slot N).  Tiles, the
    // row-end bitmap and carries then count segments instead of rows.
    u32 n_parts = 1;
    u64 E_pad = 0;              // lines including padding
    u32* seg_row = nullptr;     // segments
    u64 n_segs = 0;
    std::vector<u64> part_tile;  // n_parts + 1: first tile of each partition
    u64* d_part_tile = nullptr;  // the same on the device
    rbsc_bins* bins = nullptr;   // binned signature layout (rbsc_graph_bin) or null
};

// Per-context device workspace for one refinement of up to cap_n nodes.
struct Workspace {
    u64 cap_n = 0;
    u64* keys[2] = {nullptr, nullptr};
    u64* P[2] = {nullptr, nullptr};   // (colour << 32) | node, sorted by colour
    u64* sig = nullptr;
    u64* hc = nullptr;                 // colours of the round per node: raw u32, or
                                       // randomize(colour + salt) (u64)
    u64* hc2 = nullptr;                // the next round's, written by the rank pass
    u64* rank_status = nullptr;
    void* cold = nullptr;              // cold-colour buffer of a binned graph's rounds
    u64 cold_bytes = 0;
    u64 rank_tiles = 0;
    u64* carry = nullptr;              // per tile: partial signature of its carry row
    u64 carry_cap = 0;
    u64* tie_runs = nullptr;           // (start, end) of small runs to fix after a 32-bit sort
    u64 tie_cap = 0;
    u64* tie_runs_m = nullptr;         // mid runs
    u64* tie_runs_l = nullptr;         // long runs
    u64* tie_pre = nullptr;            // long runs: lengths -> offsets, total
    u64* tie_tmp = nullptr;            // scan scratch
    u32* tie_mark = nullptr;           // cap_n: epoch stamps of listed run starts
    u32 tie_epoch = 0;
    u64* tie_buf = nullptr;            // 8 x tie_buf_cap: compact run members
    u64 tie_buf_cap = 0;
    int pc = 0;                        // P[pc] holds the last refinement's order
    u64 n_rank = 0;                    // nodes ranked by the last refinement
    bool follow = false;               // its repeated rows took their stored rows' colours
    bool node_order = false;           // its final colours are in hc (raw, by node), not P
};

// Inputs of one refinement beyond the graph itself.
struct RefineParams {
    const u64* d_sig_fill = nullptr;  // per-node signature of nodes without lines, by
                                      // original id (else 0); needs initial colours
    u32 k0 = 1;                       // K before round 0 (rbsc.cu:339)
    u32 max_rounds = 0;               // 0 = to the fixpoint
    u64 slice_lines = 0;              // streaming slice size; 0 = lines resident
    u32 flags = 0;                    // RBSC_FLAG_*
    bool exact_colours = false;       // the caller reads the last round's raw colours
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
    bool dedup = true;         // store identical rows once (graph build)
    SortWorkspace sort;
    Workspace ws;
    u32* d_err = nullptr;
    u32* d_scalar = nullptr;   // [0] = K of the last rank pass, [1] = tie fix-up overflow,
                               // [3] = rank input out of colour order, [8..10] = small, mid
                               // and long tie runs listed
    u64 full_sort_redos = 0;   // rounds whose node side was redone (run list overflow, or
                               // equal signatures out of colour order)
    u64 long_run_rounds = 0;   // rounds with runs too long for the in-kernel tie sorts
    u64 confirmed_rounds = 0;  // last rounds confirmed from the previous round's blocks
    u32* h_scalar = nullptr;   // pinned mirror
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // streaming double buffer
    u32* sbuf[2] = {nullptr, nullptr};
    u64* swbuf[2] = {nullptr, nullptr};
    u64 sbuf_lines = 0;
    hipEvent_t ev_copy[2] = {nullptr, nullptr}, ev_done[2] = {nullptr, nullptr};
    hipEvent_t ev_chunk[4] = {nullptr, nullptr, nullptr, nullptr};  // chunked line uploads
    // fork and join of launches split over stream and copy_stream (batch tiers)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // per-phase profiling (rbsc_ctx_profile)
    bool prof = false;
    static constexpr int NPHASE = 7;
    std::vector<hipEvent_t> prof_ev;   // pairs (start, stop) recorded this refine
    std::vector<int> prof_phase;       // phase of each pair
    double prof_ms[NPHASE] = {0, 0, 0, 0, 0, 0, 0};
    u64 prof_n[NPHASE] = {0, 0, 0, 0, 0, 0, 0};
    size_t prof_used = 0;
};

// Phase markers; no-ops unless profiling is on.
void prof_begin(rbsc_ctx* ctx, int phase);
void prof_end(rbsc_ctx* ctx);
int prof_collect(rbsc_ctx* ctx);

int ws_reserve(rbsc_ctx* ctx, u64 n);
// Frees the refinement workspace (it is allocated again on demand).
void ws_release(rbsc_ctx* ctx);
void raw_release(RawGraph* r, hipStream_t s);
// Builds g from raw (raw arrays are left for the caller to release).
int dgraph_build(rbsc_ctx* ctx, RawGraph* raw, rbsc_dgraph* g);
// Regroups lines by head: n_parts interleaved partitions (head mod n_parts),
// or with hot > 0 two partitions, heads below `hot` (the most-gathered
// nodes) first.
int dgraph_relayout(rbsc_ctx* ctx, rbsc_dgraph* g, u32 n_parts, u32 hot);
int refine_core(rbsc_ctx* ctx, rbsc_dgraph* g, const RefineParams& p, RefineOut* out);
// Raw colours (dense ranks of the last round) of the last refine_core on g,
// by original node id, to host.
int download_colors(rbsc_ctx* ctx, rbsc_dgraph* g, u32* host);
// Canonical first-occurrence labels and/or raw colours by original id into
// host memory, computed on the device after a refinement (reuses workspace
// arrays: call it after anything else that reads the refinement's state).
int device_partition(rbsc_ctx* ctx, rbsc_dgraph* g, u32* partition_out, u32* colors_out);
// Block sizes of g's final colouring (K colours) in descending order, then
// zeros: N entries into host memory sizes_out.  Uses workspace arrays that
// hold no colours, so device_partition may follow it (not precede it).
int device_block_sizes(rbsc_ctx* ctx, rbsc_dgraph* g, u32 K, u32* sizes_out);
// The same colours by internal id into a device buffer of N words.
void device_colors(rbsc_ctx* ctx, rbsc_dgraph* g, u32* d_int);
// Exact stability check (verify.hip) of internal colours d_col (< K);
// *bad_out = number of violations, 0 iff stable.
int verify_partition(rbsc_ctx* ctx, rbsc_dgraph* g, const u32* d_col, u32 K, u64* bad_out);
int graph_upload_host(rbsc_ctx* ctx, const rbsc_graph* gr, bool lines_on_host, rbsc_dgraph** out,
                      bool sample_indeg = false);
int quotient_refine(rbsc_ctx* ctx, const rbsc_graph* gr, const rbsc_opts* opts, rbsc_result* res,
                    uint32_t* partition_out);
int check_kernel_errors(rbsc_ctx* ctx);
void dgraph_release(rbsc_dgraph* g, hipStream_t s);
// Canonical first-occurrence labels (rbsc.cu:431-445) of a colour vector.
u32 canonicalize_host(u32 N, const u32* colors, u32* out);
