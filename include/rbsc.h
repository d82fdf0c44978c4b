/*
 * rbsc.h -- C ABI of the MI355X (gfx950) RBSC refinement library.
 *
 * The reference tool is a single CUDA program whose hot path is the round
 * loop of src/rbsc/rbsc.cu:354-425, driven by main() (rbsc.cu:182-497), fed
 * by the stdin parser read_graph (rbsc.cu:168-180) and reporting three lines
 * (rbsc.cu:492-494).  Its internal seam is a set of CUB device calls
 * (rbsc.cu:258-261, 366-377, 386, 390, 394, 409).  This header is the seam
 * that replaces all of it: plain pointers and sizes, int status codes, no
 * exceptions and no C++ or torch types.
 *
 * Status codes: 0 on success, negative on failure; rbsc_last_error() returns
 * a thread-local message for the last failure of the calling thread.
 *
 * Ownership: the caller owns every host buffer passed in or out.  The library
 * owns device memory behind opaque handles.  One context per host thread (or
 * per stream); calls on distinct contexts are thread-safe.
 */
#ifndef RBSC_H
#define RBSC_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBSC_OK 0
#define RBSC_ERR_INVALID -1   /* bad argument */
#define RBSC_ERR_NOMEM -2     /* host or device allocation failed */
#define RBSC_ERR_HIP -3       /* HIP runtime error */
#define RBSC_ERR_BUDGET -4    /* request exceeds the HBM budget */
#define RBSC_ERR_INPUT -5     /* malformed graph: id >= N, or ungrouped lines */
#define RBSC_ERR_KERNEL -6    /* a kernel's bounded wait expired (logic error) */
#define RBSC_ERR_NODEVICE -7  /* no HIP device */

/* Edge-batch modes (opts.mode), used when max_batch_edges < n_edges. */
#define RBSC_MODE_STREAM 0    /* exact full-graph rounds, edge slices streamed
                                 from pinned host memory (done = 1) */
#define RBSC_MODE_QUOTIENT 1  /* reference multi-batch quotient passes,
                                 rbsc.cu:274-487 (done may be 0) */
#define RBSC_MODE_QUOTIENT_AGGRESSIVE 2 /* rbsc_aggressive.cu:487 stop rule */

/* A graph in the reference's line form (rbsc.cu:168-180): line e reads
 * "src[e] weight[e] dst[e]" and adds randomize(colour(dst)) * weight to the
 * signature of src (rbsc.cu:80-87, 366-377).  Lines must be grouped by src
 * (equal src contiguous), as the reference's ReduceByKey requires.
 * Either src (COO) or row_offsets (CSR over node ids, n_nodes + 1 entries)
 * is given.  weight == NULL means every weight is 1.  init_color == NULL
 * means the reference's initial colouring (rbsc.cu:51-63: N+1 for nodes with
 * a line, 0 for the others); otherwise init_color[u] is u's initial colour
 * (an extension: the reference has no initial-partition input). */
typedef struct rbsc_graph {
    uint32_t n_nodes;
    uint64_t n_edges;
    const uint64_t* row_offsets; /* CSR, nullable */
    const uint32_t* src;         /* COO, nullable when row_offsets given */
    const uint32_t* dst;
    const uint64_t* weight;      /* nullable */
    const uint32_t* init_color;  /* nullable */
} rbsc_graph;

/* Options.  max_batch_edges mirrors argv[1] of the reference (rbsc.cu:204):
 * 0 or >= n_edges means one batch. */
typedef struct rbsc_opts {
    uint64_t max_batch_edges;
    uint32_t mode;        /* RBSC_MODE_* */
    uint32_t max_rounds;  /* 0 = until the fixpoint; else stop after this many */
    uint32_t verify;      /* 1 = exact stability check of the final partition */
    uint32_t flags;       /* RBSC_FLAG_* implementation choices (results unchanged) */
} rbsc_opts;

/* opts.flags: gather per-node hashed colours randomize(colour + r) (8 bytes)
 * instead of the default raw 32-bit colours hashed on every line
 * (rbsc.cu:85): fewer multiplies, twice the gathered bytes. */
#define RBSC_FLAG_HASHED64 1u
/* opts.flags: on a graph relaid out into 8 destination partitions, run all
 * partitions in one launch with workgroup b serving partition b mod 8, so
 * each of the 8 XCDs' L2 caches its own eighth of the colours. */
#define RBSC_FLAG_XCD 2u
/* opts.flags: the default node-side sort orders the top 32 signature bits
 * (4 passes instead of 8), then re-sorts only the runs whose members collide
 * in the top half but differ in the low half (gathered into a compact array);
 * RBSC_FLAG_SPLIT names that default.  RBSC_FLAG_FULLSORT sorts all 64 bits
 * instead (results identical). */
#define RBSC_FLAG_SPLIT 4u
#define RBSC_FLAG_FULLSORT 32u
/* Verification hooks for the split sort (results unchanged, slower): split
 * at 16 bits instead, so colliding runs are common; cap the run list at 2
 * entries, so the full-sort fallback runs. */
#define RBSC_FLAG_SPLIT16 8u
#define RBSC_FLAG_TIESMALL 16u
/* opts.flags: on a binned graph (rbsc_graph_bin) compute signatures with the
 * plain line kernel instead (results identical). */
#define RBSC_FLAG_NOBINS 64u
/* opts.flags: without initial colours the node side sorts the nodes by
 * signature straight from node-id order (no gather of signatures into
 * colour order; equal signatures have equal colours from a uniform start,
 * and a round with a pair out of colour order is redone exactly).  This flag
 * keeps positions in colour order instead, as with initial colours (results
 * identical). */
#define RBSC_FLAG_COLOUR_ORDER 128u
/* Node order only: the sort moves 8 bytes per node (the signature's top 32
 * bits with the node below them) instead of 16 (signature, colour and node),
 * and the tie fix-up and rank read the signature's low bits and the colours
 * by node where neighbouring top bits are equal.  Used by default in rounds
 * whose previous block count is at least half the ranked nodes;
 * RBSC_FLAG_PACKED uses it in every such round (a test hook),
 * RBSC_FLAG_UNPACKED in none.  Results are identical. */
#define RBSC_FLAG_PACKED 512u
#define RBSC_FLAG_UNPACKED 1024u
/* A round whose previous block count is 1 (the first round from a uniform
 * start, node order) ranks the distinct signatures through a hash table
 * instead of sorting all nodes; RBSC_FLAG_SORTRANK sorts as in the other
 * rounds (a test hook).  Results are identical. */
#define RBSC_FLAG_SORTRANK 2048u
/* Node order: the last round (the one whose block count repeats the
 * previous one's) is confirmed without sorting or ranking -- the previous
 * round's rank pass lists the nodes of its blocks of two or more nodes, and
 * when none of those blocks has two signatures this round, no block split.
 * The final colours are then the previous round's, which name the same
 * blocks (the partition, block count and rounds are those of the ranked
 * round).  RBSC_FLAG_FULLROUNDS ranks the last round too (a test hook); so
 * do a round limit and a colours_out request. */
#define RBSC_FLAG_FULLROUNDS 4096u
/* Binned graph without initial colours: in the first round every gathered
 * head has colour N+1 and every other head colour 0 (rbsc.cu:51-63), so each
 * row's signature is randomize(N+1) * (weight of its gathered lines) +
 * randomize(0) * (weight of its other lines) in Z/2^64, computed per row
 * from the layout's per-row weights without reading the lines.
 * RBSC_FLAG_GATHER_FIRST gathers the first round's colours like every other
 * round's (a test hook).  Signatures and results are identical. */
#define RBSC_FLAG_GATHER_FIRST 8192u

/* Result.  nodes / done / time_s are the reference's three output lines
 * (rbsc.cu:492-494). */
typedef struct rbsc_result {
    uint64_t nodes;          /* number of blocks */
    uint32_t done;           /* 1 iff the coarsest reduction was reached */
    uint32_t rounds;         /* refinement rounds executed (all passes) */
    double time_s;           /* span after allocation to the final relabel */
    double device_ms;        /* device time of the rounds (hipEvents) */
    int32_t verified;        /* 1 stable (exact), 0 not checked, -1 unstable */
    uint32_t passes;         /* quotient passes (1 when not in quotient mode) */
    uint64_t edges_per_round;/* edge visits per round (lines) */
} rbsc_result;

typedef struct rbsc_ctx rbsc_ctx;
typedef struct rbsc_dgraph rbsc_dgraph;

/* Last error message of the calling thread ("" if none). */
const char* rbsc_last_error(void);

/* Library version string. */
const char* rbsc_version(void);

/* Page-locked, coherent host memory for callers' output buffers (labels,
 * colours, partitions): device-to-host copies into it run at the link's
 * rate, where copies into pageable memory are staged through the runtime's
 * own pinned buffers, and batched refinement writes labels into it from the
 * kernels directly (the transfer overlaps the refinement).  No counterpart
 * in the reference (its outputs are host vectors, rbsc.cu:427-445). */
int rbsc_host_alloc(uint64_t bytes, void** out);
int rbsc_host_free(void* p);

/* Context on HIP device `device`; hbm_budget = 0 means no budget.  Replaces
 * the per-GPU setup of rbsc.cu:234-270 (streams, scratch, CUB temp).  The
 * library's device memory comes from the device's default stream-ordered
 * memory pool; creating a context raises that pool's release threshold, so
 * freed memory stays cached for reuse, and destroying one hands the pool's
 * unused memory back to the device. */
int rbsc_ctx_create(int device, uint64_t hbm_budget, rbsc_ctx** out);
int rbsc_ctx_destroy(rbsc_ctx* ctx);

/* One-shot refinement of a host graph: upload, refine to the fixpoint,
 * canonical partition out.  Replaces main() of rbsc.cu:182-497 after parsing.
 * partition_out (nullable, n_nodes entries) receives the first-occurrence
 * canonical labels (rbsc.cu:431-445). */
int rbsc_refine(rbsc_ctx* ctx, const rbsc_graph* graph, const rbsc_opts* opts,
                rbsc_result* result, uint32_t* partition_out);

/* Device-resident graphs: upload once, refine many times (the bench and the
 * serving loop keep graphs resident).  Upload validates ids and grouping. */
int rbsc_graph_upload(rbsc_ctx* ctx, const rbsc_graph* graph, rbsc_dgraph** out);
/* The graph must be idle (no refinement of it in flight on any context).
 * Its device memory goes back to the device's memory pool, ordered on ctx's
 * stream (ctx may be null: then on the null stream), without waiting for
 * other streams' work. */
int rbsc_graph_free(rbsc_ctx* ctx, rbsc_dgraph* g);
/* Bytes of device memory held by a resident graph. */
uint64_t rbsc_graph_bytes(const rbsc_dgraph* g);
/* Refine a resident graph.  partition_out: host, nullable, canonical labels.
 * colors_out: host, nullable, the raw final colours (dense ranks of the last
 * round, the values rbsc.cu:429 copies back), for per-round parity checks. */
int rbsc_graph_refine(rbsc_ctx* ctx, rbsc_dgraph* g, const rbsc_opts* opts,
                      rbsc_result* result, uint32_t* partition_out,
                      uint32_t* colors_out);

/* Move a resident graph's line arrays (dst, weights) to pinned host memory;
 * later refinements stream them in slices of opts.max_batch_edges lines with
 * hipMemcpyAsync on a side stream, double-buffered against the signature
 * kernel, and compute the exact full-graph rounds (done = 1).  Replaces the
 * reference's per-batch host->device copies (rbsc.cu:341-344) for graphs
 * larger than HBM. */
int rbsc_graph_lines_to_host(rbsc_ctx* ctx, rbsc_dgraph* g);

/* Regroup a resident graph's lines by destination (head id mod n_parts,
 * n_parts a power of two <= 16; 1 keeps the plain layout).  Each round then
 * runs the signature kernel once per partition, so the hashed colours one
 * launch gathers are 1/n_parts of all of them and stay in L2.  Results are
 * unchanged.  No counterpart in the reference (its set_values gathers over
 * the whole colour array, rbsc.cu:80-87). */
int rbsc_graph_relayout(rbsc_ctx* ctx, rbsc_dgraph* g, uint32_t n_parts);
/* Two-pass line layout by head range: lines into the hot_heads most-gathered
 * nodes (internal ids are in-degree order) in one pass, whose colours stay
 * cache resident, and the rest in a second pass.  Results are unchanged.  No
 * counterpart in the reference. */
int rbsc_graph_split_hot(rbsc_ctx* ctx, rbsc_dgraph* g, uint32_t hot_heads);

/* Binned signature layout of a resident graph, an addition to its plain
 * layout (about 6 bytes per gathered line plus 8 per line into a cold head;
 * 8 more per line with weights).  Stored rows are cut into bins of
 * rows_per_bin (a power of two from 256 to 16384; 0 = 8192) whose
 * signatures are summed in LDS; lines into the hot_heads most-gathered nodes
 * (0 = 2^22) gather their colours in head order inside the bin; the colours
 * of the other gathered heads are scattered to their lines' slots once per
 * round, one workgroup per scatter_lines of them in head order: by default
 * (0 = 16384, at most 16384, 8192 on the hashed form) the workgroup reads
 * the colours in head order into LDS and writes them in slot order (stores
 * in runs); with RBSC_BIN_SLOT_ORDER (0 = 2048) the records are kept in slot
 * order inside each workgroup's share (colours read out of order); bins with
 * more than chunk_slots lines (0 = 2^19) are split over several
 * workgroups.  Of the hot heads, those below dense_heads (0 = 2^21; at least
 * hot_heads: none warm) are gathered by the bin pass itself; the lines into
 * the others (warm) get their colours once per round from a gather pass in
 * which each XCD reads only its own share of the warm colours (segments of
 * about equal lines), so they stay in that XCD's L2.  opts may be null (all defaults).  Later refinements of g use
 * the layout unless RBSC_FLAG_NOBINS or streamed lines are asked for (or
 * RBSC_FLAG_HASHED64 on the raw form); relayout, split_hot and
 * lines_to_host drop it.  Signatures and results are identical.  No counterpart in the reference (its
 * set_values + ReduceByKey, rbsc.cu:357-377, read the lines in input
 * order). */
typedef struct {
    uint32_t hot_heads;
    uint32_t rows_per_bin;
    uint64_t chunk_slots;
    uint32_t scatter_lines;
    uint32_t flags;           /* RBSC_BIN_* */
    uint32_t dense_heads;
    uint32_t reserved;        /* 0 */
} rbsc_bin_opts;
#define RBSC_BIN_SLOT_ORDER 1u
/* Hashed form of the layout: refinements keep u64 hashed colours per node,
 * randomize(c + r) computed once per node per round by the rank pass; hot
 * lines gather them and cold slots receive them, so no line hashes.  The
 * default (raw) form gathers u32 raw colours and hashes each line (four
 * 64-bit multiplies per line): it moves half the gathered bytes and is the
 * faster of the two on large graphs, where the gathers bound the pass.
 * Results are identical. */
#define RBSC_BIN_HASHED 2u
int rbsc_graph_bin(rbsc_ctx* ctx, rbsc_dgraph* g, const rbsc_bin_opts* opts);
int rbsc_graph_unbin(rbsc_ctx* ctx, rbsc_dgraph* g);
/* info[10]: binned (0/1), bins, hot lines (dense and warm), cold lines,
 * workgroup chunks, bins split over several chunks, bytes, hot_heads |
 * log2(rows_per_bin) << 32 | hashed form << 40, dense_heads, warm lines. */
int rbsc_graph_bin_info(const rbsc_dgraph* g, uint64_t* info);

/* Per-phase device timing with hipEvents on the context's stream (off by
 * default).  Phases: 0 signature kernel, 1 key gather, 2 radix sort,
 * 3 rank/scan, 4 round-0 set-up, 5 signature finish (tile carries, repeated
 * rows), 6 cold-colour scatter of a binned graph.  rbsc_ctx_profile_read
 * fills ms[7] (summed milliseconds) and launches[7] (phase executions) and
 * resets the sums when reset != 0.  No counterpart in the reference, which times
 * only the whole run (rbsc.cu:272, 489-490). */
int rbsc_ctx_profile(rbsc_ctx* ctx, int enable);
int rbsc_ctx_profile_read(rbsc_ctx* ctx, double* ms, uint64_t* launches, int reset);
/* Counters of the node side's slow paths since the context was made (or
 * last reset): out[0] rounds whose node side was redone with a full sort
 * (a tie run list overflowed, or equal signatures out of colour order),
 * out[1] rounds with tie runs too long for the in-kernel sorts, out[2]
 * last rounds confirmed without ranking (RBSC_FLAG_FULLROUNDS); out holds 3
 * words.  No counterpart in the reference. */
int rbsc_ctx_stats(rbsc_ctx* ctx, uint64_t* out, int reset);

/* Exact, hash-free stability check of a colouring of g (colors: one u32
 * label per node id, any values).  *bad_out = number of violations: 0 iff
 * every node has the same multiset of (block of head, summed weight) over
 * its lines as every other node of its block and no block mixes initial
 * colours -- then the colouring is the coarsest stable partition the round
 * function converges to, with no 64-bit signature collision.  No
 * counterpart in the reference (its collision loop rbsc.cu:404-419 checks
 * only the colour hash); rbsc_opts.verify runs it on the final colours. */
int rbsc_graph_check_partition(rbsc_ctx* ctx, rbsc_dgraph* g, const uint32_t* colors,
                               uint64_t* bad_out);

/* ---- Batches of independent graphs (no counterpart in the reference, which
 * refines one graph per process).  Graph g owns nodes
 * [node_offsets[g], node_offsets[g+1]) and lines
 * [line_offsets[g], line_offsets[g+1]); src and dst are local node ids
 * (0 .. n_g - 1) and a graph's lines are grouped by source, as for
 * rbsc_graph.  init_color (nullable) is indexed by batch node.  Graphs of at
 * most 5120 nodes are refined inside one workgroup's LDS, many graphs per
 * persistent launch; larger ones go through the single-graph path.  Each
 * graph gets exactly the colours, rounds and blocks it gets alone. */
typedef struct rbsc_batch {
    uint32_t n_graphs;
    const uint64_t* node_offsets;
    const uint64_t* line_offsets;
    const uint32_t* src;
    const uint32_t* dst;
    const uint64_t* weight;
    const uint32_t* init_color;
} rbsc_batch;
typedef struct rbsc_dbatch rbsc_dbatch;

int rbsc_batch_upload(rbsc_ctx* ctx, const rbsc_batch* batch, rbsc_dbatch** out);
int rbsc_batch_free(rbsc_ctx* ctx, rbsc_dbatch* b);
/* blocks_out, rounds_out: per graph; labels_out: first-occurrence labels per
 * batch node, canonical within each graph; colors_out: raw final colours per
 * batch node.  All nullable.  total: sums blocks, max rounds, times. */
int rbsc_batch_refine(rbsc_ctx* ctx, rbsc_dbatch* b, const rbsc_opts* opts, uint32_t* blocks_out,
                      uint32_t* rounds_out, uint32_t* labels_out, uint32_t* colors_out,
                      rbsc_result* total);
/* Outputs of rbsc_batch_refine_ex, all nullable.  blocks, rounds: per graph;
 * labels, colors: per batch node as for rbsc_batch_refine; block_sizes: per
 * batch node, the final colour histogram of each graph: graph g's entries
 * [node_offsets[g], node_offsets[g+1]) hold its block sizes in descending
 * order (blocks[g] nonzero entries, summing to n_g), then zeros.  Block sizes
 * do not depend on how blocks are numbered, so they compare across graphs and
 * tools.  Buffers from rbsc_host_alloc are written by the kernels directly. */
typedef struct rbsc_batch_outputs {
    uint32_t* blocks;
    uint32_t* rounds;
    uint32_t* labels;
    uint32_t* colors;
    uint32_t* block_sizes;
} rbsc_batch_outputs;
/* rbsc_batch_refine with the per-graph block-size histogram as a further
 * output.  No counterpart in the reference (one graph per process; its
 * output is the block count, rbsc.cu:492-494). */
int rbsc_batch_refine_ex(rbsc_ctx* ctx, rbsc_dbatch* b, const rbsc_opts* opts,
                         const rbsc_batch_outputs* out, rbsc_result* total);
/* Batched refinement of a batch already in device memory, allocating
 * nothing.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The fields of `batch` after
 * n_graphs are device pointers (weight and init_color nullable), with each
 * graph's lines sorted by source; n_nodes and n_lines are the totals
 * node_offsets[n_graphs] and line_offsets[n_graphs].  Outputs are device
 * arrays: blocks and rounds (n_graphs entries, required), labels and colors
 * (n_nodes entries, nullable) as for rbsc_batch_refine; each graph gets
 * exactly those results.  The launch layout is built in a fixed array of
 * the library on each device (about 17 MB), so a call holds at most 32768
 * graphs, 2^21 nodes and 2^22 lines, each graph at most 5120 nodes; larger
 * batches are rejected (RBSC_ERR_INVALID) before any device work.  Calls on
 * one device share that array: calls that overlap on one device race.
 * Synchronous on the context's stream. */
int rbsc_batch_refine_device(rbsc_ctx* ctx, const rbsc_batch* batch, uint64_t n_nodes,
                             uint64_t n_lines, const rbsc_opts* opts, uint32_t* blocks,
                             uint32_t* rounds, uint32_t* labels, uint32_t* colors);
/* Tunes the persistent launches on this batch and keeps the fastest for
 * later refines: the largest graphs' workgroup width (256/512/1024) and
 * count (1/2/4 per CU), then for each of the four size tiers the digit
 * width of its LDS radix sort (8 bits: 4 passes over the 32 sorted
 * signature bits; or 4 bits: 8 passes with 16-digit histograms).
 * best_ms_out: device time of one launch with the chosen parameters. */
int rbsc_batch_autotune(rbsc_ctx* ctx, rbsc_dbatch* b, uint32_t* threads_out, uint32_t* wgs_out,
                        double* best_ms_out);
int rbsc_batch_set_launch(rbsc_ctx* ctx, rbsc_dbatch* b, uint32_t threads, uint32_t wgs);
/* info[10]: workgroup width and count of the largest tier, digit bits of
 * tiers 0..3 (largest graphs first), graphs in tiers 0..3. */
int rbsc_batch_launch_info(const rbsc_dbatch* b, uint32_t* info);
/* Digit bits (4 or 8) of the LDS radix sort of tiers 0..3. */
int rbsc_batch_set_digits(rbsc_dbatch* b, const uint32_t* digit_bits);
/* One-shot: upload, refine, free. */
int rbsc_refine_batch(rbsc_ctx* ctx, const rbsc_batch* batch, const rbsc_opts* opts,
                      uint32_t* blocks_out, uint32_t* rounds_out, uint32_t* labels_out,
                      rbsc_result* total);

/* ---- Serving: a request loop on one device (no counterpart in the
 * reference, which refines one graph per process).  Requests arrive at
 * their arrival time (submit time + delay_s), wait in FIFO order until their
 * HBM estimate fits hbm_budget, then run: graphs of at most 5120 nodes are
 * gathered into batch launches (every small request queued when the batcher
 * is free goes into its next launch), larger graphs run on `workers`
 * concurrent contexts (one stream and workspace each).  Results are those
 * of rbsc_graph_refine on each graph alone. */
typedef struct rbsc_server rbsc_server;
typedef struct rbsc_server_opts {
    int device;
    uint64_t hbm_budget;      /* bytes for requests in flight; 0 = 250 GiB */
    uint32_t workers;         /* concurrent single-graph refinements; 0 = 8, at most 16 */
    uint32_t batch_max;       /* most small graphs per batch launch; 0 = 4096 */
    uint32_t keep_partition;  /* 1: results carry the canonical partition */
} rbsc_server_opts;
typedef struct rbsc_request_result {
    uint64_t ticket;
    int32_t status;           /* RBSC_OK or the request's error code */
    uint32_t nodes;           /* blocks of the final partition */
    uint32_t rounds;
    uint32_t batched;         /* 1: refined in a batch launch */
    uint32_t n_nodes;
    uint64_t n_edges;
    double latency_s;         /* arrival -> result */
    double queue_s;           /* arrival -> admission */
    double service_s;         /* start of work -> result */
    double arrival_s, done_s; /* steady-clock seconds */
} rbsc_request_result;
typedef struct rbsc_server_stats {
    uint64_t submitted, completed, failed, batches, batched_graphs, in_flight;
    uint64_t admitted_bytes, peak_admitted_bytes, budget;
} rbsc_server_stats;
#define RBSC_PENDING 1        /* rbsc_server_wait: not finished within the timeout */

int rbsc_server_create(const rbsc_server_opts* opts, rbsc_server** out);
/* Waits until every submitted request has finished, then stops. */
int rbsc_server_destroy(rbsc_server* s);
/* A host graph (copied; uploaded at admission, freed after). */
int rbsc_server_submit(rbsc_server* s, const rbsc_graph* g, double delay_s, uint64_t* ticket);
/* An R-MAT graph as rbsc_gen_rmat_device / rbsc_gen_rmat build it, made at
 * admission (small graphs on the host, others on the device). */
int rbsc_server_submit_rmat(rbsc_server* s, uint32_t scale, uint64_t n_draws, double a, double b,
                            double c, double clone_frac, uint64_t seed, double delay_s,
                            uint64_t* ticket);
/* A graph already resident on the server's device (owned by the caller,
 * read only; it must outlive the request). */
int rbsc_server_submit_resident(rbsc_server* s, rbsc_dgraph* g, double delay_s, uint64_t* ticket);
/* Waits up to timeout_s (< 0: for ever).  Returns the request's status
 * (RBSC_OK or its error) and releases it, or RBSC_PENDING.  partition_out
 * (nullable, n_nodes words) receives the canonical partition when
 * keep_partition is set. */
int rbsc_server_wait(rbsc_server* s, uint64_t ticket, double timeout_s, rbsc_request_result* out,
                     uint32_t* partition_out);
/* The same without copying the partition: *labels (n_nodes words, or null
 * without keep_partition) is page-locked memory the refinement wrote its
 * labels into, the caller's until rbsc_server_labels_free(labels) hands it
 * back for reuse; it stays valid after the server is destroyed.  No
 * counterpart in the reference (its result is the block count only). */
int rbsc_server_wait_labels(rbsc_server* s, uint64_t ticket, double timeout_s,
                            rbsc_request_result* out, uint32_t** labels);
int rbsc_server_labels_free(uint32_t* labels);
int rbsc_server_get_stats(rbsc_server* s, rbsc_server_stats* out);

/* Seeded R-MAT generated on the device (bench-scale inputs without a host
 * round trip).  2^scale nodes, n_draws lines drawn with probabilities a, b, c
 * (d = 1 - a - b - c), node ids permuted by a seeded bijection, then
 * floor(clone_frac * 2^scale) nodes replace their lines with a copy of a
 * random non-clone's lines.  Lines are sorted by (src, dst).  Identical to
 * rbsc_gen_rmat on the host for the same arguments. */
int rbsc_gen_rmat_device(rbsc_ctx* ctx, uint32_t scale, uint64_t n_draws, double a,
                         double b, double c, double clone_frac, uint64_t seed,
                         rbsc_dgraph** out);
/* Copy a resident graph back to host CSR (for checks): row_offsets has
 * n_nodes + 1 entries, dst n_edges entries.  Sizes via rbsc_graph_dims. */
int rbsc_graph_dims(const rbsc_dgraph* g, uint32_t* n_nodes, uint64_t* n_edges);
/* Rows and lines the signature pass visits: nodes with lines, one per group
 * of identical rows (same heads and weights in the same order), whose
 * signatures the repeats copy.  No counterpart in the reference. */
uint32_t rbsc_graph_rows(const rbsc_dgraph* g);
uint64_t rbsc_graph_stored_lines(const rbsc_dgraph* g);
/* Graphs built on ctx after this call store identical rows once (on, the
 * default) or every row (off); results are identical either way. */
int rbsc_ctx_set_dedup(rbsc_ctx* ctx, int on);
int rbsc_graph_download(rbsc_ctx* ctx, const rbsc_dgraph* g, uint64_t* row_offsets,
                        uint32_t* dst);

#ifdef __cplusplus
}
#endif
#endif /* RBSC_H */
