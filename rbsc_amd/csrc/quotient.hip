// quotient.hip -- the reference's multi-batch quotient mode (rbsc.cu:274-487,
// stop rule variant rbsc_aggressive.cu:487), resident on the device.
//
// What is computed is the reference's.  Per pass: the line list is cut into
// batches of max_batch_edges lines trimmed to whole source groups; each batch
// is refined to its own fixpoint with nodes outside the batch frozen as
// singletons (colour i+1, signature 2^64-1) and, in the first batch only,
// nodes without any line as colour 0 / signature 0; local colours merge into
// global labels by first occurrence; with more than one batch, nodes no batch
// labelled get fresh labels and the line list is contracted to the quotient
// graph (the first group of every label kept, its lines relabelled, parallel
// lines merged by summed weight, in order of their first line).  Passes
// repeat while more than one batch is needed and the node count shrinks.
// done = (last pass had one batch).
//
// How it is computed is the device's own.  The lines are uploaded once and
// every pass works on them in HBM; the host reads only scalars (the group
// count, the batch plan, label totals) between launches.
//   * groups: a start flag per line and an exclusive scan give every group
//     its index; the first start of each source (atomicMin) names the line
//     that breaks grouping, and each source's group index is kept per node;
//   * batch plan: two binary searches over the group starts per batch give
//     its trimmed line range and its groups [glo, ghi);
//   * a batch: one kernel writes its initial colours and signature defaults
//     from the node's group index, its rows are a shifted slice of the group
//     starts, its lines a copy of its slice; it is built as a device graph
//     and refined by the device round;
//   * merging: the first node of each local colour (atomicMin over the
//     batch's nodes), a flag scan over nodes, one write per node;
//   * contraction: the first group of each label (atomicMin over groups),
//     the kept lines keyed (group, head label) and sorted by the onesweep
//     radix sort, each run's weight a difference of one prefix sum, and the
//     runs placed in order of their first line by one more flag scan.
// This is the reference's own batch semantics, kept for `done`/`nodes`
// parity with its cuda_50 / cuda_25 columns; the exact alternative for
// graphs that exceed HBM is the streaming mode of refine.hip.
#include "internal.h"
#include "scan.h"
#include <chrono>
#include <cmath>
#include <vector>

namespace {

constexpr int BLK = 256;
inline unsigned grid_for(u64 n) { return (unsigned)((n + BLK - 1) / BLK); }
constexpr u32 NONE = 0xFFFFFFFFu;
constexpr u64 NONE64 = ~0ULL;

__device__ __forceinline__ u64 lower_bound_u64(const u64* __restrict__ a, u64 n, u64 x) {
    u64 lo = 0, hi = n;
    while (lo < hi) {
        const u64 m = lo + (hi - lo) / 2;
        if (a[m] < x) lo = m + 1; else hi = m;
    }
    return lo;
}

__global__ void fill32_kernel(u32* __restrict__ p, u64 n, u32 v) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void fill64_kernel(u64* __restrict__ p, u64 n, u64 v) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void iota32_kernel(u32* __restrict__ p, u64 n) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (u32)i;
}

// Sources of CSR lines: line e belongs to the last row u with ro[u] <= e.
// The check first: bad[2] = least index where ro stops being a monotone
// offset array from 0 to E.
__global__ void csr_src_kernel(const u64* __restrict__ ro, u64 N, u64 E, u32* __restrict__ src) {
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    // last u in [0, N) with ro[u] <= e: lower_bound(e + 1) - 1 over ro[0..N)
    src[e] = (u32)(lower_bound_u64(ro, N, e + 1) - 1);
}
__global__ void csr_check_kernel(const u64* __restrict__ ro, u64 N, u64 E, u64* __restrict__ bad) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u > N) return;
    const bool ok = u == 0 ? ro[0] == 0 : (u == N ? ro[N] == E && ro[u - 1] <= ro[u]
                                                  : ro[u - 1] <= ro[u]);
    if (!ok) atomicMin((unsigned long long*)&bad[2], (unsigned long long)u);
}

// Group start flags and the input checks of the first pass: bad[0] = first
// line with a node id >= N; first_start[v] = first line starting a group of v.
__global__ void start_flag_kernel(const u32* __restrict__ src, const u32* __restrict__ dst, u64 E,
                                  u64 N, u64* __restrict__ flag, u64* __restrict__ first_start,
                                  u64* __restrict__ bad) {
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const u32 s = src[e];
    if (s >= N || dst[e] >= N) atomicMin((unsigned long long*)&bad[0], (unsigned long long)e);
    const bool start = e == 0 || src[e - 1] != s;
    flag[e] = start;
    if (start && s < N) atomicMin((unsigned long long*)&first_start[s], (unsigned long long)e);
}
// Group g = pos[e] of each start e: its first line, its source, and the
// source's group index; bad[1] = first start of a source seen before.
__global__ void groups_kernel(const u32* __restrict__ src, u64 E, u64 N, const u64* __restrict__ pos,
                              const u64* __restrict__ first_start, u64* __restrict__ gs,
                              u32* __restrict__ gsrc, u32* __restrict__ node_group,
                              u64* __restrict__ bad) {
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const u32 s = src[e];
    if (!(e == 0 || src[e - 1] != s)) return;
    const u64 g = pos[e];
    gs[g] = e;
    gsrc[g] = s;
    if (s >= N) return;
    if (first_start[s] != e) atomicMin((unsigned long long*)&bad[1], (unsigned long long)e);
    else node_group[s] = (u32)g;
}

// Batch b's nominal lines [b B, min((b + 1) B, E)) trimmed to whole groups
// (rbsc.cu:315-334): it starts at the first group start inside the range
// and ends at the range's end when that is a group boundary, else at the
// last group start inside it -- counting a start at line 0 only as a start,
// so a first group longer than the range leaves batch 0 without an end.
// plan[4b..4b+3] = (glo, ghi, first line, end line), glo = NONE64 when the
// batch is skipped.  gs holds G + 1 entries, gs[G] = E.
__global__ void plan_kernel(const u64* __restrict__ gs, u64 G, u64 E, u64 B, u64 nb,
                            u64* __restrict__ plan) {
    const u64 b = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const u64 bs = b * B;
    const u64 be = bs + B < E ? bs + B : E;
    u64 glo = lower_bound_u64(gs, G, bs), ghi = 0;
    bool ok = glo < G && gs[glo] < be;
    if (ok) {
        if (be == E) {
            ghi = G;
        } else {
            const u64 j = lower_bound_u64(gs, G, be);
            if (j < G && gs[j] == be) {
                ghi = j;
            } else {
                ghi = j - 1;  // >= glo, since gs[glo] < be
                ok = gs[ghi] != 0;
            }
        }
    }
    plan[4 * b] = ok ? glo : NONE64;
    plan[4 * b + 1] = ok ? ghi : 0;
    plan[4 * b + 2] = ok ? gs[glo] : 0;
    plan[4 * b + 3] = ok ? gs[ghi] : 0;
}

// A node's part in batch (glo, ghi): in the batch (its group is one of the
// batch's), dangling (no lines, first batch only) or off-batch.
__device__ __forceinline__ bool in_batch(u32 grp, u64 glo, u64 ghi) {
    return grp != NONE && grp >= glo && grp < ghi;
}
__device__ __forceinline__ bool selected(u32 grp, u64 glo, u64 ghi, int first) {
    return in_batch(grp, glo, ghi) || (first && grp == NONE);
}

// Initial colours (rbsc.cu:51-63 with the batch mask of 310-336): in the
// batch N+1, dangling 0, off-batch i+1; signature defaults (104-116):
// off-batch 2^64-1, the rest 0.
__global__ void batch_init_kernel(const u32* __restrict__ node_group, u64 n, u64 glo, u64 ghi,
                                  int first, u32* __restrict__ init, u64* __restrict__ fill) {
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const u32 grp = node_group[v];
    const bool in = in_batch(grp, glo, ghi), dang = first && grp == NONE;
    init[v] = in ? (u32)(n + 1) : (dang ? 0u : (u32)(v + 1));
    fill[v] = (in || dang) ? 0ULL : ~0ULL;
}

// The batch's rows: groups [glo, glo + U), row i starting at gs[glo + i] - base.
__global__ void batch_rows_kernel(const u64* __restrict__ gs, const u32* __restrict__ gsrc, u64 glo,
                                  u64 U, u64 base, u64* __restrict__ row_ptr,
                                  u32* __restrict__ row_node) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > U) return;
    row_ptr[i] = gs[glo + i] - base;
    if (i < U) row_node[i] = gsrc[glo + i];
}

__global__ void map_kernel(const u32* __restrict__ by_internal, const u32* __restrict__ pi, u64 n,
                           u32* __restrict__ out) {
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) out[v] = by_internal[pi[v]];
}

// First-occurrence labels of the selected nodes' colours (rbsc.cu:431-445):
// first[c] = least selected node of colour c, then a flag at each such node,
// an exclusive scan, and label = base + rank of the colour's first node.
// all != 0 selects every node.
__global__ void first_kernel(const u32* __restrict__ col, u64 n, const u32* __restrict__ node_group,
                             u64 glo, u64 ghi, int first_batch, int all, u32* __restrict__ first) {
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    if (all || selected(node_group[v], glo, ghi, first_batch)) atomicMin(&first[col[v]], (u32)v);
}
__global__ void first_flag_kernel(const u32* __restrict__ col, u64 n,
                                  const u32* __restrict__ node_group, u64 glo, u64 ghi,
                                  int first_batch, int all, const u32* __restrict__ first,
                                  u64* __restrict__ flag) {
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const bool sel = all || selected(node_group[v], glo, ghi, first_batch);
    flag[v] = sel && first[col[v]] == (u32)v;
}
__global__ void first_label_kernel(const u32* __restrict__ col, u64 n,
                                   const u32* __restrict__ node_group, u64 glo, u64 ghi,
                                   int first_batch, int all, const u32* __restrict__ first,
                                   const u64* __restrict__ pos, u64 base, u32* __restrict__ w) {
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    if (all || selected(node_group[v], glo, ghi, first_batch))
        w[v] = (u32)(base + pos[first[col[v]]]);
}

// Nodes no batch labelled (w == n) get fresh labels in node order
// (rbsc.cu:457-461).
__global__ void unlabelled_flag_kernel(const u32* __restrict__ w, u64 n, u64* __restrict__ flag) {
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) flag[v] = w[v] == (u32)n;
}
__global__ void unlabelled_label_kernel(u32* __restrict__ w, u64 n, const u64* __restrict__ pos,
                                        u64 base) {
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && w[v] == (u32)n) w[v] = (u32)(base + pos[v]);
}

// Contraction (rbsc.cu:463-485): the first group of each label is kept.
__global__ void rep_kernel(const u32* __restrict__ gsrc, u64 G, const u32* __restrict__ w,
                           u32* __restrict__ rep) {
    const u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < G) atomicMin(&rep[w[gsrc[g]]], (u32)g);
}
__global__ void keep_flag_kernel(const u32* __restrict__ src, u64 E, const u32* __restrict__ w,
                                 const u32* __restrict__ node_group, const u32* __restrict__ rep,
                                 u64* __restrict__ flag) {
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const u32 s = src[e];
    flag[e] = rep[w[s]] == node_group[s];
}
// Kept line e -> key (group << bl) | label(head), value e, at its rank k.
__global__ void keep_keys_kernel(const u32* __restrict__ src, const u32* __restrict__ dst, u64 E,
                                 const u32* __restrict__ w, const u32* __restrict__ node_group,
                                 const u32* __restrict__ rep, const u64* __restrict__ pos, u32 bl,
                                 u64* __restrict__ keys, u64* __restrict__ vals) {
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const u32 s = src[e];
    const u32 g = node_group[s];
    if (rep[w[s]] != g) return;
    const u64 k = pos[e];
    keys[k] = ((u64)g << bl) | w[dst[e]];
    vals[k] = e;
}
// Sorted kept lines: their weights in sorted order, and a flag at the head
// of every run of equal keys.
__global__ void run_flag_kernel(const u64* __restrict__ keys, const u64* __restrict__ vals, u64 n,
                                const u64* __restrict__ wt, u64* __restrict__ wsorted,
                                u64* __restrict__ head) {
    const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    wsorted[j] = wt ? wt[vals[j]] : 1ULL;
    head[j] = j == 0 || keys[j] != keys[j - 1];
}
// Run r starts at sorted position runs[r]; runs[nr] = n.  The run's first
// line (the least e, the sort being stable over lines in order) is flagged
// in mark (zeroed over the E lines beforehand).
__global__ void run_start_kernel(const u64* __restrict__ keys, const u64* __restrict__ vals, u64 n,
                                 const u64* __restrict__ rid, u64* __restrict__ runs,
                                 u64* __restrict__ mark) {
    const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (j == 0 || keys[j] != keys[j - 1]) {
        runs[rid[j]] = j;
        mark[vals[j]] = 1;
    }
}
// Run r -> quotient line slot[first line]: (label(source), label(head),
// summed weight), the sum as a difference of the exclusive prefix of the
// sorted weights (mod 2^64, as the reference's u64 sums).
__global__ void emit_kernel(const u64* __restrict__ keys, const u64* __restrict__ vals,
                            const u64* __restrict__ runs, u64 nr, const u64* __restrict__ wpre,
                            const u64* __restrict__ slot, const u32* __restrict__ src,
                            const u32* __restrict__ w, u32 bl, u32* __restrict__ osrc,
                            u32* __restrict__ odst, u64* __restrict__ owt) {
    const u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const u64 j0 = runs[r], j1 = runs[r + 1];
    const u64 e0 = vals[j0];
    const u64 s = slot[e0];
    osrc[s] = w[src[e0]];
    odst[s] = (u32)(keys[j0] & ((1ULL << bl) - 1));
    owt[s] = wpre[j1] - wpre[j0];
}

// The class of every original node: its label in this pass.
__global__ void compose_kernel(u32* __restrict__ comp, u64 N0, const u32* __restrict__ w, u64 n) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N0) return;
    const u32 c = comp[i];
    if (c < n) comp[i] = w[c];
}

// Device buffers of one quotient run, freed together.
struct Bufs {
    hipStream_t s = nullptr;  // frees are ordered on this stream
    std::vector<void*> held;
    ~Bufs() {
        for (void* p : held)
            ctx_free(p, s);
    }
    template <class T>
    hipError_t get(T** out, u64 n) {
        *out = nullptr;
        const hipError_t e = ctx_alloc((void**)out, sizeof(T) * (n ? n : 1), s);
        if (e == hipSuccess) held.push_back(*out);
        return e;
    }
};

u32 bits_for(u64 n) {  // bits of the values below n (at least 1)
    u32 b = 1;
    while (b < 64 && (1ULL << b) < n) ++b;
    return b;
}

}  // namespace

int quotient_refine(rbsc_ctx* ctx, const rbsc_graph* gr, const rbsc_opts* opts, rbsc_result* res,
                    uint32_t* partition_out) {
    const u64 N0 = gr->n_nodes;
    const u64 E0 = gr->n_edges;
    const u64 max_batch = opts->max_batch_edges;
    const bool aggressive = opts->mode == RBSC_MODE_QUOTIENT_AGGRESSIVE;
    if (N0 >= 0xFFFFFFFFull) {
        rbsc_set_error("n_nodes must be below 2^32-1");
        return RBSC_ERR_INVALID;
    }
    if (E0 && !gr->dst) {
        rbsc_set_error("graph needs dst");
        return RBSC_ERR_INVALID;
    }
    if (E0 && !gr->src && !gr->row_offsets) {
        rbsc_set_error("graph needs src or row_offsets");
        return RBSC_ERR_INVALID;
    }
    const u64 NW = std::max(N0, E0) + 1;
    // lines twice (contraction ping-pong), sort keys and values twice, run
    // starts, two scan arrays, one batch's lines and the per-node arrays
    const u64 need = 32 * E0 + 40 * (E0 + 1) + 16 * NW + 12 * std::min(max_batch, E0) + 64 * (N0 + 2);
    if (ctx->budget && need > ctx->budget) {
        rbsc_set_error("quotient mode needs %llu bytes, budget is %llu", (unsigned long long)need,
                       (unsigned long long)ctx->budget);
        return RBSC_ERR_BUDGET;
    }
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = ctx->stream;
    Bufs B;
    B.s = s;
    u32 *src[2], *dst[2], *node_group, *gsrc, *init, *d_int, *col, *first, *w, *comp, *rep, *brn,
        *bdst;
    u64 *wt[2], *flag, *pos, *tmp, *first_start, *gs, *fill, *brp, *bwt, *keys[2], *vals[2], *runs,
        *plan = nullptr, *scal;
    u64 plan_cap = 0;
    const u64 BL = std::min(max_batch, E0);
#define Q_TRY(expr)                                                                       \
    do {                                                                                  \
        const hipError_t _e = (expr);                                                     \
        if (_e != hipSuccess) {                                                           \
            rbsc_set_error("%s failed: %s", #expr, hipGetErrorString(_e));                \
            return _e == hipErrorOutOfMemory ? RBSC_ERR_NOMEM : RBSC_ERR_HIP;             \
        }                                                                                 \
    } while (0)
#define Q_RC(expr)                  \
    do {                            \
        const int _rc = (expr);     \
        if (_rc) return _rc;        \
    } while (0)
    Q_TRY(B.get(&src[0], E0));
    Q_TRY(B.get(&src[1], E0));
    Q_TRY(B.get(&dst[0], E0));
    Q_TRY(B.get(&dst[1], E0));
    Q_TRY(B.get(&wt[0], E0));
    Q_TRY(B.get(&wt[1], E0));
    // lines without weights weigh 1 until the first contraction sums them
    bool weighted = gr->weight != nullptr;
    Q_TRY(B.get(&flag, NW));
    Q_TRY(B.get(&pos, NW));
    Q_TRY(B.get(&tmp, scan_tmp_words(NW)));
    Q_TRY(B.get(&first_start, N0));
    Q_TRY(B.get(&gs, N0 + 1));
    Q_TRY(B.get(&gsrc, N0));
    Q_TRY(B.get(&node_group, N0));
    Q_TRY(B.get(&init, N0));
    Q_TRY(B.get(&fill, N0));
    Q_TRY(B.get(&d_int, N0));
    Q_TRY(B.get(&col, N0));
    Q_TRY(B.get(&first, N0 + 2));
    Q_TRY(B.get(&w, N0));
    Q_TRY(B.get(&comp, N0));
    Q_TRY(B.get(&rep, N0));
    Q_TRY(B.get(&brp, N0 + 1));
    Q_TRY(B.get(&brn, N0));
    Q_TRY(B.get(&bdst, BL));
    Q_TRY(B.get(&bwt, BL));
    for (int i = 0; i < 2; ++i) {
        Q_TRY(B.get(&keys[i], E0 + 1));
        Q_TRY(B.get(&vals[i], E0 + 1));
    }
    Q_TRY(B.get(&runs, E0 + 1));
    Q_TRY(B.get(&scal, 4));
    u64* h = reinterpret_cast<u64*>(ctx->h_scalar);  // pinned mirror, >= 4 u64
    // Upload once.
    if (E0) {
        Q_TRY(hipMemcpyAsync(dst[0], gr->dst, 4 * E0, hipMemcpyHostToDevice, s));
        if (gr->weight) Q_TRY(hipMemcpyAsync(wt[0], gr->weight, 8 * E0, hipMemcpyHostToDevice, s));
    }
    fill64_kernel<<<1, 4, 0, s>>>(scal, 4, NONE64);
    if (gr->row_offsets) {
        Q_TRY(hipMemcpyAsync(gs, gr->row_offsets, 8 * (N0 + 1), hipMemcpyHostToDevice, s));
        csr_check_kernel<<<grid_for(N0 + 1), BLK, 0, s>>>(gs, N0, E0, scal);
        Q_TRY(hipMemcpyAsync(h, scal, 8 * 4, hipMemcpyDeviceToHost, s));
        Q_TRY(hipStreamSynchronize(s));
        if (h[2] != NONE64) {
            rbsc_set_error("row_offsets[%llu] breaks the offsets from 0 to E", (unsigned long long)h[2]);
            return RBSC_ERR_INPUT;
        }
        if (E0) csr_src_kernel<<<grid_for(E0), BLK, 0, s>>>(gs, N0, E0, src[0]);
    } else if (E0) {
        Q_TRY(hipMemcpyAsync(src[0], gr->src, 4 * E0, hipMemcpyHostToDevice, s));
    }
    if (N0) iota32_kernel<<<grid_for(N0), BLK, 0, s>>>(comp, N0);
    Q_TRY(hipGetLastError());

    int cur = 0;
    u64 node_n, new_node_n = N0, edge_n, new_edge_n = E0, batches = 1;
    u32 passes = 0, rounds = 0;
    double dev_ms = 0;
    std::vector<u64> hplan;
    for (;;) {
        node_n = new_node_n;
        edge_n = new_edge_n;
        new_node_n = 0;
        batches = (u64)ceil((double)edge_n / (double)max_batch);
        ++passes;
        // Groups of the current line list.
        u64 G = 0;
        if (edge_n) {
            fill64_kernel<<<grid_for(node_n), BLK, 0, s>>>(first_start, node_n, NONE64);
            fill32_kernel<<<grid_for(node_n), BLK, 0, s>>>(node_group, node_n, NONE);
            fill64_kernel<<<1, 4, 0, s>>>(scal, 4, NONE64);
            start_flag_kernel<<<grid_for(edge_n), BLK, 0, s>>>(src[cur], dst[cur], edge_n, node_n,
                                                               flag, first_start, scal);
            Q_RC(scan_excl_u64(flag, pos, edge_n, scal + 3, tmp, s));
            Q_TRY(hipMemcpyAsync(h, scal, 8 * 4, hipMemcpyDeviceToHost, s));
            Q_TRY(hipStreamSynchronize(s));
            if (h[0] != NONE64) {
                rbsc_set_error("line %llu: node id >= N=%llu", (unsigned long long)h[0],
                               (unsigned long long)node_n);
                return RBSC_ERR_INPUT;
            }
            G = h[3];
            groups_kernel<<<grid_for(edge_n), BLK, 0, s>>>(src[cur], edge_n, node_n, pos, first_start,
                                                           gs, gsrc, node_group, scal);
            fill64_kernel<<<1, 1, 0, s>>>(gs + G, 1, edge_n);
            Q_TRY(hipMemcpyAsync(h, scal, 8 * 4, hipMemcpyDeviceToHost, s));
            Q_TRY(hipStreamSynchronize(s));
            if (h[1] != NONE64) {
                u32 sv = 0;
                Q_TRY(hipMemcpy(&sv, src[cur] + h[1], 4, hipMemcpyDeviceToHost));
                rbsc_set_error("line %llu: lines of source %u are not grouped",
                               (unsigned long long)h[1], sv);
                return RBSC_ERR_INPUT;
            }
        } else {
            fill32_kernel<<<grid_for(node_n), BLK, 0, s>>>(node_group, node_n, NONE);
        }
        // Batch plan.
        if (batches > plan_cap) {
            plan_cap = batches;
            Q_TRY(B.get(&plan, 4 * plan_cap));
        }
        hplan.assign(4 * batches, 0);
        if (batches) {
            plan_kernel<<<grid_for(batches), BLK, 0, s>>>(gs, G, edge_n, max_batch, batches, plan);
            Q_TRY(hipMemcpyAsync(hplan.data(), plan, 8 * 4 * batches, hipMemcpyDeviceToHost, s));
            Q_TRY(hipStreamSynchronize(s));
        }
        fill32_kernel<<<grid_for(node_n), BLK, 0, s>>>(w, node_n, (u32)node_n);
        for (u64 b = 0; b < batches; ++b) {
            const u64 glo = hplan[4 * b], ghi = hplan[4 * b + 1];
            const u64 nbs = hplan[4 * b + 2], nbe = hplan[4 * b + 3];
            // skipped, or empty: an empty trimmed range freezes every node and
            // so labels none (rbsc.cu:336, 441)
            if (glo == NONE64 || ghi == glo) continue;
            const int fb = b == 0;
            const u64 U = ghi - glo, bE = nbe - nbs;
            const u64 off_n = node_n - U - (fb ? node_n - G : 0);
            batch_init_kernel<<<grid_for(node_n), BLK, 0, s>>>(node_group, node_n, glo, ghi, fb, init,
                                                               fill);
            batch_rows_kernel<<<grid_for(U + 1), BLK, 0, s>>>(gs, gsrc, glo, U, nbs, brp, brn);
            Q_TRY(hipMemcpyAsync(bdst, dst[cur] + nbs, 4 * bE, hipMemcpyDeviceToDevice, s));
            const u64* bw = nullptr;
            if (weighted) {
                Q_TRY(hipMemcpyAsync(bwt, wt[cur] + nbs, 8 * bE, hipMemcpyDeviceToDevice, s));
                bw = bwt;
            }
            RawGraph raw;
            raw.N = node_n;
            raw.E = bE;
            raw.U = U;
            raw.row_ptr = brp;
            raw.row_node = brn;
            raw.dst = bdst;
            raw.wt = const_cast<u64*>(bw);
            raw.init = init;
            raw.sample_indeg = true;  // refined once
            rbsc_dgraph* g = new rbsc_dgraph();
            int rc = dgraph_build(ctx, &raw, g);
            RefineOut o;
            if (!rc) {
                RefineParams prm;
                prm.d_sig_fill = fill;
                prm.k0 = (u32)(off_n + 1);
                rc = refine_core(ctx, g, prm, &o);
            }
            if (!rc) {
                device_colors(ctx, g, d_int);
                map_kernel<<<grid_for(node_n), BLK, 0, s>>>(d_int, g->pi, node_n, col);
            }
            rbsc_graph_free(ctx, g);
            if (rc) return rc;
            rounds += o.rounds;
            dev_ms += o.device_ms;
            // merge the batch's colours into global labels
            fill32_kernel<<<grid_for(node_n + 2), BLK, 0, s>>>(first, node_n + 2, NONE);
            first_kernel<<<grid_for(node_n), BLK, 0, s>>>(col, node_n, node_group, glo, ghi, fb, 0,
                                                          first);
            first_flag_kernel<<<grid_for(node_n), BLK, 0, s>>>(col, node_n, node_group, glo, ghi, fb,
                                                               0, first, flag);
            Q_RC(scan_excl_u64(flag, pos, node_n, scal, tmp, s));
            first_label_kernel<<<grid_for(node_n), BLK, 0, s>>>(col, node_n, node_group, glo, ghi, fb,
                                                                0, first, pos, new_node_n, w);
            Q_TRY(hipMemcpyAsync(h, scal, 8, hipMemcpyDeviceToHost, s));
            Q_TRY(hipStreamSynchronize(s));
            new_node_n += h[0];
        }
        if (batches > 1) {
            unlabelled_flag_kernel<<<grid_for(node_n), BLK, 0, s>>>(w, node_n, flag);
            Q_RC(scan_excl_u64(flag, pos, node_n, scal, tmp, s));
            unlabelled_label_kernel<<<grid_for(node_n), BLK, 0, s>>>(w, node_n, pos, new_node_n);
            Q_TRY(hipMemcpyAsync(h, scal, 8, hipMemcpyDeviceToHost, s));
            Q_TRY(hipStreamSynchronize(s));
            new_node_n += h[0];
            // contraction into the other line buffers
            fill32_kernel<<<grid_for(new_node_n), BLK, 0, s>>>(rep, new_node_n, NONE);
            rep_kernel<<<grid_for(G), BLK, 0, s>>>(gsrc, G, w, rep);
            keep_flag_kernel<<<grid_for(edge_n), BLK, 0, s>>>(src[cur], edge_n, w, node_group, rep,
                                                              flag);
            Q_RC(scan_excl_u64(flag, pos, edge_n, scal, tmp, s));
            Q_TRY(hipMemcpyAsync(h, scal, 8, hipMemcpyDeviceToHost, s));
            Q_TRY(hipStreamSynchronize(s));
            const u64 n_kept = h[0];
            const u32 bl = bits_for(new_node_n), bg = bits_for(G);
            const int bit_end = (int)((bl + bg + 7) / 8 * 8);
            keep_keys_kernel<<<grid_for(edge_n), BLK, 0, s>>>(src[cur], dst[cur], edge_n, w, node_group,
                                                              rep, pos, bl, keys[0], vals[0]);
            Q_RC(sort_ws_init(&ctx->sort, n_kept, ctx->d_err, s));
            int which = 0;
            Q_RC(radix_sort_u64(&ctx->sort, keys, vals, n_kept, 0, bit_end, &which, s));
            const u64* K = keys[which];
            const u64* V = vals[which];
            u64* wpre = keys[which ^ 1];  // n_kept + 1
            u64* rid = vals[which ^ 1];
            run_flag_kernel<<<grid_for(n_kept), BLK, 0, s>>>(K, V, n_kept,
                                                             weighted ? wt[cur] : nullptr, wpre, rid);
            fill64_kernel<<<1, 1, 0, s>>>(wpre + n_kept, 1, 0);
            Q_RC(scan_excl_u64(wpre, wpre, n_kept + 1, nullptr, tmp, s));
            Q_RC(scan_excl_u64(rid, rid, n_kept, scal, tmp, s));
            fill64_kernel<<<grid_for(edge_n), BLK, 0, s>>>(flag, edge_n, 0);
            run_start_kernel<<<grid_for(n_kept), BLK, 0, s>>>(K, V, n_kept, rid, runs, flag);
            Q_TRY(hipMemcpyAsync(h, scal, 8, hipMemcpyDeviceToHost, s));
            Q_TRY(hipStreamSynchronize(s));
            const u64 nr = h[0];
            fill64_kernel<<<1, 1, 0, s>>>(runs + nr, 1, n_kept);
            Q_RC(scan_excl_u64(flag, pos, edge_n, nullptr, tmp, s));
            emit_kernel<<<grid_for(nr), BLK, 0, s>>>(K, V, runs, nr, wpre, pos, src[cur], w, bl,
                                                     src[cur ^ 1], dst[cur ^ 1], wt[cur ^ 1]);
            new_edge_n = nr;
        }
        compose_kernel<<<grid_for(N0), BLK, 0, s>>>(comp, N0, w, node_n);
        Q_TRY(hipGetLastError());
        if (batches > 1) {  // the contracted lines are the next pass's
            cur ^= 1;
            weighted = true;
        }
        if (!(batches > 1 && node_n > new_node_n)) break;
        if (aggressive && !((float)batches > ceilf((float)new_edge_n / (float)max_batch))) break;
    }
    Q_TRY(hipStreamSynchronize(s));
    if (partition_out && N0) {
        // canonical first-occurrence labels of the classes
        fill32_kernel<<<grid_for(N0 + 2), BLK, 0, s>>>(first, N0 + 2, NONE);
        first_kernel<<<grid_for(N0), BLK, 0, s>>>(comp, N0, nullptr, 0, 0, 0, 1, first);
        first_flag_kernel<<<grid_for(N0), BLK, 0, s>>>(comp, N0, nullptr, 0, 0, 0, 1, first, flag);
        Q_RC(scan_excl_u64(flag, pos, N0, nullptr, tmp, s));
        first_label_kernel<<<grid_for(N0), BLK, 0, s>>>(comp, N0, nullptr, 0, 0, 0, 1, first, pos, 0,
                                                        col);
        Q_TRY(hipMemcpyAsync(partition_out, col, 4 * N0, hipMemcpyDeviceToHost, s));
        Q_TRY(hipStreamSynchronize(s));
    }
    Q_RC(check_kernel_errors(ctx));
    const auto t1 = std::chrono::steady_clock::now();
    rbsc_result r{};
    r.nodes = (u32)new_node_n;
    r.done = batches == 1;
    r.rounds = rounds;
    r.passes = passes;
    r.device_ms = dev_ms;
    r.edges_per_round = E0;
    r.time_s = std::chrono::duration<double>(t1 - t0).count();
    if (res) *res = r;
    return RBSC_OK;
#undef Q_TRY
#undef Q_RC
}
