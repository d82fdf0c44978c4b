// quotient.hip -- the reference's multi-batch quotient mode (rbsc.cu:274-487,
// stop rule variant rbsc_aggressive.cu:487) with each batch refined by the
// device round of refine.hip.
//
// Per pass: the line list is cut into batches of max_batch_edges lines
// trimmed to whole source groups; each batch is refined to its own fixpoint
// with nodes outside the batch frozen as singletons (colour i+1, signature
// 2^64-1) and nodes without any line as colour 0 / signature 0; local labels
// merge into global ones by first occurrence; then the line list is
// contracted to the quotient graph (first representative's lines, parallel
// lines merged by summed weight).  Passes repeat while more than one batch
// is needed and the node count shrinks.  done = (last pass had one batch).
//
// This is the reference's own batch semantics, kept for `done`/`nodes`
// parity with its cuda_50 / cuda_25 columns; the exact alternative for
// graphs that exceed HBM is the streaming mode of refine.hip.
#include "internal.h"
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace {
constexpr unsigned char MB_OFF = 0, MB_IN = 1, MB_DANG = 2;
}

int quotient_refine(rbsc_ctx* ctx, const rbsc_graph* gr, const rbsc_opts* opts, rbsc_result* res,
                    uint32_t* partition_out) {
    const u32 N0 = gr->n_nodes;
    const u64 E0 = gr->n_edges;
    const u64 max_batch = opts->max_batch_edges;
    const bool aggressive = opts->mode == RBSC_MODE_QUOTIENT_AGGRESSIVE;
    if (N0 == 0xFFFFFFFFu) {
        rbsc_set_error("n_nodes must be below 2^32-1");
        return RBSC_ERR_INVALID;
    }
    // Working copies of the line list (contracted in place between passes).
    std::vector<u32> src(E0), dst(E0);
    std::vector<u64> wt(E0);
    {
        std::vector<unsigned char> seen(N0, 0);
        if (gr->row_offsets) {
            for (u32 u = 0; u < N0; ++u)
                for (u64 e = gr->row_offsets[u]; e < gr->row_offsets[u + 1]; ++e) src[e] = u;
        } else {
            if (E0 && !gr->src) {
                rbsc_set_error("graph needs src or row_offsets");
                return RBSC_ERR_INVALID;
            }
            for (u64 e = 0; e < E0; ++e) src[e] = gr->src[e];
        }
        for (u64 e = 0; e < E0; ++e) {
            if (src[e] >= N0 || gr->dst[e] >= N0) {
                rbsc_set_error("line %llu: node id >= N=%u", (unsigned long long)e, N0);
                return RBSC_ERR_INPUT;
            }
            if (e == 0 || src[e] != src[e - 1]) {
                if (seen[src[e]]) {
                    rbsc_set_error("line %llu: lines of source %u are not grouped",
                                   (unsigned long long)e, src[e]);
                    return RBSC_ERR_INPUT;
                }
                seen[src[e]] = 1;
            }
            dst[e] = gr->dst[e];
            wt[e] = gr->weight ? gr->weight[e] : 1ULL;
        }
    }
    auto t0 = std::chrono::steady_clock::now();
    std::vector<u32> w(N0), label_map(N0 + 2), comp(N0);
    std::vector<unsigned char> mask(N0);
    std::vector<u64> considered(N0), windex(N0), sig_fill;
    std::vector<u32> c0, colors;
    for (u32 i = 0; i < N0; ++i) comp[i] = i;
    u32 node_n, new_node_n = N0;
    u64 edge_n, new_edge_n = E0, batches = 1;
    u32 passes = 0, rounds = 0;
    double dev_ms = 0;
    const u64 NOT_INIT = ~0ULL;
    u64* d_fill = nullptr;
    u64 d_fill_n = 0;
    int rc = RBSC_OK;
    for (;;) {
        node_n = new_node_n;
        edge_n = new_edge_n;
        new_node_n = 0;
        batches = (u64)ceil((double)edge_n / (double)max_batch);
        ++passes;
        for (u32 i = 0; i < node_n; ++i) w[i] = node_n;
        for (u64 batch = 0; batch < batches; ++batch) {
            u64 bs = batch * max_batch;
            u64 be = std::min<u64>(bs + max_batch, edge_n);
            for (u32 i = 0; i < node_n; ++i) mask[i] = batch == 0 ? MB_DANG : MB_OFF;
            u64 nbs = NOT_INIT, nbe = NOT_INIT;
            for (u64 i = bs; i < be; ++i) {
                if ((!i || src[i - 1] != src[i]) && nbs == NOT_INIT) nbs = i;
                if (i && nbs != NOT_INIT && src[i - 1] != src[i]) nbe = i;
                mask[src[i]] = MB_IN;
            }
            if (be == edge_n || src[be] != src[be - 1]) nbe = be;
            bs = nbs;
            be = nbe;
            if (bs == NOT_INIT || be == NOT_INIT) continue;
            for (u64 i = 0; i < bs; ++i) mask[src[i]] = MB_OFF;
            for (u64 i = be; i < edge_n; ++i) mask[src[i]] = MB_OFF;
            u32 batch_node_n = 0;
            for (u32 i = 0; i < node_n; ++i) batch_node_n += mask[i] != MB_OFF;
            const u32 off_n = node_n - batch_node_n;
            c0.resize(node_n);
            sig_fill.resize(node_n);
            for (u32 i = 0; i < node_n; ++i) {
                c0[i] = mask[i] == MB_IN ? node_n + 1 : (mask[i] == MB_DANG ? 0 : i + 1);
                sig_fill[i] = mask[i] == MB_OFF ? ~0ULL : 0ULL;
            }
            rbsc_graph bg{};
            bg.n_nodes = node_n;
            bg.n_edges = be - bs;
            bg.src = src.data() + bs;
            bg.dst = dst.data() + bs;
            bg.weight = wt.data() + bs;
            bg.init_color = c0.data();
            rbsc_dgraph* g = nullptr;
            rc = graph_upload_host(ctx, &bg, false, &g);
            if (rc) goto out;
            if (d_fill_n < node_n) {
                if (d_fill) (void)hipFree(d_fill);
                d_fill = nullptr;
                if (hipMalloc(&d_fill, sizeof(u64) * std::max<u32>(node_n, 1)) != hipSuccess) {
                    rbsc_set_error("hipMalloc failed for %u signatures", node_n);
                    rbsc_graph_free(ctx, g);
                    rc = RBSC_ERR_NOMEM;
                    goto out;
                }
                d_fill_n = node_n;
            }
            if (hipMemcpyAsync(d_fill, sig_fill.data(), sizeof(u64) * node_n, hipMemcpyHostToDevice,
                               ctx->stream) != hipSuccess) {
                rbsc_set_error("signature upload failed");
                rbsc_graph_free(ctx, g);
                rc = RBSC_ERR_HIP;
                goto out;
            }
            {
                RefineParams prm;
                prm.d_sig_fill = d_fill;
                prm.k0 = off_n + 1;
                RefineOut o;
                rc = refine_core(ctx, g, prm, &o);
                if (!rc) {
                    colors.resize(node_n);
                    rc = download_colors(ctx, g, colors.data());
                }
                rbsc_graph_free(ctx, g);
                if (rc) goto out;
                rounds += o.rounds;
                dev_ms += o.device_ms;
            }
            for (u32 i = 0; i < node_n; ++i) label_map[i] = node_n;
            for (u32 i = 0; i < node_n; ++i) {
                if (mask[i] != MB_OFF) {
                    const u32 local = colors[i];
                    if (label_map[local] == node_n) label_map[local] = new_node_n++;
                    w[i] = label_map[local];
                }
            }
        }
        if (batches > 1) {
            new_edge_n = 0;
            u64 first_edge_i = 0;
            u32 current_node = 0;
            bool counting = false;
            for (u32 i = 0; i < node_n; ++i) {
                windex[i] = edge_n;
                considered[i] = 0;
                if (w[i] == node_n) w[i] = new_node_n++;
            }
            for (u64 i = 0; i < edge_n; ++i) {
                if (!i || src[i] != current_node) {
                    current_node = src[i];
                    counting = !considered[w[current_node]];
                    first_edge_i = new_edge_n;
                    considered[w[src[i]]] = 1;
                }
                if (counting) {
                    src[i] = w[src[i]];
                    dst[i] = w[dst[i]];
                    if (windex[dst[i]] == edge_n || windex[dst[i]] < first_edge_i) {
                        src[new_edge_n] = src[i];
                        dst[new_edge_n] = dst[i];
                        wt[new_edge_n] = wt[i];
                        windex[dst[i]] = new_edge_n;
                        ++new_edge_n;
                    } else {
                        wt[windex[dst[i]]] += wt[i];
                    }
                }
            }
        }
        // Composition of labels for the partition of the original nodes.
        for (u32 i = 0; i < N0; ++i) comp[i] = comp[i] < node_n ? w[comp[i]] : comp[i];
        if (!(batches > 1 && node_n > new_node_n)) break;
        if (aggressive && !((float)batches > ceilf((float)new_edge_n / (float)max_batch))) break;
    }
    {
        auto t1 = std::chrono::steady_clock::now();
        rbsc_result r{};
        r.nodes = new_node_n;
        r.done = batches == 1;
        r.rounds = rounds;
        r.passes = passes;
        r.device_ms = dev_ms;
        r.edges_per_round = E0;
        r.time_s = std::chrono::duration<double>(t1 - t0).count();
        if (partition_out && N0) canonicalize_host(N0, comp.data(), partition_out);
        if (res) *res = r;
    }
out:
    if (d_fill) (void)hipFree(d_fill);
    return rc;
}
