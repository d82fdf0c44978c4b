/*
 * rbsc_gen.h -- seeded synthetic workloads for the RBSC library (host side).
 *
 * No counterpart in the reference, which reads LAW datasets through
 * webgraph-rs (bench.py:8, 29-30); these generators make the benchmark
 * workloads offline.  Every generator is deterministic in its seed.
 * Output graphs are CSR over node ids with lines sorted by (src, dst), the
 * order the reference's producer emits (webgraph-rs cli/src/to/arcscuda.rs).
 */
#ifndef RBSC_GEN_H
#define RBSC_GEN_H
#include <stdint.h>
#include "rbsc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A host graph allocated by the library; free with rbsc_host_graph_free. */
typedef struct rbsc_host_graph {
    uint32_t n_nodes;
    uint64_t n_edges;
    uint64_t* row_offsets;  /* n_nodes + 1 */
    uint32_t* dst;          /* n_edges */
    uint64_t* weight;       /* n_edges or NULL */
    uint32_t* init_color;   /* n_nodes or NULL */
} rbsc_host_graph;

void rbsc_host_graph_free(rbsc_host_graph* g);

/* R-MAT: 2^scale nodes, n_draws lines, quadrant probabilities a, b, c
 * (d = 1-a-b-c), seeded node-id bijection, clone_frac of the nodes replace
 * their lines with a copy of a random non-clone's lines.  Duplicates and
 * self-loops are kept (multiplicities).  Same graph as rbsc_gen_rmat_device. */
int rbsc_gen_rmat(uint32_t scale, uint64_t n_draws, double a, double b, double c,
                  double clone_frac, uint64_t seed, rbsc_host_graph* out);

/* Barabasi-Albert in-degree growth (directed): node t >= m+1 emits m lines
 * t -> v with v drawn proportional to in-degree(v)+1 among earlier nodes;
 * nodes 0..m form a complete seed digraph.  dup_frac of the lines get 1-3
 * extra copies (multiplicity 2-4).  n_init_colors > 0 draws initial colours
 * uniformly from [0, n_init_colors). */
int rbsc_gen_ba(uint32_t n, uint32_t m, double dup_frac, uint32_t n_init_colors, uint64_t seed,
                rbsc_host_graph* out);

/* Undirected graphs, each edge emitted as both directed lines:
 * Erdos-Renyi with the given average degree, Barabasi-Albert with m edges
 * per new node, and random k-regular (pairing model with rejection). */
int rbsc_gen_er_undirected(uint32_t n, double avg_deg, uint64_t seed, rbsc_host_graph* out);
int rbsc_gen_ba_undirected(uint32_t n, uint32_t m, uint64_t seed, rbsc_host_graph* out);
int rbsc_gen_regular_undirected(uint32_t n, uint32_t k, uint64_t seed, rbsc_host_graph* out);

/* Writes g as text: RBSC_TEXT_LINES is the reference's stdin format
 * (rbsc.cu:168-180: "N\nE\n" then "src weight dst" lines); RBSC_TEXT_VALMARI
 * is the 1-based MDP format the reference's valmari-adapter feeds part_ref
 * (valmari-adapter:9-16, MDP_read.cc:51-64), with the initial colouring as
 * initial blocks (MDP_read.cc:140-150). */
#define RBSC_TEXT_LINES 0
#define RBSC_TEXT_VALMARI 1
int rbsc_host_graph_write_text(const rbsc_host_graph* g, const char* path, int format);

/* A host batch allocated by the library (see rbsc_batch in rbsc.h). */
typedef struct rbsc_host_batch {
    uint32_t n_graphs;
    uint64_t* node_offsets;  /* n_graphs + 1 */
    uint64_t* line_offsets;  /* n_graphs + 1 */
    uint32_t* src;           /* local ids */
    uint32_t* dst;           /* local ids */
} rbsc_host_batch;
void rbsc_host_batch_free(rbsc_host_batch* b);

/* n_graphs undirected graphs, node counts log-uniform in [n_min, n_max];
 * graph i is Erdos-Renyi (average degree uniform in [2, 8]) when i % 3 == 0,
 * Barabasi-Albert (m uniform in {1..4}) when i % 3 == 1, random k-regular
 * (k in {3, 4}) when i % 3 == 2.  Each from its own seed. */
int rbsc_gen_batch_mixed(uint32_t n_graphs, uint32_t n_min, uint32_t n_max, uint64_t seed,
                         rbsc_host_batch* out);

#ifdef __cplusplus
}
#endif
#endif
