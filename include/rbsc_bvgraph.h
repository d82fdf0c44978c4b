/*
 * rbsc_bvgraph.h -- reading WebGraph BVGraph files (the LAW datasets the
 * reference benchmarks) without the webgraph-rs tool chain.
 *
 * The reference feeds its tools through `webgraph to arcs-cuda NAME` /
 * `arcs-valmari` / `gdata` (bench.py:8, benchlarge.py:10, webgraph-rs
 * cli/src/to/arcscuda.rs:61-66, arcsvalmari.rs:61-65, gdata.rs); these calls
 * replace that decoder: BVGraph sequential decoding as in webgraph-rs
 * webgraph/src/graphs/bvgraph/sequential.rs:213-313 with the instantaneous
 * codes of its .properties (compflags, zetak; defaults of
 * comp/flags.rs:43-54), big-endian bit order (the Java WebGraph layout).
 */
#ifndef RBSC_BVGRAPH_H
#define RBSC_BVGRAPH_H
#include <stdint.h>
#include "rbsc_gen.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Decodes BASENAME.graph (+ BASENAME.properties) into a host CSR graph with
 * successors sorted, the line order `webgraph to arcs-cuda` prints.  Free
 * with rbsc_host_graph_free. */
int rbsc_bvgraph_load(const char* basename, rbsc_host_graph* out);

/* Checks BASENAME.offsets (gamma-coded gaps of each node's first bit) against
 * the positions found while decoding BASENAME.graph: *n_offsets = offsets
 * read, *mismatches = nodes whose position differs (0 when consistent). */
int rbsc_bvgraph_check_offsets(const char* basename, uint64_t* n_offsets, uint64_t* mismatches);

#ifdef __cplusplus
}
#endif
#endif
