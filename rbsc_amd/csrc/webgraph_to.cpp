// webgraph_to.cpp -- `webgraph-to FORMAT BASENAME`: prints a BVGraph as the
// text the reference's harness pipes into its tools, in place of
// `webgraph to FORMAT NAME/NAME.graph` (bench.py:8, benchlarge.py:10):
//   arcs-cuda     "N\nE\n" then "src 1 dst" lines   (cli/src/to/arcscuda.rs:61-66)
//   arcs-valmari  "N 0 0 E 1" then 1-based lines    (cli/src/to/arcsvalmari.rs:61-65)
//   gdata         "N\nE\n"                           (cli/src/to/gdata.rs:61-62)
#include "../../include/rbsc_bvgraph.h"
#include "../../include/rbsc_gen.h"
#include <cstdio>
#include <cstring>

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: webgraph-to arcs-cuda|arcs-valmari|gdata BASENAME\n");
        return 2;
    }
    const char* fmt = argv[1];
    if (strcmp(fmt, "arcs-cuda") && strcmp(fmt, "arcs-valmari") && strcmp(fmt, "gdata")) {
        fprintf(stderr, "webgraph-to: unknown format %s\n", fmt);
        return 2;
    }
    rbsc_host_graph g;
    if (rbsc_bvgraph_load(argv[2], &g) != 0) {
        fprintf(stderr, "webgraph-to: %s\n", rbsc_last_error());
        return 1;
    }
    int rc = 0;
    if (!strcmp(fmt, "gdata")) {
        printf("%u\n%llu\n", g.n_nodes, (unsigned long long)g.n_edges);
    } else {
        fflush(stdout);
        rc = rbsc_host_graph_write_text(&g, "/dev/stdout",
                                        strcmp(fmt, "arcs-cuda") ? RBSC_TEXT_VALMARI : RBSC_TEXT_LINES);
        if (rc) fprintf(stderr, "webgraph-to: %s\n", rbsc_last_error());
    }
    rbsc_host_graph_free(&g);
    return rc ? 1 : 0;
}
