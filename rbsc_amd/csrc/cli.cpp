// cli.cpp -- the `rbsc` command: drop-in for the reference binary.
//
//   producer | rbsc [max_batch_edges] [gpus] [--flags...]
//
// Reference contract (src/rbsc/rbsc.cu):
//   argv[1] = max batch edge count, atoll, default E            (rbsc.cu:204)
//   argv[2] = number of GPUs, read when present, default 1       (rbsc.cu:205)
//   stdin   = "N\nE\n" then E lines "src weight dst"              (rbsc.cu:156-180)
//   stdout  = "%lf\n%u\n%lu\n": time, done, nodes                 (rbsc.cu:492-494)
// Diagnostics go to stderr only, so a harness reading the first three stdout
// lines (bench.py:44-48) sees either a result or nothing parseable.
//
// Extra options (after the positionals):
//   --mode quotient|aggressive|stream  edge-batch semantics when B < E:
//       quotient (default) is the reference's multi-batch contraction
//       (rbsc.cu:274-487: `done` may be 0), aggressive its rbsc_aggressive.cu
//       stop rule, stream the exact refinement with lines streamed from
//       pinned host memory in batches of B (done = 1)
//   --partition-out FILE   write the canonical partition, one label per line
//   --init-colors FILE     N initial colours, whitespace separated
//   --verify               exact stability check of the final partition
//   --device K             HIP device index
#include "../../include/rbsc.h"
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Input {
    std::vector<char> buf;
    size_t pos = 0;
    bool read_all(FILE* f) {
        buf.clear();
        const size_t chunk = 1 << 24;
        for (;;) {
            const size_t old = buf.size();
            buf.resize(old + chunk);
            const size_t got = fread(buf.data() + old, 1, chunk, f);
            buf.resize(old + got);
            if (got < chunk) break;
        }
        buf.push_back('\0');
        return !ferror(f);
    }
    // Next unsigned integer; false at end of input or on a malformed token.
    bool next(uint64_t* out) {
        const char* p = buf.data() + pos;
        while (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t') ++p;
        if (*p < '0' || *p > '9') return false;
        uint64_t v = 0;
        while (*p >= '0' && *p <= '9') {
            const uint64_t d = (uint64_t)(*p - '0');
            if (v > (UINT64_MAX - d) / 10) return false;
            v = v * 10 + d;
            ++p;
        }
        pos = p - buf.data();
        *out = v;
        return true;
    }
};

int fail(const char* what) {
    fprintf(stderr, "rbsc: %s\n", what);
    return EXIT_FAILURE;
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> pos;
    std::string mode = "quotient", part_out, init_file;
    int verify = 0, device = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto need = [&](const char* flag) -> const char* {
            if (i + 1 >= argc) {
                fprintf(stderr, "rbsc: %s needs a value\n", flag);
                exit(EXIT_FAILURE);
            }
            return argv[++i];
        };
        if (a == "--mode") mode = need("--mode");
        else if (a == "--partition-out") part_out = need("--partition-out");
        else if (a == "--init-colors") init_file = need("--init-colors");
        else if (a == "--verify") verify = 1;
        else if (a == "--device") device = atoi(need("--device"));
        else if (a.rfind("--", 0) == 0) {
            fprintf(stderr, "rbsc: unknown option %s\n", a.c_str());
            return EXIT_FAILURE;
        } else pos.push_back(a);
    }

    Input in;
    if (!in.read_all(stdin)) return fail("error reading stdin");
    uint64_t N64 = 0, E = 0;
    if (!in.next(&N64) || !in.next(&E)) return fail("input must start with N and E");
    if (N64 >= 0xFFFFFFFFull) return fail("N must be below 2^32-1");
    const uint32_t N = (uint32_t)N64;
    std::vector<uint32_t> src(E), dst(E);
    std::vector<uint64_t> wt(E);
    bool unit = true;
    for (uint64_t e = 0; e < E; ++e) {
        uint64_t u, w, v;
        if (!in.next(&u) || !in.next(&w) || !in.next(&v)) return fail("truncated or malformed line");
        if (u >= N || v >= N) return fail("node id out of range");
        src[e] = (uint32_t)u;
        dst[e] = (uint32_t)v;
        wt[e] = w;
        unit = unit && w == 1;
    }
    std::vector<uint32_t> init;
    if (!init_file.empty()) {
        FILE* f = fopen(init_file.c_str(), "rb");
        if (!f) return fail("cannot open --init-colors file");
        Input ic;
        ic.read_all(f);
        fclose(f);
        init.resize(N);
        for (uint32_t i = 0; i < N; ++i) {
            uint64_t c;
            if (!ic.next(&c) || c > 0xFFFFFFFEull) return fail("bad --init-colors file");
            init[i] = (uint32_t)c;
        }
    }
    const uint64_t max_batch = pos.size() >= 1 ? (uint64_t)atoll(pos[0].c_str()) : E;
    const long long gpus = pos.size() >= 2 ? atoll(pos[1].c_str()) : 1;
    if (gpus > 1) fprintf(stderr, "rbsc: %lld GPUs requested; this build runs on one device\n", gpus);

    rbsc_ctx* ctx = nullptr;
    if (rbsc_ctx_create(device, 0, &ctx) != RBSC_OK) {
        fprintf(stderr, "rbsc: %s\n", rbsc_last_error());
        return EXIT_FAILURE;
    }
    rbsc_graph g{};
    g.n_nodes = N;
    g.n_edges = E;
    g.src = src.data();
    g.dst = dst.data();
    g.weight = unit ? nullptr : wt.data();
    g.init_color = init.empty() ? nullptr : init.data();
    rbsc_opts o{};
    o.max_batch_edges = max_batch;
    o.mode = mode == "quotient" ? RBSC_MODE_QUOTIENT
           : mode == "aggressive" ? RBSC_MODE_QUOTIENT_AGGRESSIVE : RBSC_MODE_STREAM;
    o.verify = (uint32_t)verify;
    rbsc_result r{};
    // the partition lands in page-locked memory (a copy at the link's rate)
    uint32_t* part = nullptr;
    if (!part_out.empty() && rbsc_host_alloc(sizeof(uint32_t) * (uint64_t)N, (void**)&part) != RBSC_OK) {
        fprintf(stderr, "rbsc: %s\n", rbsc_last_error());
        rbsc_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }
    const int rc = rbsc_refine(ctx, &g, &o, &r, part);
    if (rc != RBSC_OK) {
        fprintf(stderr, "rbsc: %s (status %d)\n", rbsc_last_error(), rc);
        rbsc_host_free(part);
        rbsc_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }
    rbsc_ctx_destroy(ctx);
    printf("%lf\n", r.time_s);
    printf("%u\n", r.done);
    printf("%lu\n", (unsigned long)r.nodes);
    fflush(stdout);
    fprintf(stderr, "rbsc: rounds=%u passes=%u device_ms=%.3f verified=%d\n", r.rounds, r.passes,
            r.device_ms, r.verified);
    if (!part_out.empty()) {
        FILE* f = fopen(part_out.c_str(), "wb");
        if (!f) {
            rbsc_host_free(part);
            return fail("cannot open --partition-out file");
        }
        for (uint32_t i = 0; i < N; ++i) fprintf(f, "%u\n", part[i]);
        fclose(f);
    }
    rbsc_host_free(part);
    return 0;
}
