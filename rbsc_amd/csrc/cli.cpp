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
//   (a pipe on stdin is parsed in chunks of 256 MiB, or of RBSC_READ_CHUNK
//   bytes when that is set in the environment)
//   --check-input          parse and check stdin only (no device): print
//       "N E line-digest csr|coo unit|weighted offsets-digest" and the parse
//       time on stderr; the digests are sums mod 2^64 of
//       src*A + w*B + dst*C + e*D over lines e and of ro[u]*(u+1) over the
//       row offsets (0 when the lines go over as COO)
#include "../../include/rbsc.h"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <sys/mman.h>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>
#include <vector>

namespace {

// The whole of a file descriptor's bytes: mapped when it is a regular file
// (stdin redirected from a file: no copy, pages fault in on the parsing
// threads), else read into an anonymous mapping grown by mremap (a pipe;
// growth moves page tables, not bytes).
struct Text {
    char* p = nullptr;
    size_t n = 0, cap = 0;
    bool mapped_file = false;
    bool load(int fd) {
        struct stat st;
        if (fstat(fd, &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0) {
            const off_t at = lseek(fd, 0, SEEK_CUR);
            if (at >= 0 && at < st.st_size) {
                void* m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
                if (m != MAP_FAILED) {
                    p = (char*)m + at;
                    n = (size_t)(st.st_size - at);
                    cap = (size_t)st.st_size;
                    mapped_file = true;
                    return true;
                }
            }
        }
        cap = (size_t)1 << 30;
        void* m = mmap(nullptr, cap, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        if (m == MAP_FAILED) return false;
        p = (char*)m;
        for (;;) {
            if (n == cap) {
                void* g = mremap(p, cap, cap * 2, MREMAP_MAYMOVE);
                if (g == MAP_FAILED) return false;
                p = (char*)g;
                cap *= 2;
            }
            const ssize_t got = read(fd, p + n, cap - n);
            if (got < 0) {
                if (errno == EINTR) continue;
                return false;
            }
            if (got == 0) return true;
            n += (size_t)got;
        }
    }
    ~Text() {
        if (!p) return;
        if (mapped_file) munmap(p - (cap - n), cap);
        else munmap(p, cap);
    }
};

inline bool is_digit(char c) { return (unsigned char)(c - '0') < 10; }
inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t'; }

// Next unsigned integer at or after *pos in [0, n); false at the end, on a
// byte that is neither a digit nor white space, or on overflow.
bool next_u64(const char* p, size_t n, size_t* pos, uint64_t* out) {
    size_t i = *pos;
    while (i < n && is_space(p[i])) ++i;
    if (i >= n || !is_digit(p[i])) return false;
    uint64_t v = 0;
    while (i < n && is_digit(p[i])) {
        const uint64_t d = (uint64_t)(p[i] - '0');
        if (v > (UINT64_MAX - d) / 10) return false;
        v = v * 10 + d;
        ++i;
    }
    *pos = i;
    *out = v;
    return true;
}

// The E lines "src weight dst" after the header (rbsc.cu:156-180), fed as
// byte ranges that end at a token boundary.  Each range is parsed by T
// threads in two passes over T sub-ranges cut at token boundaries: the
// first counts each sub-range's tokens (and notes its first byte that is
// neither a digit nor white space), a prefix sum gives every sub-range its
// first token index, the second parses token t of the body into field t % 3
// of line t / 3.  Bytes past the 3E-th token are ignored, as a reader that
// stops after E lines ignores them.
// Host memory for what goes to the device: page-locked (rbsc_host_alloc)
// when `pinned`, so the upload inside the timed span runs at the link's
// pinned rate instead of through the runtime's staging copies; plain memory
// otherwise, or when page-locked memory cannot be had.  Left uninitialised.
class HostBuf {
  public:
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { reset(); }
    // zeroed: plain memory that reads as zeros (calloc: fresh pages, so
    // untouched parts cost nothing)
    void* alloc(size_t bytes, bool pinned, bool zeroed = false) {
        reset();
        if (pinned && rbsc_host_alloc(bytes ? bytes : 1, &p_) == RBSC_OK) {
            pinned_ = true;
            return p_;
        }
        p_ = zeroed ? calloc(bytes ? bytes : 1, 1) : malloc(bytes ? bytes : 1);
        pinned_ = false;
        if (!p_) {
            fprintf(stderr, "rbsc: out of host memory (%zu bytes)\n", bytes);
            exit(EXIT_FAILURE);
        }
        return p_;
    }
    void reset() {
        if (!p_) return;
        if (pinned_) rbsc_host_free(p_);
        else free(p_);
        p_ = nullptr;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p_); }

  private:
    void* p_ = nullptr;
    bool pinned_ = false;
};

struct Lines {
    bool pinned = false;  // heads and row offsets in page-locked memory
    HostBuf src_buf, dst_buf, wt_buf;
    uint32_t *src = nullptr, *dst = nullptr;
    uint64_t* wt = nullptr;
    bool unit = true;    // every weight is 1
    bool sorted = true;  // sources never decrease
};

class LineParser {
  public:
    LineParser(uint64_t N, uint64_t E, Lines* L) : N_(N), E_(E), L_(L) {
        T_ = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
        L->src = static_cast<uint32_t*>(L->src_buf.alloc(4 * E, false));
        L->dst = static_cast<uint32_t*>(L->dst_buf.alloc(4 * E, L->pinned));
        // weights other than 1 only, stored as w ^ 1 (so an untouched 0 is a
        // weight of 1), restored by finish() when there are any
        L->wt = static_cast<uint64_t*>(L->wt_buf.alloc(8 * E, false, true));
    }
    bool done() const { return t_ >= 3 * E_; }
    // Parses [p, p + n); nullptr, or the reason the text is refused.
    const char* feed(const char* p, size_t n) {
        if (done() || n == 0) return nullptr;
        const unsigned T = n < ((size_t)1 << 22) ? 1 : T_;
        std::vector<size_t> cut(T + 1);
        cut[0] = 0;
        cut[T] = n;
        for (unsigned k = 1; k < T; ++k) {
            size_t c = n / T * k;
            while (c < n && is_digit(p[c])) ++c;  // never inside a token
            cut[k] = std::max(c, cut[k - 1]);
        }
        std::vector<uint64_t> tok(T + 1, 0), bad_tok(T, UINT64_MAX);
        run(T, [&](unsigned k) { count(p + cut[k], cut[k + 1] - cut[k], &tok[k + 1], &bad_tok[k]); });
        tok[0] = t_;
        for (unsigned k = 0; k < T; ++k) {
            if (bad_tok[k] != UINT64_MAX && tok[k] + bad_tok[k] < 3 * E_)
                return "truncated or malformed line";
            tok[k + 1] += tok[k];
        }
        std::vector<unsigned char> unit(T, 1);
        run(T, [&](unsigned k) { unit[k] = parse(p + cut[k], cut[k + 1] - cut[k], tok[k]); });
        if (err_.load() & 1) return "truncated or malformed line";
        if (err_.load() & 2) return "node id out of range";
        for (unsigned k = 0; k < T; ++k) L_->unit = L_->unit && unit[k];
        t_ = tok[T];
        return nullptr;
    }
    // After the last range: enough tokens?  Then whether sources are sorted.
    const char* finish() {
        if (!done()) return "truncated or malformed line";
        std::vector<unsigned char> sorted(T_, 1);
        const uint64_t E = E_;
        const unsigned T = T_;
        run(T, [&](unsigned k) {
            const uint64_t a0 = std::max<uint64_t>(1, E / T * k);
            const uint64_t z = k + 1 == T ? E : E / T * (k + 1);
            for (uint64_t e = a0; e < z; ++e)
                if (L_->src[e] < L_->src[e - 1]) {
                    sorted[k] = 0;
                    return;
                }
        });
        for (unsigned k = 0; k < T; ++k) L_->sorted = L_->sorted && sorted[k];
        if (!L_->unit)
            run(T, [&](unsigned k) {
                const uint64_t a0 = E / T * k, z = k + 1 == T ? E : E / T * (k + 1);
                for (uint64_t e = a0; e < z; ++e) L_->wt[e] ^= 1;
            });
        return nullptr;
    }

  private:
    template <class F>
    static void run(unsigned T, F f) {
        if (T == 1) return f(0u);
        std::vector<std::thread> th;
        for (unsigned k = 0; k < T; ++k) th.emplace_back(f, k);
        for (auto& x : th) x.join();
    }
    // Tokens of q[0, len) (a range starts after a non-digit), and into *bad
    // the tokens read before its first stray byte (UINT64_MAX if none).
    static void count(const char* q, size_t len, uint64_t* out, uint64_t* bad_out) {
        uint64_t c = 0, bad = UINT64_MAX;
        bool prev = false;
        for (size_t i = 0; i < len; i += 4096) {
            const size_t stop = std::min(len, i + 4096);
            const uint64_t c0 = c;
            const bool prev0 = prev;
            unsigned stray = 0;
            for (size_t j = i; j < stop; ++j) {  // branch-free over the block
                const bool d = is_digit(q[j]);
                c += d & !prev;
                prev = d;
                stray |= !(d | is_space(q[j]));
            }
            if (stray && bad == UINT64_MAX) {
                uint64_t cb = c0;
                bool pv = prev0;
                for (size_t j = i; j < stop; ++j) {
                    const bool d = is_digit(q[j]);
                    if (!(d || is_space(q[j]))) break;
                    cb += d & !pv;
                    pv = d;
                }
                bad = cb;
            }
        }
        *out = c;
        *bad_out = bad;
    }
    // Tokens of q[0, len) from global token index t on; returns whether
    // every weight seen was 1.
    bool parse(const char* q, size_t len, uint64_t t) {
        bool u1 = true;
        size_t i = 0;
        const uint64_t need = 3 * E_;
        while (t < need) {
            while (i < len && !is_digit(q[i])) ++i;
            if (i >= len) break;
            uint64_t v = 0;
            while (i < len && is_digit(q[i])) {
                const uint64_t d = (uint64_t)(q[i] - '0');
                if (v > (UINT64_MAX - d) / 10) {
                    err_.fetch_or(1);
                    return u1;
                }
                v = v * 10 + d;
                ++i;
            }
            const uint64_t e = t / 3;
            switch (t % 3) {
                case 0:
                    if (v >= N_) err_.fetch_or(2);
                    L_->src[e] = (uint32_t)v;
                    break;
                case 1:
                    if (v != 1) {
                        L_->wt[e] = v ^ 1;
                        u1 = false;
                    }
                    break;
                default:
                    if (v >= N_) err_.fetch_or(2);
                    L_->dst[e] = (uint32_t)v;
            }
            ++t;
        }
        return u1;
    }
    uint64_t N_, E_, t_ = 0;
    unsigned T_ = 1;
    Lines* L_;
    std::atomic<int> err_{0};  // 1 overflow, 2 id out of range
};

// Fills buf[0, want) from fd as far as the input goes; bytes read, or -1.
ssize_t read_full(int fd, char* buf, size_t want) {
    size_t got = 0;
    while (got < want) {
        const ssize_t r = read(fd, buf + got, want - got);
        if (r < 0) {
            if (errno == EINTR) continue;
            return -1;
        }
        if (r == 0) break;
        got += (size_t)r;
    }
    return (ssize_t)got;
}

// The header and lines of stdin.  A regular file is mapped and parsed in
// place (pages fault in on the parsing threads).  Anything else (a pipe) is
// read in chunks of CHUNK bytes by a reader thread while the previous chunk
// is parsed; a chunk ends after its last white space, the partial token
// after it moving to the front of the next chunk, so the text is never held
// whole.
const char* read_input(int fd, uint64_t* N_out, uint64_t* E_out, Lines* L, size_t* bytes) {
    struct stat st;
    if (fstat(fd, &st) == 0 && S_ISREG(st.st_mode)) {
        Text in;
        if (!in.load(fd)) return "error reading stdin";
        size_t at = 0;
        uint64_t N = 0, E = 0;
        if (!next_u64(in.p, in.n, &at, &N) || !next_u64(in.p, in.n, &at, &E))
            return "input must start with N and E";
        if (N >= 0xFFFFFFFFull) return "N must be below 2^32-1";
        *N_out = N;
        *E_out = E;
        *bytes = in.n;
        LineParser lp(N, E, L);
        if (const char* why = lp.feed(in.p + at, in.n - at)) return why;
        return lp.finish();
    }
    // RBSC_READ_CHUNK overrides the chunk size (bytes, at least 4096)
    size_t CHUNK = (size_t)256 << 20;
    if (const char* c = getenv("RBSC_READ_CHUNK")) CHUNK = std::max<size_t>(4096, strtoull(c, nullptr, 10));
    constexpr size_t SLACK = 1 << 16;  // room for a carried partial token
    std::unique_ptr<char[]> buf[2] = {std::unique_ptr<char[]>(new char[CHUNK + SLACK]),
                                      std::unique_ptr<char[]>(new char[CHUNK + SLACK])};
    size_t carry = 0;
    ssize_t got = read_full(fd, buf[0].get(), CHUNK);
    if (got < 0) return "error reading stdin";
    size_t len = (size_t)got;
    bool eof = len < CHUNK;
    *bytes = len;
    size_t at = 0;
    uint64_t N = 0, E = 0;
    if (!next_u64(buf[0].get(), len, &at, &N) || !next_u64(buf[0].get(), len, &at, &E))
        return "input must start with N and E";
    if (!eof && at == len) return "input must start with N and E";  // header cut by the chunk
    if (N >= 0xFFFFFFFFull) return "N must be below 2^32-1";
    *N_out = N;
    *E_out = E;
    LineParser lp(N, E, L);
    int cur = 0;
    size_t start = at;
    for (;;) {
        // the chunk ends after its last white space unless the input ended
        size_t end = len;
        if (!eof) {
            while (end > start && !is_space(buf[cur][end - 1])) --end;
            if (end == start) return "truncated or malformed line";  // a token longer than a chunk
        }
        carry = len - end;
        if (carry > SLACK) return "truncated or malformed line";
        std::memcpy(buf[cur ^ 1].get(), buf[cur].get() + end, carry);
        ssize_t next_got = 0;
        std::thread reader;
        const bool more = !eof && !lp.done();
        if (more)
            reader = std::thread([&, cur]() {
                next_got = read_full(fd, buf[cur ^ 1].get() + carry, CHUNK);
            });
        const char* why = lp.feed(buf[cur].get() + start, end - start);
        if (reader.joinable()) reader.join();
        if (why) return why;
        if (!more) break;
        if (next_got < 0) return "error reading stdin";
        cur ^= 1;
        start = 0;
        len = carry + (size_t)next_got;
        eof = (size_t)next_got < CHUNK;
        *bytes += (size_t)next_got;
    }
    return lp.finish();
}

int fail(const char* what) {
    fprintf(stderr, "rbsc: %s\n", what);
    return EXIT_FAILURE;
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> pos;
    std::string mode = "quotient", part_out, init_file;
    int verify = 0, device = 0, check_input = 0;
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
        else if (a == "--check-input") check_input = 1;
        else if (a == "--device") device = atoi(need("--device"));
        else if (a.rfind("--", 0) == 0) {
            fprintf(stderr, "rbsc: unknown option %s\n", a.c_str());
            return EXIT_FAILURE;
        } else pos.push_back(a);
    }

    const auto t_read = std::chrono::steady_clock::now();
    uint64_t N64 = 0, E = 0;
    size_t bytes = 0;
    Lines L;
    L.pinned = !check_input;  // --check-input touches no device
    if (const char* why = read_input(0, &N64, &E, &L, &bytes)) return fail(why);
    const uint32_t N = (uint32_t)N64;
    uint64_t line_digest = 0;
    if (check_input)
        for (uint64_t e = 0; e < E; ++e)
            line_digest += (uint64_t)L.src[e] * 0x9e3779b97f4a7c15ULL +
                           (L.unit ? 1 : L.wt[e]) * 0xbf58476d1ce4e5b9ULL +
                           (uint64_t)L.dst[e] * 0x94d049bb133111ebULL + e * 0x2545f4914f6cdd1dULL;
    // sources in order: hand the lines over as CSR (row offsets instead of
    // one source per line: 8 (N + 1) bytes to upload instead of 4 E)
    HostBuf ro_buf;
    uint64_t* row_offsets = nullptr;
    if (L.sorted && E) {
        row_offsets = static_cast<uint64_t*>(ro_buf.alloc(8 * ((size_t)N + 1), L.pinned));
        const unsigned T = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (unsigned k = 0; k < T; ++k)
            th.emplace_back([&, k]() {
                // line e starting a row (or e = E) ends the rows before it
                const uint64_t a = E / T * k, z = k + 1 == T ? E + 1 : E / T * (k + 1);
                for (uint64_t e = a; e < z; ++e) {
                    const uint64_t lo = e == 0 ? 0 : (uint64_t)L.src[e - 1] + 1;
                    const uint64_t hi = e == E ? (uint64_t)N : (uint64_t)L.src[e];
                    if (e == 0 || e == E || L.src[e] != L.src[e - 1])
                        for (uint64_t u = lo; u <= hi && u <= N; ++u) row_offsets[u] = e;
                }
            });
        for (auto& t : th) t.join();
        row_offsets[N] = E;
        L.src_buf.reset();
        L.src = nullptr;
    }
    const double parse_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_read).count();
    if (check_input) {
        uint64_t ro_digest = 0;
        if (row_offsets)
            for (size_t u = 0; u <= N; ++u) ro_digest += row_offsets[u] * (uint64_t)(u + 1);
        const double dt =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t_read).count();
        printf("%u %llu %016llx %s %s %016llx\n", N, (unsigned long long)E,
               (unsigned long long)line_digest, row_offsets ? "csr" : "coo",
               L.unit ? "unit" : "weighted", (unsigned long long)ro_digest);
        fprintf(stderr, "rbsc: read and parsed %zu bytes in %.3f s\n", bytes, dt);
        return 0;
    }
    std::vector<uint32_t> init;
    if (!init_file.empty()) {
        FILE* f = fopen(init_file.c_str(), "rb");
        if (!f) return fail("cannot open --init-colors file");
        Text ic;
        const bool ok = ic.load(fileno(f));
        fclose(f);
        if (!ok) return fail("cannot read --init-colors file");
        init.resize(N);
        size_t q = 0;
        for (uint32_t i = 0; i < N; ++i) {
            uint64_t c;
            if (!next_u64(ic.p, ic.n, &q, &c) || c > 0xFFFFFFFEull) return fail("bad --init-colors file");
            init[i] = (uint32_t)c;
        }
    }
    const uint64_t max_batch = pos.size() >= 1 ? (uint64_t)atoll(pos[0].c_str()) : E;
    const long long gpus = pos.size() >= 2 ? atoll(pos[1].c_str()) : 1;
    if (gpus > 1) fprintf(stderr, "rbsc: %lld GPUs requested; running on one device\n", gpus);

    rbsc_ctx* ctx = nullptr;
    if (rbsc_ctx_create(device, 0, &ctx) != RBSC_OK) {
        fprintf(stderr, "rbsc: %s\n", rbsc_last_error());
        return EXIT_FAILURE;
    }
    rbsc_graph g{};
    g.n_nodes = N;
    g.n_edges = E;
    g.row_offsets = row_offsets;
    g.src = row_offsets ? nullptr : L.src;
    g.dst = L.dst;
    g.weight = L.unit ? nullptr : L.wt;
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
    fprintf(stderr, "rbsc: read and parsed %zu bytes in %.3f s (%s, %s)\n", bytes, parse_s,
            row_offsets ? "csr" : "coo", L.unit ? "unit weights" : "weights");
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
