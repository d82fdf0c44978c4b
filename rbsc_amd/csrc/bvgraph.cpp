// bvgraph.cpp -- sequential decoding of WebGraph BVGraph files (host side).
//
// Layout (Boldi-Vigna compression; the field order and copy/interval/
// residual rules follow webgraph-rs webgraph/src/graphs/bvgraph/sequential.rs
// 213-313): per node x, its outdegree d; if d > 0 and the window is not 0,
// a reference offset r; for r > 0 a block count and blocks that alternately
// copy and skip the successors of node x - r (copy first; the first block
// may be 0, the others are stored minus one; an even block count copies the
// rest of the list); then, while successors remain and intervals are on, an
// interval count and intervals (first start x + zigzag(v), later starts
// previous end + 1 + v, lengths v + min_interval_length); then residuals
// (first x + zigzag(v), later previous + 1 + v).  Bits are read most
// significant first (the Java WebGraph layout).  Codes per field come from
// the .properties compflags (defaults: outdegrees gamma, references unary,
// blocks gamma, intervals gamma, residuals zeta_k with k = zetak, 3).
#include "../../include/rbsc_bvgraph.h"
#include "../../include/rbsc.h"
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

extern "C" void rbsc_set_error(const char* fmt, ...);

namespace {

enum Code { GAMMA, DELTA, UNARY, ZETA };

struct Props {
    uint64_t nodes = 0, arcs = 0;
    uint32_t window = 7, min_interval = 4, zeta_k = 3;
    Code outdeg = GAMMA, refs = UNARY, blocks = GAMMA, intervals = GAMMA, residuals = ZETA;
};

std::string strip_graph(const char* basename) {
    std::string b = basename;
    const std::string suf = ".graph";
    if (b.size() > suf.size() && b.compare(b.size() - suf.size(), suf.size(), suf) == 0)
        b.resize(b.size() - suf.size());
    return b;
}

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    out.assign((size_t)(n > 0 ? n : 0) + 16, 0);  // zero padding for the bit reader
    const size_t got = n > 0 ? fread(out.data(), 1, (size_t)n, f) : 0;
    fclose(f);
    return got == (size_t)(n > 0 ? n : 0);
}

int parse_code(const std::string& name, Code* c) {
    if (name == "GAMMA") *c = GAMMA;
    else if (name == "DELTA") *c = DELTA;
    else if (name == "UNARY") *c = UNARY;
    else if (name == "ZETA") *c = ZETA;
    else return -1;
    return 0;
}

int read_props(const std::string& base, Props* p) {
    std::ifstream in(base + ".properties");
    if (!in) {
        rbsc_set_error("cannot read %s.properties", base.c_str());
        return RBSC_ERR_INPUT;
    }
    std::map<std::string, std::string> kv;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string v = line.substr(eq + 1);
        while (!v.empty() && (v.back() == '\r' || v.back() == ' ')) v.pop_back();
        kv[line.substr(0, eq)] = v;
    }
    auto num = [&](const char* k, uint64_t dflt) -> uint64_t {
        auto it = kv.find(k);
        return it == kv.end() || it->second.empty() ? dflt : strtoull(it->second.c_str(), nullptr, 10);
    };
    if (!kv.count("nodes") || !kv.count("arcs")) {
        rbsc_set_error("%s.properties lacks nodes or arcs", base.c_str());
        return RBSC_ERR_INPUT;
    }
    p->nodes = num("nodes", 0);
    p->arcs = num("arcs", 0);
    p->window = (uint32_t)num("windowsize", 7);
    p->min_interval = (uint32_t)num("minintervallength", 4);
    p->zeta_k = (uint32_t)num("zetak", 3);
    if (p->nodes >= 0xFFFFFFFFull || p->zeta_k == 0 || p->zeta_k > 16) {
        rbsc_set_error("%s.properties: unsupported nodes or zetak", base.c_str());
        return RBSC_ERR_INPUT;
    }
    std::stringstream flags(kv["compressionflags"]);
    std::string f;
    while (std::getline(flags, f, '|')) {
        if (f.empty()) continue;
        const size_t us = f.find('_');
        if (us == std::string::npos) continue;
        const std::string what = f.substr(0, us), code = f.substr(us + 1);
        Code c;
        if (what == "OFFSETS" || what == "BLOCK" || what == "BLOCK_COUNT") continue;
        if (parse_code(code, &c)) {
            rbsc_set_error("compression flag %s: code %s not supported", f.c_str(), code.c_str());
            return RBSC_ERR_INPUT;
        }
        if (what == "OUTDEGREES") p->outdeg = c;
        else if (what == "REFERENCES") p->refs = c;
        else if (what == "BLOCKS") p->blocks = c;
        else if (what == "INTERVALS") p->intervals = c;
        else if (what == "RESIDUALS") p->residuals = c;
        else {
            rbsc_set_error("unknown compression flag %s", f.c_str());
            return RBSC_ERR_INPUT;
        }
    }
    return RBSC_OK;
}

// Most-significant-bit-first reader over a zero-padded byte buffer.
struct Bits {
    const uint8_t* d;
    uint64_t nbits;
    uint64_t p = 0;
    bool over = false;  // read past the end of the data
    uint64_t peek64() const {
        const uint64_t b = p >> 3;
        uint64_t v = 0;
        for (int i = 0; i < 8; ++i) v = (v << 8) | d[b + i];
        const unsigned s = (unsigned)(p & 7);
        if (s) v = (v << s) | (uint64_t)(d[b + 8] >> (8 - s));
        return v;
    }
    uint64_t read(unsigned n) {  // n <= 64
        if (n == 0) return 0;
        const uint64_t v = peek64() >> (64 - n);
        p += n;
        if (p > nbits) over = true;
        return v;
    }
    uint64_t unary() {
        uint64_t zeros = 0;
        for (;;) {
            if (p >= nbits) {
                over = true;
                return 0;
            }
            const uint64_t w = peek64();
            if (w) {
                const unsigned lz = (unsigned)__builtin_clzll(w);
                zeros += lz;
                p += lz + 1;
                if (p > nbits) over = true;
                return zeros;
            }
            zeros += 64;
            p += 64;
        }
    }
    uint64_t gamma() {
        const uint64_t msb = unary();
        if (msb >= 64) {
            over = true;
            return 0;
        }
        return ((1ULL << msb) | read((unsigned)msb)) - 1;
    }
    uint64_t delta() {
        const uint64_t msb = gamma();
        if (msb >= 64) {
            over = true;
            return 0;
        }
        return ((1ULL << msb) | read((unsigned)msb)) - 1;
    }
    // zeta_k: h in unary, then the minimal binary code of x + 1 - 2^(hk) in
    // an interval of 2^((h+1)k) - 2^(hk) values.
    uint64_t zeta(unsigned k) {
        const uint64_t h = unary();
        if (h * k + k > 63) {
            over = true;
            return 0;
        }
        const uint64_t left = 1ULL << (h * k);
        const uint64_t m = read((unsigned)(h * k + k - 1));
        if (m < left) return m + left - 1;
        return (m << 1) + read(1) - 1;
    }
    uint64_t code(Code c, unsigned k) {
        switch (c) {
            case GAMMA: return gamma();
            case DELTA: return delta();
            case UNARY: return unary();
            default: return zeta(k);
        }
    }
};

inline int64_t zigzag(uint64_t v) { return (v & 1) ? -(int64_t)((v + 1) >> 1) : (int64_t)(v >> 1); }

// Decodes every node; succ receives the sorted lists in CSR form and
// starts[x] the first bit of node x (starts[n] = bits consumed).
int decode(const std::string& base, const Props& pr, std::vector<uint64_t>& row,
           std::vector<uint32_t>& succ, std::vector<uint64_t>* starts) {
    std::vector<uint8_t> data;
    if (!read_file(base + ".graph", data)) {
        rbsc_set_error("cannot read %s.graph", base.c_str());
        return RBSC_ERR_INPUT;
    }
    Bits bs{data.data(), (uint64_t)(data.size() - 16) * 8};
    const uint64_t n = pr.nodes;
    const uint32_t W = pr.window;
    std::vector<std::vector<uint32_t>> win(W + 1);  // lists of the last W + 1 nodes
    row.assign(n + 1, 0);
    succ.clear();
    succ.reserve(pr.arcs);
    if (starts) starts->assign(n + 1, 0);
    std::vector<uint32_t> res;
    auto bad = [&](uint64_t x, const char* what) {
        rbsc_set_error("%s.graph: node %llu: %s", base.c_str(), (unsigned long long)x, what);
        return RBSC_ERR_INPUT;
    };
    for (uint64_t x = 0; x < n; ++x) {
        if (starts) (*starts)[x] = bs.p;
        res.clear();
        const uint64_t d = bs.code(pr.outdeg, pr.zeta_k);
        if (d > n) return bad(x, "outdegree larger than the node count");
        if (d) {
            const uint64_t r = W ? bs.code(pr.refs, pr.zeta_k) : 0;
            if (r > W || r > x) return bad(x, "reference outside the window");
            if (r) {
                const std::vector<uint32_t>& nb = win[(x - r) % (W + 1)];
                const uint64_t nblocks = bs.code(pr.blocks, pr.zeta_k);
                if (nblocks == 0) {
                    res.insert(res.end(), nb.begin(), nb.end());
                } else {
                    uint64_t idx = bs.code(pr.blocks, pr.zeta_k);
                    if (idx > nb.size()) return bad(x, "block past the reference list");
                    res.insert(res.end(), nb.begin(), nb.begin() + idx);
                    for (uint64_t b = 1; b < nblocks; ++b) {
                        const uint64_t end = idx + bs.code(pr.blocks, pr.zeta_k) + 1;
                        if (end > nb.size()) return bad(x, "block past the reference list");
                        if (b % 2 == 0) res.insert(res.end(), nb.begin() + idx, nb.begin() + end);
                        idx = end;
                    }
                    if (nblocks % 2 == 0) res.insert(res.end(), nb.begin() + idx, nb.end());
                }
                if (res.size() > d) return bad(x, "more copied successors than the outdegree");
            }
            if (res.size() < d && pr.min_interval) {
                const uint64_t nint = bs.code(pr.intervals, pr.zeta_k);
                int64_t start = 0;
                for (uint64_t i = 0; i < nint; ++i) {
                    const uint64_t v = bs.code(pr.intervals, pr.zeta_k);
                    start = i == 0 ? (int64_t)x + zigzag(v) : start + 1 + (int64_t)v;
                    const uint64_t len = bs.code(pr.intervals, pr.zeta_k) + pr.min_interval;
                    if (start < 0 || (uint64_t)start + len > n || res.size() + len > d)
                        return bad(x, "interval out of range");
                    for (uint64_t j = 0; j < len; ++j) res.push_back((uint32_t)(start + (int64_t)j));
                    start += (int64_t)len;
                }
            }
            if (res.size() < d) {
                const uint64_t left = d - res.size();
                int64_t extra = (int64_t)x + zigzag(bs.code(pr.residuals, pr.zeta_k));
                for (uint64_t i = 0; i < left; ++i) {
                    if (i) extra += 1 + (int64_t)bs.code(pr.residuals, pr.zeta_k);
                    if (extra < 0 || (uint64_t)extra >= n) return bad(x, "residual out of range");
                    res.push_back((uint32_t)extra);
                }
            }
            std::sort(res.begin(), res.end());
        }
        if (bs.over) return bad(x, "truncated bit stream");
        win[x % (W + 1)] = res;
        succ.insert(succ.end(), res.begin(), res.end());
        row[x + 1] = succ.size();
    }
    if (starts) (*starts)[n] = bs.p;
    if (succ.size() != pr.arcs) {
        rbsc_set_error("%s: decoded %llu arcs, properties say %llu", base.c_str(),
                       (unsigned long long)succ.size(), (unsigned long long)pr.arcs);
        return RBSC_ERR_INPUT;
    }
    return RBSC_OK;
}

}  // namespace

extern "C" int rbsc_bvgraph_load(const char* basename, rbsc_host_graph* out) {
    if (!basename || !out) {
        rbsc_set_error("rbsc_bvgraph_load: null argument");
        return RBSC_ERR_INVALID;
    }
    memset(out, 0, sizeof(*out));
    const std::string base = strip_graph(basename);
    Props pr;
    int rc = read_props(base, &pr);
    if (rc) return rc;
    std::vector<uint64_t> row;
    std::vector<uint32_t> succ;
    rc = decode(base, pr, row, succ, nullptr);
    if (rc) return rc;
    out->n_nodes = (uint32_t)pr.nodes;
    out->n_edges = succ.size();
    out->row_offsets = (uint64_t*)malloc(sizeof(uint64_t) * row.size());
    out->dst = (uint32_t*)malloc(sizeof(uint32_t) * std::max<size_t>(succ.size(), 1));
    if (!out->row_offsets || !out->dst) {
        rbsc_host_graph_free(out);
        rbsc_set_error("host allocation failed");
        return RBSC_ERR_NOMEM;
    }
    memcpy(out->row_offsets, row.data(), sizeof(uint64_t) * row.size());
    if (!succ.empty()) memcpy(out->dst, succ.data(), sizeof(uint32_t) * succ.size());
    return RBSC_OK;
}

extern "C" int rbsc_bvgraph_check_offsets(const char* basename, uint64_t* n_offsets,
                                          uint64_t* mismatches) {
    if (!basename || !n_offsets || !mismatches) {
        rbsc_set_error("rbsc_bvgraph_check_offsets: null argument");
        return RBSC_ERR_INVALID;
    }
    const std::string base = strip_graph(basename);
    Props pr;
    int rc = read_props(base, &pr);
    if (rc) return rc;
    std::vector<uint64_t> row, starts;
    std::vector<uint32_t> succ;
    rc = decode(base, pr, row, succ, &starts);
    if (rc) return rc;
    std::vector<uint8_t> data;
    if (!read_file(base + ".offsets", data)) {
        rbsc_set_error("cannot read %s.offsets", base.c_str());
        return RBSC_ERR_INPUT;
    }
    Bits bs{data.data(), (uint64_t)(data.size() - 16) * 8};
    uint64_t pos = 0, k = 0, bad = 0;
    while (k <= pr.nodes) {
        const uint64_t save = bs.p;
        const uint64_t gap = bs.gamma();
        if (bs.over) {
            bs.p = save;
            break;
        }
        pos += gap;
        if (pos != starts[k]) ++bad;
        ++k;
    }
    *n_offsets = k;
    *mismatches = bad + (k < pr.nodes ? pr.nodes - k : 0);
    return RBSC_OK;
}
