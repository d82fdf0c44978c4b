// gen.cpp -- host-side seeded generators (see include/rbsc_gen.h).
#include "gen.h"
#include "../../include/rbsc_gen.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

extern "C" void rbsc_set_error(const char* fmt, ...);

namespace {

struct Rng {  // xoshiro256** seeded by SplitMix64
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) {
            x += 0x9e3779b97f4a7c15ULL;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    double unit() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

template <class T>
T* dup_array(const std::vector<T>& v) {
    T* p = (T*)malloc(sizeof(T) * (v.size() ? v.size() : 1));
    if (p && !v.empty()) memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

// Parallel LSD radix sort of u64 keys (8 passes of 8 bits).
void sort_keys(std::vector<uint64_t>& a) {
    const size_t n = a.size();
    if (n < (1u << 16)) {
        std::sort(a.begin(), a.end());
        return;
    }
    std::vector<uint64_t> b(n);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (int pass = 0; pass < 8; ++pass) {
        const int sh = pass * 8;
        std::vector<std::vector<size_t>> cnt(nt, std::vector<size_t>(256, 0));
        std::vector<std::thread> th;
        const size_t chunk = (n + nt - 1) / nt;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                const size_t lo = t * chunk, hi = std::min(n, lo + chunk);
                for (size_t i = lo; i < hi; ++i) cnt[t][(a[i] >> sh) & 255]++;
            });
        for (auto& x : th) x.join();
        th.clear();
        size_t run = 0;
        bool trivial = false;
        for (int d = 0; d < 256; ++d) {
            size_t tot = 0;
            for (unsigned t = 0; t < nt; ++t) tot += cnt[t][d];
            if (tot == n) trivial = true;
            for (unsigned t = 0; t < nt; ++t) {
                const size_t c = cnt[t][d];
                cnt[t][d] = run;
                run += c;
            }
        }
        if (trivial) continue;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                const size_t lo = t * chunk, hi = std::min(n, lo + chunk);
                for (size_t i = lo; i < hi; ++i) b[cnt[t][(a[i] >> sh) & 255]++] = a[i];
            });
        for (auto& x : th) x.join();
        a.swap(b);
    }
}

// CSR from sorted (src << 32 | dst) keys over n nodes.
int csr_from_keys(uint32_t n, const std::vector<uint64_t>& keys, rbsc_host_graph* out) {
    std::vector<uint64_t> ro((size_t)n + 1, 0);
    for (uint64_t k : keys) ro[(k >> 32) + 1]++;
    for (uint32_t u = 0; u < n; ++u) ro[u + 1] += ro[u];
    std::vector<uint32_t> dst(keys.size());
    for (size_t i = 0; i < keys.size(); ++i) dst[i] = (uint32_t)keys[i];
    memset(out, 0, sizeof(*out));
    out->n_nodes = n;
    out->n_edges = keys.size();
    out->row_offsets = dup_array(ro);
    out->dst = dup_array(dst);
    if (!out->row_offsets || !out->dst) {
        rbsc_host_graph_free(out);
        rbsc_set_error("host allocation failed");
        return RBSC_ERR_NOMEM;
    }
    return RBSC_OK;
}

}  // namespace

extern "C" void rbsc_host_graph_free(rbsc_host_graph* g) {
    if (!g) return;
    free(g->row_offsets);
    free(g->dst);
    free(g->weight);
    free(g->init_color);
    memset(g, 0, sizeof(*g));
}

extern "C" int rbsc_gen_rmat(uint32_t scale, uint64_t n_draws, double a, double b, double c,
                             double clone_frac, uint64_t seed, rbsc_host_graph* out) {
    if (!out || scale == 0 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0 ||
        clone_frac < 0 || clone_frac >= 1.0) {
        rbsc_set_error("rbsc_gen_rmat: invalid arguments");
        return RBSC_ERR_INVALID;
    }
    const RmatSpec s = rmat_spec(scale, n_draws, a, b, c, clone_frac, seed);
    const uint32_t N = (uint32_t)(s.mask + 1);
    std::vector<uint64_t> keys(n_draws);
    {
        unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        const uint64_t chunk = (n_draws + nt - 1) / nt;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                const uint64_t lo = t * chunk, hi = std::min(n_draws, lo + chunk);
                for (uint64_t e = lo; e < hi; ++e) keys[e] = rmat_draw(s, e);
            });
        for (auto& x : th) x.join();
    }
    sort_keys(keys);
    // rows of the draws, then clone rows copy their source's row
    std::vector<uint64_t> rp((size_t)N + 1, 0);
    for (uint64_t k : keys) rp[(k >> 32) + 1]++;
    for (uint32_t u = 0; u < N; ++u) rp[u + 1] += rp[u];
    std::vector<uint64_t> ro((size_t)N + 1, 0);
    for (uint32_t x = 0; x < N; ++x) {
        const uint64_t y = rmat_source(s, x);
        ro[x + 1] = ro[x] + (rp[y + 1] - rp[y]);
    }
    std::vector<uint32_t> dst(ro[N]);
    for (uint32_t x = 0; x < N; ++x) {
        const uint64_t y = rmat_source(s, x);
        for (uint64_t i = rp[y], o = ro[x]; i < rp[y + 1]; ++i, ++o) dst[o] = (uint32_t)keys[i];
    }
    memset(out, 0, sizeof(*out));
    out->n_nodes = N;
    out->n_edges = ro[N];
    out->row_offsets = dup_array(ro);
    out->dst = dup_array(dst);
    if (!out->row_offsets || !out->dst) {
        rbsc_host_graph_free(out);
        rbsc_set_error("host allocation failed");
        return RBSC_ERR_NOMEM;
    }
    return RBSC_OK;
}

extern "C" int rbsc_gen_ba(uint32_t n, uint32_t m, double dup_frac, uint32_t n_init_colors,
                           uint64_t seed, rbsc_host_graph* out) {
    if (!out || m == 0 || n < m + 1) {
        rbsc_set_error("rbsc_gen_ba: need n >= m + 1 and m >= 1");
        return RBSC_ERR_INVALID;
    }
    Rng rng(seed);
    std::vector<uint32_t> pool;  // node v appears 1 + in-degree(v) times
    pool.reserve((size_t)n * (m + 1));
    std::vector<uint64_t> keys;
    keys.reserve((size_t)n * m * 6 / 5 + 16);
    for (uint32_t t = 0; t <= m; ++t) pool.push_back(t);
    for (uint32_t t = 0; t <= m; ++t)
        for (uint32_t v = 0; v <= m; ++v)
            if (v != t) {
                keys.push_back(((uint64_t)t << 32) | v);
                pool.push_back(v);
            }
    std::vector<uint32_t> tg(m);
    for (uint32_t t = m + 1; t < n; ++t) {
        for (uint32_t j = 0; j < m; ++j) tg[j] = pool[rng.below(pool.size())];
        pool.push_back(t);
        for (uint32_t j = 0; j < m; ++j) {
            keys.push_back(((uint64_t)t << 32) | tg[j]);
            pool.push_back(tg[j]);
        }
    }
    const size_t base_n = keys.size();
    for (size_t i = 0; i < base_n; ++i) {
        if (rng.unit() < dup_frac) {
            const uint64_t extra = 1 + rng.below(3);
            for (uint64_t r = 0; r < extra; ++r) keys.push_back(keys[i]);
        }
    }
    sort_keys(keys);
    int rc = csr_from_keys(n, keys, out);
    if (rc) return rc;
    if (n_init_colors) {
        std::vector<uint32_t> ic(n);
        Rng cr(seed ^ 0x1c0104ULL);
        for (uint32_t u = 0; u < n; ++u) ic[u] = (uint32_t)cr.below(n_init_colors);
        out->init_color = dup_array(ic);
    }
    return RBSC_OK;
}

extern "C" int rbsc_gen_er_undirected(uint32_t n, double avg_deg, uint64_t seed,
                                      rbsc_host_graph* out) {
    if (!out || n < 2 || avg_deg < 0) {
        rbsc_set_error("rbsc_gen_er_undirected: invalid arguments");
        return RBSC_ERR_INVALID;
    }
    Rng rng(seed);
    const uint64_t m = (uint64_t)(avg_deg * n / 2.0 + 0.5);
    std::vector<uint64_t> keys;
    keys.reserve(2 * m);
    for (uint64_t i = 0; i < m; ++i) {
        uint32_t u = (uint32_t)rng.below(n), v = (uint32_t)rng.below(n - 1);
        if (v >= u) ++v;
        keys.push_back(((uint64_t)u << 32) | v);
        keys.push_back(((uint64_t)v << 32) | u);
    }
    sort_keys(keys);
    return csr_from_keys(n, keys, out);
}

extern "C" int rbsc_gen_ba_undirected(uint32_t n, uint32_t m, uint64_t seed, rbsc_host_graph* out) {
    if (!out || m == 0 || n < m + 1) {
        rbsc_set_error("rbsc_gen_ba_undirected: need n >= m + 1 and m >= 1");
        return RBSC_ERR_INVALID;
    }
    Rng rng(seed);
    std::vector<uint32_t> pool;
    std::vector<uint64_t> keys;
    for (uint32_t t = 0; t <= m; ++t)
        for (uint32_t v = t + 1; v <= m; ++v) {
            keys.push_back(((uint64_t)t << 32) | v);
            keys.push_back(((uint64_t)v << 32) | t);
            pool.push_back(t);
            pool.push_back(v);
        }
    if (pool.empty()) pool.push_back(0);
    std::vector<uint32_t> tg;
    for (uint32_t t = m + 1; t < n; ++t) {
        tg.clear();
        while (tg.size() < m) {  // distinct targets
            const uint32_t v = pool[rng.below(pool.size())];
            if (std::find(tg.begin(), tg.end(), v) == tg.end()) tg.push_back(v);
        }
        for (uint32_t v : tg) {
            keys.push_back(((uint64_t)t << 32) | v);
            keys.push_back(((uint64_t)v << 32) | t);
            pool.push_back(t);
            pool.push_back(v);
        }
    }
    sort_keys(keys);
    return csr_from_keys(n, keys, out);
}

extern "C" int rbsc_gen_regular_undirected(uint32_t n, uint32_t k, uint64_t seed,
                                           rbsc_host_graph* out) {
    if (!out || k == 0 || k >= n || ((uint64_t)n * k) % 2) {
        rbsc_set_error("rbsc_gen_regular_undirected: need 0 < k < n and n*k even");
        return RBSC_ERR_INVALID;
    }
    Rng rng(seed);
    std::vector<uint32_t> stubs((size_t)n * k);
    std::vector<uint64_t> keys;
    for (int attempt = 0; attempt < 64; ++attempt) {
        for (uint32_t u = 0; u < n; ++u)
            for (uint32_t j = 0; j < k; ++j) stubs[(size_t)u * k + j] = u;
        for (size_t i = stubs.size(); i > 1; --i) std::swap(stubs[i - 1], stubs[rng.below(i)]);
        keys.clear();
        bool ok = true;
        for (size_t i = 0; i + 1 < stubs.size(); i += 2) {
            const uint32_t u = stubs[i], v = stubs[i + 1];
            if (u == v) { ok = false; break; }
            keys.push_back(((uint64_t)u << 32) | v);
            keys.push_back(((uint64_t)v << 32) | u);
        }
        if (ok) {
            std::sort(keys.begin(), keys.end());
            if (std::adjacent_find(keys.begin(), keys.end()) == keys.end()) break;
        }
        if (attempt == 63) {  // accept a multigraph rather than fail
            keys.clear();
            for (size_t i = 0; i + 1 < stubs.size(); i += 2) {
                keys.push_back(((uint64_t)stubs[i] << 32) | stubs[i + 1]);
                keys.push_back(((uint64_t)stubs[i + 1] << 32) | stubs[i]);
            }
        }
    }
    sort_keys(keys);
    return csr_from_keys(n, keys, out);
}

extern "C" void rbsc_host_batch_free(rbsc_host_batch* b) {
    if (!b) return;
    free(b->node_offsets);
    free(b->line_offsets);
    free(b->src);
    free(b->dst);
    memset(b, 0, sizeof(*b));
}

extern "C" int rbsc_gen_batch_mixed(uint32_t n_graphs, uint32_t n_min, uint32_t n_max,
                                    uint64_t seed, rbsc_host_batch* out) {
    if (!out || n_min < 5 || n_max < n_min) {
        rbsc_set_error("rbsc_gen_batch_mixed: need 5 <= n_min <= n_max");
        return RBSC_ERR_INVALID;
    }
    std::vector<rbsc_host_graph> gs(n_graphs);
    std::vector<int> rcs(n_graphs, 0);
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (uint64_t i = t; i < n_graphs; i += nt) {
                Rng r(seed * 0x9e3779b97f4a7c15ULL + i);
                const double lo = log((double)n_min), hi = log((double)n_max);
                uint32_t n = (uint32_t)floor(exp(lo + (hi - lo) * r.unit()) + 0.5);
                n = std::min(std::max(n, n_min), n_max);
                const uint64_t gseed = r.next();
                memset(&gs[i], 0, sizeof(gs[i]));
                if (i % 3 == 0) {
                    rcs[i] = rbsc_gen_er_undirected(n, 2.0 + 6.0 * r.unit(), gseed, &gs[i]);
                } else if (i % 3 == 1) {
                    const uint32_t m = 1 + (uint32_t)r.below(4);
                    rcs[i] = rbsc_gen_ba_undirected(std::max(n, m + 1), m, gseed, &gs[i]);
                } else {
                    uint32_t k = 3 + (uint32_t)r.below(2);
                    if ((uint64_t)n * k % 2) k = 4;
                    rcs[i] = rbsc_gen_regular_undirected(n, k, gseed, &gs[i]);
                }
            }
        });
    for (auto& x : th) x.join();
    for (uint32_t i = 0; i < n_graphs; ++i)
        if (rcs[i]) {
            for (auto& g : gs) rbsc_host_graph_free(&g);
            return rcs[i];
        }
    memset(out, 0, sizeof(*out));
    out->n_graphs = n_graphs;
    out->node_offsets = (uint64_t*)malloc(sizeof(uint64_t) * (n_graphs + 1));
    out->line_offsets = (uint64_t*)malloc(sizeof(uint64_t) * (n_graphs + 1));
    out->node_offsets[0] = out->line_offsets[0] = 0;
    for (uint32_t i = 0; i < n_graphs; ++i) {
        out->node_offsets[i + 1] = out->node_offsets[i] + gs[i].n_nodes;
        out->line_offsets[i + 1] = out->line_offsets[i] + gs[i].n_edges;
    }
    const uint64_t L = out->line_offsets[n_graphs];
    out->src = (uint32_t*)malloc(sizeof(uint32_t) * (L ? L : 1));
    out->dst = (uint32_t*)malloc(sizeof(uint32_t) * (L ? L : 1));
    if (!out->node_offsets || !out->line_offsets || !out->src || !out->dst) {
        for (auto& g : gs) rbsc_host_graph_free(&g);
        rbsc_host_batch_free(out);
        rbsc_set_error("host allocation failed");
        return RBSC_ERR_NOMEM;
    }
    th.clear();
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (uint64_t i = t; i < n_graphs; i += nt) {
                const rbsc_host_graph& g = gs[i];
                uint64_t o = out->line_offsets[i];
                for (uint32_t u = 0; u < g.n_nodes; ++u)
                    for (uint64_t e = g.row_offsets[u]; e < g.row_offsets[u + 1]; ++e, ++o) {
                        out->src[o] = u;
                        out->dst[o] = g.dst[e];
                    }
                rbsc_host_graph_free(&gs[i]);
            }
        });
    for (auto& x : th) x.join();
    return RBSC_OK;
}

// ------------------------------------------------------------ text output
namespace {
inline char* put_u64(char* p, uint64_t v) {
    char tmp[24];
    int k = 0;
    do {
        tmp[k++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (k) *p++ = tmp[--k];
    return p;
}
}  // namespace

extern "C" int rbsc_host_graph_write_text(const rbsc_host_graph* g, const char* path, int format) {
    if (!g || !path || (format != RBSC_TEXT_LINES && format != RBSC_TEXT_VALMARI) ||
        (g->n_edges && (!g->row_offsets || !g->dst))) {
        rbsc_set_error("rbsc_host_graph_write_text: bad arguments");
        return RBSC_ERR_INVALID;
    }
    FILE* f = fopen(path, "wb");
    if (!f) {
        rbsc_set_error("rbsc_host_graph_write_text: cannot open %s", path);
        return RBSC_ERR_INPUT;
    }
    const int one = format == RBSC_TEXT_VALMARI ? 1 : 0;  // Valmari ids are 1-based
    std::vector<char> buf(1 << 22);
    char* p = buf.data();
    char* end = buf.data() + buf.size() - 96;
    bool ok = true;
    auto flush = [&]() {
        if (fwrite(buf.data(), 1, (size_t)(p - buf.data()), f) != (size_t)(p - buf.data())) ok = false;
        p = buf.data();
    };
    if (format == RBSC_TEXT_LINES) {
        // rbsc.cu:168-180: N, E, then "src weight dst"
        p = put_u64(p, g->n_nodes);
        *p++ = '\n';
        p = put_u64(p, g->n_edges);
        *p++ = '\n';
    } else {
        // MDP_read.cc:51-64 as valmari-adapter:9-16 writes it: "N 0 0 E blocks"
        uint32_t blocks = 1;
        if (g->init_color) {
            std::vector<uint32_t> c(g->init_color, g->init_color + g->n_nodes);
            std::sort(c.begin(), c.end());
            blocks = (uint32_t)(std::unique(c.begin(), c.end()) - c.begin());
            if (!blocks) blocks = 1;
        }
        p = put_u64(p, g->n_nodes);
        memcpy(p, " 0 0 ", 5);
        p += 5;
        p = put_u64(p, g->n_edges);
        *p++ = ' ';
        p = put_u64(p, blocks);
        *p++ = '\n';
    }
    flush();
    // The lines, formatted by T threads in blocks of BL consecutive lines
    // (block b by thread b mod T, into its own buffer) and written in order.
    {
        const uint64_t E = g->n_edges, BL = 1 << 20;
        const unsigned T = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
        std::vector<std::vector<char>> tb(T, std::vector<char>(BL * 64));
        std::vector<size_t> used(T, 0);
        auto fmt = [&](unsigned k, uint64_t b) {
            const uint64_t e0 = b * BL, e1 = std::min(E, e0 + BL);
            // the row holding line e0: last u with row_offsets[u] <= e0
            const uint64_t* ro = g->row_offsets;
            uint64_t u = (uint64_t)(std::upper_bound(ro, ro + g->n_nodes + 1, e0) - ro) - 1;
            char* q = tb[k].data();
            for (uint64_t e = e0; e < e1; ++e) {
                while (ro[u + 1] <= e) ++u;
                q = put_u64(q, u + one);
                *q++ = ' ';
                q = put_u64(q, g->weight ? g->weight[e] : 1);
                *q++ = ' ';
                q = put_u64(q, (uint64_t)g->dst[e] + one);
                *q++ = '\n';
            }
            used[k] = (size_t)(q - tb[k].data());
        };
        const uint64_t nb = (E + BL - 1) / BL;
        for (uint64_t b0 = 0; b0 < nb && ok; b0 += T) {
            std::vector<std::thread> th;
            for (unsigned k = 0; k < T && b0 + k < nb; ++k) th.emplace_back(fmt, k, b0 + k);
            for (auto& t : th) t.join();
            for (unsigned k = 0; k < T && b0 + k < nb && ok; ++k)
                if (fwrite(tb[k].data(), 1, used[k], f) != used[k]) ok = false;
        }
    }
    if (format == RBSC_TEXT_VALMARI && g->init_color) {
        // MDP_read.cc:140-150: every initial block but the first, members then 0
        std::vector<std::pair<uint32_t, uint32_t>> cu(g->n_nodes);
        for (uint32_t u = 0; u < g->n_nodes; ++u) cu[u] = {g->init_color[u], u};
        std::sort(cu.begin(), cu.end());
        for (size_t i = 0; i < cu.size() && ok; ++i) {
            if (i == 0 || cu[i].first == cu[0].first) continue;
            p = put_u64(p, (uint64_t)cu[i].second + 1);
            const bool last = i + 1 == cu.size() || cu[i + 1].first != cu[i].first;
            if (last) {
                memcpy(p, " 0\n", 3);
                p += 3;
            } else {
                *p++ = ' ';
            }
            if (p >= end) flush();
        }
    }
    flush();
    if (fclose(f) != 0) ok = false;
    if (!ok) {
        rbsc_set_error("rbsc_host_graph_write_text: write to %s failed", path);
        return RBSC_ERR_INPUT;
    }
    return RBSC_OK;
}
