// serve.hip -- continuous-batching request loop on one device (host code).
//
// Requests arrive at their arrival times (submit delay), wait in FIFO order
// until their HBM estimate fits the budget, and are then admitted:
//   * graphs of at most 5120 nodes join the small queue; a batcher thread
//     takes every small request queued at that moment and refines them in
//     one persistent batch launch (rbsc_batch_*), while new ones queue for
//     the next launch;
//   * larger graphs go to a pool of workers, each with its own context
//     (stream and workspace), so several refinements and the batch launch
//     overlap on the device.
// Every graph is refined to its own fixpoint by the same round function, so
// a result is identical alone and in any mix.  No counterpart in the
// reference, which refines one graph per process.
#include "internal.h"
#include "batch.h"
#include "../../include/rbsc_gen.h"
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <queue>
#include <string>
#include <thread>
#include <vector>

using Clock = std::chrono::steady_clock;

namespace {

enum ReqKind { REQ_HOST = 0, REQ_RMAT = 1, REQ_RESIDENT = 2 };

struct Req {
    u64 ticket = 0;
    int kind = REQ_HOST;
    u32 n = 0;
    u64 e = 0;
    // REQ_HOST: a copy of the graph, lines grouped by source (COO)
    std::vector<u32> src, dst, init;
    std::vector<u64> wt;
    // REQ_RMAT
    u32 scale = 0;
    u64 draws = 0;
    double a = 0, b = 0, c = 0, clone = 0;
    u64 seed = 0;
    // REQ_RESIDENT (owned by the caller)
    rbsc_dgraph* g = nullptr;
    bool small = false;
    u64 need = 0;  // HBM estimate charged while in flight
    Clock::time_point t_arrival, t_admit, t_start, t_done;
    // result
    bool done = false;
    int status = RBSC_OK;
    std::string error;
    u32 blocks = 0, rounds = 0, batched = 0;
    std::vector<u32> partition;  // batched requests
    u32* part = nullptr;         // single-graph requests: a page-locked pool buffer
};

// Page-locked buffers for partitions, kept for reuse by size class (powers
// of two, from 4096 labels): a refinement writes its labels straight into
// them (rbsc_host_alloc memory, no staging copy), a large graph's buffer is
// not pinned again for every request, and rbsc_server_wait_labels hands a
// buffer to the caller instead of copying it.  One pool serves the process
// (servers come and go; buffers handed out may outlive their server); it
// keeps at most POOL_KEEP bytes of free buffers and is never destroyed, so
// no buffer is freed after the runtime shuts down.
class PartPool {
  public:
    static constexpr u64 POOL_KEEP = 8ull << 30;
    u32* get(u64 n) {
        u64 c = 4096;
        while (c < n) c *= 2;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto& v = free_[c];
            if (!v.empty()) {
                u32* p = v.back();
                v.pop_back();
                kept_ -= 4 * c;
                out_[p] = c;
                return p;
            }
        }
        void* p = nullptr;
        if (rbsc_host_alloc(sizeof(u32) * c, &p) != RBSC_OK) return nullptr;
        std::lock_guard<std::mutex> lk(mu_);
        out_[static_cast<u32*>(p)] = c;
        return static_cast<u32*>(p);
    }
    // false: p is not a buffer of this pool that is out
    bool put(u32* p) {
        if (!p) return true;
        std::unique_lock<std::mutex> lk(mu_);
        auto it = out_.find(p);
        if (it == out_.end()) return false;
        const u64 c = it->second;
        out_.erase(it);
        if (kept_ + 4 * c <= POOL_KEEP) {
            free_[c].push_back(p);
            kept_ += 4 * c;
            return true;
        }
        lk.unlock();
        rbsc_host_free(p);
        return true;
    }

  private:
    std::mutex mu_;
    std::map<u64, std::vector<u32*>> free_;
    std::map<u32*, u64> out_;
    u64 kept_ = 0;
};
PartPool& part_pool() {
    static PartPool* p = new PartPool();
    return *p;
}

struct ByArrival {
    bool operator()(const Req* x, const Req* y) const {
        return x->t_arrival > y->t_arrival || (x->t_arrival == y->t_arrival && x->ticket > y->ticket);
    }
};

// HBM estimates (bytes) of a request while it is in flight: its graph as
// built on the device (with the build's temporaries) and the refinement
// workspace (about 56 bytes per node).
u64 need_host(u64 n, u64 e, bool wt) { return 9 * e + (wt ? 16 * e : 0) + 160 * n + 4096; }
u64 need_rmat(u64 n, u64 draws) { return 24 * draws + 160 * n + 4096; }
u64 need_resident(u64 n) { return 64 * n + 4096; }
u64 need_small(u64 n, u64 e, bool wt) { return 2 * e + (wt ? 8 * e : 0) + 24 * n + 64; }

}  // namespace

struct rbsc_server {
    rbsc_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv_sched, cv_work, cv_small, cv_done;
    bool stop = false;
    u64 next_ticket = 1;
    std::map<u64, std::unique_ptr<Req>> reqs;
    std::priority_queue<Req*, std::vector<Req*>, ByArrival> future;
    std::deque<Req*> pending, large_q, small_q;
    u64 admitted = 0, peak_admitted = 0;
    u64 submitted = 0, completed = 0, failed = 0, batches = 0, batched_graphs = 0;
    u64 in_flight = 0;  // admitted, not completed
    std::vector<rbsc_ctx*> ctxs;  // workers..., batcher last
    std::vector<std::thread> threads;
};

namespace {

void finish(rbsc_server* s, Req* r, int status, const char* err) {
    r->status = status;
    if (status) r->error = err ? err : "";
    r->t_done = Clock::now();
    r->done = true;
    s->admitted -= r->need;
    --s->in_flight;
    ++s->completed;
    if (status) ++s->failed;
    // inputs are not needed any more
    std::vector<u32>().swap(r->src);
    std::vector<u32>().swap(r->dst);
    std::vector<u32>().swap(r->init);
    std::vector<u64>().swap(r->wt);
}

// Moves arrived requests to the FIFO and admits from its head while the
// head fits the budget.  Called with the lock held.
void admit_locked(rbsc_server* s, Clock::time_point now) {
    while (!s->future.empty() && s->future.top()->t_arrival <= now) {
        s->pending.push_back(s->future.top());
        s->future.pop();
    }
    while (!s->pending.empty()) {
        Req* r = s->pending.front();
        if (r->need > s->opts.hbm_budget) {  // can never fit
            s->pending.pop_front();
            r->t_admit = r->t_start = now;
            ++s->in_flight;
            s->admitted += r->need;
            finish(s, r, RBSC_ERR_BUDGET, "request needs more HBM than the server budget");
            s->cv_done.notify_all();
            continue;
        }
        if (s->admitted + r->need > s->opts.hbm_budget) break;  // wait for memory
        s->pending.pop_front();
        s->admitted += r->need;
        s->peak_admitted = std::max(s->peak_admitted, s->admitted);
        ++s->in_flight;
        r->t_admit = now;
        if (r->small) {
            s->small_q.push_back(r);
            s->cv_small.notify_one();
        } else {
            s->large_q.push_back(r);
            s->cv_work.notify_one();
        }
    }
}

void scheduler_main(rbsc_server* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        admit_locked(s, Clock::now());
        if (s->stop && s->future.empty() && s->pending.empty()) break;
        if (!s->future.empty())
            s->cv_sched.wait_until(lk, s->future.top()->t_arrival);
        else
            s->cv_sched.wait(lk);
    }
}

void worker_main(rbsc_server* s, rbsc_ctx* ctx) {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        Req* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->large_q.empty(); });
            if (s->large_q.empty()) return;
            r = s->large_q.front();
            s->large_q.pop_front();
            r->t_start = Clock::now();
        }
        rbsc_dgraph* g = nullptr;
        int rc = RBSC_OK;
        if (r->kind == REQ_HOST) {
            rbsc_graph gr{};
            gr.n_nodes = r->n;
            gr.n_edges = r->e;
            gr.src = r->src.data();
            gr.dst = r->dst.data();
            gr.weight = r->wt.empty() ? nullptr : r->wt.data();
            gr.init_color = r->init.empty() ? nullptr : r->init.data();
            rc = graph_upload_host(ctx, &gr, false, &g, true);  // refined once
        } else if (r->kind == REQ_RMAT) {
            rc = rbsc_gen_rmat_device(ctx, r->scale, r->draws, r->a, r->b, r->c, r->clone, r->seed, &g);
        } else {
            g = r->g;
        }
        rbsc_result res{};
        if (!rc) {
            rbsc_opts o{};
            if (s->opts.keep_partition && r->n) {
                r->part = part_pool().get(r->n);
                if (!r->part) rc = RBSC_ERR_NOMEM;
            }
            if (!rc) rc = rbsc_graph_refine(ctx, g, &o, &res, r->part, nullptr);
        }
        const std::string err = rc ? rbsc_last_error() : "";
        if (g && r->kind != REQ_RESIDENT) rbsc_graph_free(ctx, g);
        if (ctx->ws.cap_n > (1u << 20)) ws_release(ctx);  // hand large workspaces back
        std::lock_guard<std::mutex> lk(s->mu);
        r->blocks = res.nodes;
        r->rounds = res.rounds;
        finish(s, r, rc, err.c_str());
        s->cv_done.notify_all();
        s->cv_sched.notify_one();
    }
}

// One batch launch over a group of small requests with the same kind of
// initial colouring.
int run_small_group(rbsc_server* s, rbsc_ctx* ctx, std::vector<Req*>& grp) {
    const u32 G = (u32)grp.size();
    std::vector<u64> no(G + 1, 0), lo(G + 1, 0);
    bool any_wt = false, has_init = false;
    for (u32 i = 0; i < G; ++i) {
        no[i + 1] = no[i] + grp[i]->n;
        lo[i + 1] = lo[i] + grp[i]->e;
        any_wt |= !grp[i]->wt.empty();
        has_init = !grp[i]->init.empty();
    }
    std::vector<u32> src(lo[G]), dst(lo[G]), init(has_init ? no[G] : 0);
    std::vector<u64> wt(any_wt ? lo[G] : 0);
    for (u32 i = 0; i < G; ++i) {
        Req* r = grp[i];
        std::copy(r->src.begin(), r->src.end(), src.begin() + lo[i]);
        std::copy(r->dst.begin(), r->dst.end(), dst.begin() + lo[i]);
        if (any_wt) {
            if (r->wt.empty()) std::fill(wt.begin() + lo[i], wt.begin() + lo[i + 1], 1ULL);
            else std::copy(r->wt.begin(), r->wt.end(), wt.begin() + lo[i]);
        }
        if (has_init) std::copy(r->init.begin(), r->init.end(), init.begin() + no[i]);
    }
    rbsc_batch b{};
    b.n_graphs = G;
    b.node_offsets = no.data();
    b.line_offsets = lo.data();
    b.src = src.data();
    b.dst = dst.data();
    b.weight = any_wt ? wt.data() : nullptr;
    b.init_color = has_init ? init.data() : nullptr;
    std::vector<u32> blocks(G), rounds(G), labels(s->opts.keep_partition ? std::max<u64>(no[G], 1) : 0);
    rbsc_dbatch* db = nullptr;
    int rc = rbsc_batch_upload(ctx, &b, &db);
    if (!rc) {
        rbsc_opts o{};
        rc = rbsc_batch_refine(ctx, db, &o, blocks.data(), rounds.data(),
                               s->opts.keep_partition ? labels.data() : nullptr, nullptr, nullptr);
        rbsc_batch_free(ctx, db);
    }
    const std::string err = rc ? rbsc_last_error() : "";
    for (u32 i = 0; i < G; ++i) {
        Req* r = grp[i];
        r->blocks = blocks[i];
        r->rounds = rounds[i];
        r->batched = 1;
        if (!rc && s->opts.keep_partition)
            r->partition.assign(labels.begin() + no[i], labels.begin() + no[i + 1]);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    ++s->batches;
    s->batched_graphs += G;
    for (Req* r : grp) finish(s, r, rc, err.c_str());
    s->cv_done.notify_all();
    s->cv_sched.notify_one();
    return rc;
}

void batcher_main(rbsc_server* s, rbsc_ctx* ctx) {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        std::vector<Req*> take;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_small.wait(lk, [&] { return s->stop || !s->small_q.empty(); });
            if (s->small_q.empty()) return;
            const size_t m = std::min<size_t>(s->small_q.size(), s->opts.batch_max);
            const Clock::time_point now = Clock::now();
            for (size_t i = 0; i < m; ++i) {
                take.push_back(s->small_q.front());
                s->small_q.pop_front();
                take.back()->t_start = now;
            }
        }
        // R-MAT requests of small graphs are generated here, on the host.
        std::vector<Req*> ready;
        for (Req* r : take) {
            if (r->kind != REQ_RMAT) {
                ready.push_back(r);
                continue;
            }
            rbsc_host_graph hg{};
            const int rc = rbsc_gen_rmat(r->scale, r->draws, r->a, r->b, r->c, r->clone, r->seed, &hg);
            if (rc) {
                const std::string err = rbsc_last_error();
                std::lock_guard<std::mutex> lk(s->mu);
                finish(s, r, rc, err.c_str());
                s->cv_done.notify_all();
                s->cv_sched.notify_one();
                continue;
            }
            ready.push_back(r);
            r->src.resize(hg.n_edges);
            r->dst.assign(hg.dst, hg.dst + hg.n_edges);
            for (u32 u = 0; u < hg.n_nodes; ++u)
                for (u64 k = hg.row_offsets[u]; k < hg.row_offsets[u + 1]; ++k) r->src[k] = u;
            r->e = hg.n_edges;
            rbsc_host_graph_free(&hg);
        }
        std::vector<Req*> plain, colored;
        for (Req* r : ready) (r->init.empty() ? plain : colored).push_back(r);
        if (!plain.empty()) run_small_group(s, ctx, plain);
        if (!colored.empty()) run_small_group(s, ctx, colored);
    }
}

u64 submit_locked(rbsc_server* s, std::unique_ptr<Req> r, double delay_s) {
    r->ticket = s->next_ticket++;
    r->t_arrival = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                      std::chrono::duration<double>(std::max(0.0, delay_s)));
    Req* p = r.get();
    const u64 t = p->ticket;
    s->reqs[t] = std::move(r);
    s->future.push(p);
    ++s->submitted;
    s->cv_sched.notify_one();
    return t;
}

}  // namespace

extern "C" int rbsc_server_create(const rbsc_server_opts* o, rbsc_server** out) {
    if (!o || !out) {
        rbsc_set_error("rbsc_server_create: null argument");
        return RBSC_ERR_INVALID;
    }
    *out = nullptr;
    auto* s = new rbsc_server();
    s->opts = *o;
    if (!s->opts.hbm_budget) s->opts.hbm_budget = 250ull << 30;
    if (!s->opts.workers) s->opts.workers = 8;
    if (s->opts.workers > 16) s->opts.workers = 16;
    if (!s->opts.batch_max) s->opts.batch_max = 4096;
    for (u32 i = 0; i <= s->opts.workers; ++i) {
        rbsc_ctx* c = nullptr;
        const int rc = rbsc_ctx_create(s->opts.device, 0, &c);
        if (rc) {
            for (rbsc_ctx* x : s->ctxs) rbsc_ctx_destroy(x);
            delete s;
            return rc;
        }
        s->ctxs.push_back(c);
    }
    s->threads.emplace_back(scheduler_main, s);
    for (u32 i = 0; i < s->opts.workers; ++i) s->threads.emplace_back(worker_main, s, s->ctxs[i]);
    s->threads.emplace_back(batcher_main, s, s->ctxs.back());
    *out = s;
    return RBSC_OK;
}

extern "C" int rbsc_server_destroy(rbsc_server* s) {
    if (!s) return RBSC_OK;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        // every submitted request runs to completion first
        s->cv_done.wait(lk, [&] { return s->completed == s->submitted; });
        s->stop = true;
    }
    s->cv_sched.notify_all();
    s->cv_work.notify_all();
    s->cv_small.notify_all();
    for (auto& t : s->threads) t.join();
    for (rbsc_ctx* c : s->ctxs) rbsc_ctx_destroy(c);
    // partitions never collected go back to the pool
    for (auto& kv : s->reqs) {
        part_pool().put(kv.second->part);
        kv.second->part = nullptr;
    }
    delete s;
    return RBSC_OK;
}

extern "C" int rbsc_server_submit(rbsc_server* s, const rbsc_graph* g, double delay_s,
                                  uint64_t* ticket) {
    if (!s || !g || !ticket || (g->n_edges && (!g->dst || (!g->src && !g->row_offsets)))) {
        rbsc_set_error("rbsc_server_submit: null argument");
        return RBSC_ERR_INVALID;
    }
    auto r = std::make_unique<Req>();
    r->kind = REQ_HOST;
    r->n = g->n_nodes;
    r->e = g->n_edges;
    r->dst.assign(g->dst, g->dst + g->n_edges);
    if (g->src) {
        r->src.assign(g->src, g->src + g->n_edges);
    } else {
        r->src.resize(g->n_edges);
        if (g->row_offsets[0] != 0 || g->row_offsets[g->n_nodes] != g->n_edges) {
            rbsc_set_error("rbsc_server_submit: row_offsets must span the lines");
            return RBSC_ERR_INPUT;
        }
        for (u32 u = 0; u < g->n_nodes; ++u) {
            if (g->row_offsets[u + 1] < g->row_offsets[u]) {
                rbsc_set_error("rbsc_server_submit: row_offsets not monotone");
                return RBSC_ERR_INPUT;
            }
            for (u64 k = g->row_offsets[u]; k < g->row_offsets[u + 1]; ++k) r->src[k] = u;
        }
    }
    if (g->weight) r->wt.assign(g->weight, g->weight + g->n_edges);
    if (g->init_color) r->init.assign(g->init_color, g->init_color + g->n_nodes);
    r->small = r->n <= BATCH_NMAX;
    r->need = r->small ? need_small(r->n, r->e, g->weight != nullptr)
                       : need_host(r->n, r->e, g->weight != nullptr);
    std::lock_guard<std::mutex> lk(s->mu);
    *ticket = submit_locked(s, std::move(r), delay_s);
    return RBSC_OK;
}

extern "C" int rbsc_server_submit_rmat(rbsc_server* s, uint32_t scale, uint64_t n_draws, double a,
                                       double b, double c, double clone_frac, uint64_t seed,
                                       double delay_s, uint64_t* ticket) {
    if (!s || !ticket || scale == 0 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0 ||
        clone_frac < 0 || clone_frac >= 1.0) {
        rbsc_set_error("rbsc_server_submit_rmat: invalid arguments");
        return RBSC_ERR_INVALID;
    }
    auto r = std::make_unique<Req>();
    r->kind = REQ_RMAT;
    r->n = 1u << scale;
    r->e = n_draws;
    r->scale = scale;
    r->draws = n_draws;
    r->a = a;
    r->b = b;
    r->c = c;
    r->clone = clone_frac;
    r->seed = seed;
    r->small = r->n <= BATCH_NMAX;
    // clones can lengthen rows: bound the line count by twice the draws
    r->need = r->small ? need_small(r->n, 2 * n_draws, false) : need_rmat(r->n, n_draws);
    std::lock_guard<std::mutex> lk(s->mu);
    *ticket = submit_locked(s, std::move(r), delay_s);
    return RBSC_OK;
}

extern "C" int rbsc_server_submit_resident(rbsc_server* s, rbsc_dgraph* g, double delay_s,
                                           uint64_t* ticket) {
    if (!s || !g || !ticket) {
        rbsc_set_error("rbsc_server_submit_resident: null argument");
        return RBSC_ERR_INVALID;
    }
    auto r = std::make_unique<Req>();
    r->kind = REQ_RESIDENT;
    r->g = g;
    r->n = g->N;
    r->e = g->E_in;
    r->small = false;  // already laid out for the single-graph path
    r->need = need_resident(g->N);
    std::lock_guard<std::mutex> lk(s->mu);
    *ticket = submit_locked(s, std::move(r), delay_s);
    return RBSC_OK;
}

namespace {
// The request's result once it has finished (or RBSC_PENDING); the request
// leaves the table and is handed to the caller, whose copy of its labels
// then runs without the server's lock.
int take_result(rbsc_server* s, uint64_t ticket, double timeout_s, rbsc_request_result* out,
                std::unique_ptr<Req>* own, const char* who) {
    if (!s || !out) {
        rbsc_set_error("%s: null argument", who);
        return RBSC_ERR_INVALID;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) {
        rbsc_set_error("%s: unknown ticket %llu", who, (unsigned long long)ticket);
        return RBSC_ERR_INVALID;
    }
    Req* r = it->second.get();
    auto ready = [&] { return r->done; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
        rbsc_set_error("%s: request %llu not finished", who, (unsigned long long)ticket);
        return RBSC_PENDING;
    }
    rbsc_request_result o{};
    o.ticket = r->ticket;
    o.status = r->status;
    o.nodes = r->blocks;
    o.rounds = r->rounds;
    o.batched = r->batched;
    o.n_nodes = r->n;
    o.n_edges = r->e;
    auto sec = [](Clock::duration d) { return std::chrono::duration<double>(d).count(); };
    o.latency_s = sec(r->t_done - r->t_arrival);
    o.queue_s = sec(r->t_admit - r->t_arrival);
    o.service_s = sec(r->t_done - r->t_start);
    o.arrival_s = sec(r->t_arrival.time_since_epoch());
    o.done_s = sec(r->t_done.time_since_epoch());
    *out = o;
    *own = std::move(it->second);
    s->reqs.erase(it);
    return RBSC_OK;
}
}  // namespace

extern "C" int rbsc_server_wait(rbsc_server* s, uint64_t ticket, double timeout_s,
                                rbsc_request_result* out, uint32_t* partition_out) {
    std::unique_ptr<Req> own;
    const int rc = take_result(s, ticket, timeout_s, out, &own, "rbsc_server_wait");
    if (rc) return rc;
    const int st = own->status;
    if (partition_out && !st) {
        if (own->part)
            std::memcpy(partition_out, own->part, sizeof(u32) * own->n);
        else if (!own->partition.empty())
            std::memcpy(partition_out, own->partition.data(), sizeof(u32) * own->partition.size());
    }
    part_pool().put(own->part);
    own->part = nullptr;
    if (st) rbsc_set_error("%s", own->error.c_str());
    return st ? st : RBSC_OK;
}

extern "C" int rbsc_server_wait_labels(rbsc_server* s, uint64_t ticket, double timeout_s,
                                       rbsc_request_result* out, uint32_t** labels) {
    if (!labels) {
        rbsc_set_error("rbsc_server_wait_labels: null argument");
        return RBSC_ERR_INVALID;
    }
    *labels = nullptr;
    std::unique_ptr<Req> own;
    const int rc = take_result(s, ticket, timeout_s, out, &own, "rbsc_server_wait_labels");
    if (rc) return rc;
    const int st = own->status;
    if (st) {
        part_pool().put(own->part);
        own->part = nullptr;
        rbsc_set_error("%s", own->error.c_str());
        return st;
    }
    if (own->part) {
        *labels = own->part;  // the caller's until rbsc_server_labels_free
        own->part = nullptr;
    } else if (!own->partition.empty()) {
        u32* p = part_pool().get(own->partition.size());
        if (!p) {
            rbsc_set_error("rbsc_server_wait_labels: page-locked memory exhausted");
            return RBSC_ERR_NOMEM;
        }
        std::memcpy(p, own->partition.data(), sizeof(u32) * own->partition.size());
        *labels = p;
    }
    return RBSC_OK;
}

extern "C" int rbsc_server_labels_free(uint32_t* labels) {
    if (!part_pool().put(labels)) {
        rbsc_set_error("rbsc_server_labels_free: not a buffer from rbsc_server_wait_labels");
        return RBSC_ERR_INVALID;
    }
    return RBSC_OK;
}

extern "C" int rbsc_server_get_stats(rbsc_server* s, rbsc_server_stats* out) {
    if (!s || !out) {
        rbsc_set_error("rbsc_server_get_stats: null argument");
        return RBSC_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    rbsc_server_stats o{};
    o.submitted = s->submitted;
    o.completed = s->completed;
    o.failed = s->failed;
    o.batches = s->batches;
    o.batched_graphs = s->batched_graphs;
    o.admitted_bytes = s->admitted;
    o.peak_admitted_bytes = s->peak_admitted;
    o.budget = s->opts.hbm_budget;
    o.in_flight = s->in_flight;
    *out = o;
    return RBSC_OK;
}
