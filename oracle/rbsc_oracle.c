/*
 * rbsc_oracle.c -- CPU restatement of the RBSC refinement round.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: only tests, the smoke check and the CPU-baseline leg of the bench may
 * load it.  Nothing in the product library calls or links it.
 *
 * Every function restates the reference algorithm with plain loops and
 * qsort, citing the reference file:line it follows.  How each function is
 * pinned (beyond restating the reference) is stated next to it and in the
 * tests under tests/test_oracle*.py:
 *
 *   oracle_randomize       pinned by closed-form inversion (bijectivity) and
 *                          known-answer vectors of the published SplitMix64
 *                          finalizer / xoshiro** scrambler steps.
 *   oracle_rbsc_refine     pinned by the exact hash-free refiner (final
 *                          partition), by the compiled reference rsc and
 *                          Valmari block counts on committed fixtures, and by
 *                          closed-form graphs (cycle, path, star, regular).
 *                          Its per-round colour vectors are pinned only by
 *                          self-consistency: the CUDA binary cannot be built
 *                          without nvcc/CUB (parity unpinned per round).
 *   oracle_rbsc_multibatch pinned by block counts on fixtures that
 *                          reduce to the single-batch case (B >= E) and by
 *                          hand-worked small multi-batch cases; the CUDA
 *                          multi-batch `done`/`nodes` cannot be run here
 *                          (parity unpinned against the binary).
 *   oracle_rsc_refine      pinned against the compiled reference rsc binary
 *                          (oracle/_ref/rsc) on every fixture.
 *   oracle_exact_refine    pinned by brute force / closed forms and against
 *                          the compiled Valmari tool (oracle/_ref/part_ref).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef uint32_t u32;
typedef uint64_t u64;

/* rbsc.cu:65-71 and rsc.cpp:63-69: SplitMix64 finalizer followed by the
 * xoshiro** output scrambler (rotl 7, times 9), every step a bijection. */
u64 oracle_randomize(u64 v) {
    v = v + 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    v = (v ^ (v >> 31)) * 5ULL;
    return ((v << 7) | (v >> 57)) * 9ULL;
}

/* ---------------------------------------------------------------------- */
/* Sorting helpers: a (sig, colour, index) triple ordered lexicographically.
 * rbsc.cu:386-390 sorts stably by the low 32 colour bits, then stably by the
 * 64-bit signature, which yields lexicographic (sig, colour, index) order. */
typedef struct { u64 sig; u64 col; u32 idx; } trip_t;

static int trip_cmp(const void* a, const void* b) {
    const trip_t* x = (const trip_t*)a;
    const trip_t* y = (const trip_t*)b;
    if (x->sig != y->sig) return x->sig < y->sig ? -1 : 1;
    u32 cx = (u32)x->col, cy = (u32)y->col;
    if (cx != cy) return cx < cy ? -1 : 1;
    if (x->idx != y->idx) return x->idx < y->idx ? -1 : 1;
    return 0;
}

/* One rank step shared by the rbsc restatements: new colour = dense rank of
 * (sig, colour) ascending; returns the number of distinct pairs.
 * rbsc.cu:384-402 (create_index, the two SortPairs, gather,
 * compute_thruths_lex, InclusiveSum, reorder, K = cumsum[N-1] + 1). */
static u32 rank_pairs(u32 n, const u64* sig, u64* col, trip_t* tmp) {
    if (n == 0) return 0;
    for (u32 i = 0; i < n; ++i) { tmp[i].sig = sig[i]; tmp[i].col = col[i]; tmp[i].idx = i; }
    qsort(tmp, n, sizeof(trip_t), trip_cmp);
    u32 rank = 0;
    u64 psig = tmp[0].sig, pcol = tmp[0].col;
    for (u32 i = 0; i < n; ++i) {
        /* rbsc.cu:121 compares the full 64-bit colour words. */
        if (i && (psig != tmp[i].sig || pcol != tmp[i].col)) ++rank;
        psig = tmp[i].sig;
        pcol = tmp[i].col;
        col[tmp[i].idx] = rank;
    }
    return rank + 1;
}

/* ---------------------------------------------------------------------- */
/* Single-batch rbsc refinement (rbsc.cu:272-445 with batches == 1):
 *   c0(u) = N+1 if u has a line, 0 otherwise          (init_partition 51-63)
 *   sig_r(u) = sum_e randomize(c_r(v_e) + r) * wt_e     (set_values 80-87 and
 *                                                       ReduceByKey 366-377,
 *                                                       mend_partition 104-116)
 *   c_{r+1}(u) = dense rank of (sig_r(u), c_r(u))       (384-396)
 *   K starts at 1 (339) and the loop stops at the first round whose K equals
 *   the previous one (425).  The salt is the round number because the
 *   collision loop 404-419 runs exactly once (randomize is a bijection).
 * When `init` is non-NULL the initial colour is init[u] (an extension used by
 * the initial-partition workload; the reference has no such input).
 * src must be grouped (equal values contiguous), as ReduceByKey requires.
 * wt may be NULL (all weights 1).
 * trace, when non-NULL, receives c_{r+1} for r = 0 .. trace_rounds-1
 * (row-major, N u32 per round).  Returns 0, or -1 on allocation failure. */
int oracle_rbsc_refine(u32 N, u64 E, const u32* src, const u64* wt, const u32* dst,
                       const u32* init, u32 max_rounds,
                       u32* colors_out, u32* K_out, u32* rounds_out,
                       u32* trace, u32 trace_rounds) {
    u64* col = (u64*)malloc(sizeof(u64) * (N ? N : 1));
    u64* sig = (u64*)malloc(sizeof(u64) * (N ? N : 1));
    trip_t* tmp = (trip_t*)malloc(sizeof(trip_t) * (N ? N : 1));
    unsigned char* has = (unsigned char*)calloc(N ? N : 1, 1);
    if (!col || !sig || !tmp || !has) { free(col); free(sig); free(tmp); free(has); return -1; }
    for (u64 e = 0; e < E; ++e) has[src[e]] = 1;
    for (u32 i = 0; i < N; ++i) col[i] = init ? (u64)init[i] : (has[i] ? (u64)N + 1 : 0);
    u32 K = 1, Knew = 1, it = 0;
    do {
        K = Knew;
        memset(sig, 0, sizeof(u64) * (N ? N : 1));
        for (u64 e = 0; e < E; ++e) {
            u64 w = wt ? wt[e] : 1ULL;
            sig[src[e]] += oracle_randomize(col[dst[e]] + (u64)it) * w;
        }
        Knew = rank_pairs(N, sig, col, tmp);
        if (trace && it < trace_rounds)
            for (u32 i = 0; i < N; ++i) trace[(u64)it * N + i] = (u32)col[i];
        ++it;
        if (max_rounds && it >= max_rounds) break;
    } while (K != Knew);
    for (u32 i = 0; i < N; ++i) colors_out[i] = (u32)col[i];
    if (K_out) *K_out = Knew;
    if (rounds_out) *rounds_out = it;
    free(col); free(sig); free(tmp); free(has);
    return 0;
}

/* First-occurrence relabel over node ids 0..N-1 (rbsc.cu:431-445): the
 * canonical partition the parity checks compare.  Returns #blocks. */
u32 oracle_canonicalize(u32 N, const u32* colors, u32* out) {
    u32 maxc = 0;
    for (u32 i = 0; i < N; ++i) if (colors[i] > maxc) maxc = colors[i];
    u32* map = (u32*)malloc(sizeof(u32) * ((u64)maxc + 1));
    if (!map) return 0;
    for (u64 c = 0; c <= maxc; ++c) map[c] = UINT32_MAX;
    u32 next = 0;
    for (u32 i = 0; i < N; ++i) {
        u32 c = colors[i];
        if (map[c] == UINT32_MAX) map[c] = next++;
        out[i] = map[c];
    }
    free(map);
    return next;
}

/* ---------------------------------------------------------------------- */
/* rsc.cpp:103-150 restated: c0 = 1, K starts at N, the new colour is the
 * dense rank of the signature alone (rsc.cpp:88-97 ranks s_z only), stop
 * when K repeats.  Weights are u32 as in rsc.cpp:46-60.  Dangling nodes keep
 * signature 0 (rsc.cpp:118-121, spmv only touches grouped sources). */
typedef struct { u64 sig; u64 col; u32 idx; } rsc_t;
static int rsc_cmp(const void* a, const void* b) {
    const rsc_t* x = (const rsc_t*)a; const rsc_t* y = (const rsc_t*)b;
    if (x->sig != y->sig) return x->sig < y->sig ? -1 : 1;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}
int oracle_rsc_refine(u32 N, u64 E, const u32* src, const u32* wt, const u32* dst,
                      u32* colors_out, u32* K_out, u32* rounds_out) {
    if (N == 0) { if (K_out) *K_out = 0; if (rounds_out) *rounds_out = 0; return 0; }
    u64* w = (u64*)malloc(sizeof(u64) * N);
    u64* z = (u64*)malloc(sizeof(u64) * N);
    rsc_t* t = (rsc_t*)malloc(sizeof(rsc_t) * N);
    if (!w || !z || !t) { free(w); free(z); free(t); return -1; }
    for (u32 i = 0; i < N; ++i) { w[i] = 1; z[i] = 0; }
    u32 K, Knew = N; u64 it = 0;
    do {
        K = Knew;
        for (u64 e = 0; e < E; ++e) {       /* spmv, rsc.cpp:71-78 */
            if (!e || src[e] != src[e - 1]) z[src[e]] = 0;
            z[src[e]] += oracle_randomize(w[dst[e]] + it) * (u64)(wt ? wt[e] : 1u);
        }
        for (u32 i = 0; i < N; ++i) { t[i].sig = z[i]; t[i].col = w[i]; t[i].idx = i; }
        qsort(t, N, sizeof(rsc_t), rsc_cmp);
        u64 cur = 0, prec = t[0].sig;     /* sum, rsc.cpp:88-97 */
        t[0].col = 0;
        for (u32 i = 1; i < N; ++i) {
            if (t[i].sig != prec) ++cur;
            prec = t[i].sig;
            t[i].col = cur;
        }
        Knew = (u32)(cur + 1);
        for (u32 i = 0; i < N; ++i) w[t[i].idx] = t[i].col;
        ++it;
    } while (K != Knew);
    for (u32 i = 0; i < N; ++i) colors_out[i] = (u32)w[i];
    if (K_out) *K_out = Knew;
    if (rounds_out) *rounds_out = (u32)it;
    free(w); free(z); free(t);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Exact, hash-free refiner: the signature of u is the sorted list of
 * (colour of head, summed weight) over u's lines, and u's new colour is the
 * rank of (old colour, signature) compared exactly.  Iterates to the
 * fixpoint, so the result is the coarsest partition refining the initial one
 * in which all members of a block send equal total weight into every block
 * (the lumping definition of MDP_minimize.cc:55-65).  init NULL means the
 * all-equal initial colouring.  Intended for small graphs (tests). */
typedef struct { u32 col; u64 w; } cw_t;
static int cw_cmp(const void* a, const void* b) {
    const cw_t* x = (const cw_t*)a; const cw_t* y = (const cw_t*)b;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return 0;
}
static int u32_cmp(const void* a, const void* b) {
    u32 x = *(const u32*)a, y = *(const u32*)b;
    return x < y ? -1 : (x > y);
}
/* qsort has no context argument: the exact comparator reads these. */
static const cw_t* g_sig;
static const u64* g_off;
static const u32* g_col;
static int exact_key_cmp(u32 x, u32 y) {
    if (g_col[x] != g_col[y]) return g_col[x] < g_col[y] ? -1 : 1;
    u64 lx = g_off[x + 1] - g_off[x], ly = g_off[y + 1] - g_off[y];
    u64 l = lx < ly ? lx : ly;
    for (u64 k = 0; k < l; ++k) {
        const cw_t* p = &g_sig[g_off[x] + k];
        const cw_t* q = &g_sig[g_off[y] + k];
        if (p->col != q->col) return p->col < q->col ? -1 : 1;
        if (p->w != q->w) return p->w < q->w ? -1 : 1;
    }
    if (lx != ly) return lx < ly ? -1 : 1;
    return 0;
}
static int exact_cmp(const void* a, const void* b) {
    u32 x = *(const u32*)a, y = *(const u32*)b;
    int c = exact_key_cmp(x, y);
    if (c) return c;
    return x < y ? -1 : (x > y);
}
int oracle_exact_refine(u32 N, u64 E, const u32* src, const u64* wt, const u32* dst,
                        const u32* init, u32* colors_out, u32* K_out, u32* rounds_out) {
    if (N == 0) { if (K_out) *K_out = 0; if (rounds_out) *rounds_out = 0; return 0; }
    u64* off = (u64*)calloc((u64)N + 1, sizeof(u64));
    u64* pos = (u64*)malloc(sizeof(u64) * ((u64)N + 1));
    u64* poff = (u64*)malloc(sizeof(u64) * ((u64)N + 1));
    cw_t* sig = (cw_t*)malloc(sizeof(cw_t) * (E ? E : 1));
    cw_t* packed = (cw_t*)malloc(sizeof(cw_t) * (E ? E : 1));
    u32* col = (u32*)malloc(sizeof(u32) * N);
    u32* ncol = (u32*)malloc(sizeof(u32) * N);
    u32* order = (u32*)malloc(sizeof(u32) * N);
    if (!off || !pos || !poff || !sig || !packed || !col || !ncol || !order) {
        free(off); free(pos); free(poff); free(sig); free(packed);
        free(col); free(ncol); free(order); return -1;
    }
    for (u64 e = 0; e < E; ++e) off[src[e] + 1]++;
    for (u32 i = 0; i < N; ++i) off[i + 1] += off[i];
    for (u32 i = 0; i < N; ++i) col[i] = init ? init[i] : 0;
    /* K_prev = number of distinct initial colours */
    memcpy(order, col, sizeof(u32) * N);
    qsort(order, N, sizeof(u32), u32_cmp);
    u32 Kprev = 1;
    for (u32 i = 1; i < N; ++i) if (order[i] != order[i - 1]) ++Kprev;
    u32 Kr = Kprev, rounds = 0;
    for (;;) {
        memcpy(pos, off, sizeof(u64) * ((u64)N + 1));
        for (u64 e = 0; e < E; ++e) {
            cw_t c; c.col = col[dst[e]]; c.w = wt ? wt[e] : 1ULL;
            sig[pos[src[e]]++] = c;
        }
        poff[0] = 0;
        for (u32 u = 0; u < N; ++u) {
            u64 b = off[u], n = off[u + 1] - off[u];
            if (n > 1) qsort(sig + b, n, sizeof(cw_t), cw_cmp);
            u64 w = poff[u];
            for (u64 k = b; k < b + n; ++k) {   /* merge equal colours */
                if (w > poff[u] && packed[w - 1].col == sig[k].col) packed[w - 1].w += sig[k].w;
                else packed[w++] = sig[k];
            }
            poff[u + 1] = w;
        }
        g_sig = packed; g_off = poff; g_col = col;
        for (u32 i = 0; i < N; ++i) order[i] = i;
        qsort(order, N, sizeof(u32), exact_cmp);
        u32 r = 0;
        for (u32 i = 0; i < N; ++i) {
            if (i && exact_key_cmp(order[i - 1], order[i]) != 0) ++r;
            ncol[order[i]] = r;
        }
        Kr = r + 1;
        ++rounds;
        memcpy(col, ncol, sizeof(u32) * N);
        if (Kr == Kprev) break;
        Kprev = Kr;
    }
    for (u32 i = 0; i < N; ++i) colors_out[i] = col[i];
    if (K_out) *K_out = Kr;
    if (rounds_out) *rounds_out = rounds;
    free(off); free(pos); free(poff); free(sig); free(packed);
    free(col); free(ncol); free(order);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Multi-batch quotient mode with one device (rbsc.cu:274-487, the OpenMP
 * region with max_gpus = 1 processes batches in order).  Per pass:
 *   batches = ceil(E / B) in double (278); global labels w[] = node_n (280);
 *   per batch: trim [start, end) to whole source groups (315-328), mark
 *   nodes IN / OFF / DANGLING (310-336), refine the batch to its own fixpoint
 *   with OFF nodes as singletons of colour i+1 and signature UINT64_MAX
 *   (51-63, 104-116, 354-425), merge local labels by first occurrence into
 *   the global counter (431-445);
 *   if batches > 1: fresh labels for unprocessed nodes (457-461) and
 *   in-place contraction to the quotient graph, parallel edges merged by
 *   summed weight (463-485);
 * repeat while batches > 1 && node_n > new_node_n (487); the `aggressive`
 * variant also requires batches > ceilf(new_edge_n / (float)B)
 * (rbsc_aggressive.cu:487).  done = (batches == 1) (493).
 * partition_out (nullable): the first-occurrence canonical form of each
 * original node's class, the composition of its labels over the passes
 * (the reference keeps only the node count; the composition is what a
 * caller of the quotient would read back).
 * Inputs are copied; weights are u64 (rbsc.cu:173).  Returns 0 or -1. */
#define MB_OFF 0
#define MB_IN 1
#define MB_DANG 2
int oracle_rbsc_multibatch(u32 N0, u64 E0, const u32* src0, const u64* wt0, const u32* dst0,
                           u64 max_batch, int aggressive,
                           u32* nodes_out, u32* done_out, u32* passes_out,
                           u32* partition_out) {
    u32* src = (u32*)malloc(sizeof(u32) * (E0 ? E0 : 1));
    u32* dst = (u32*)malloc(sizeof(u32) * (E0 ? E0 : 1));
    u64* wt = (u64*)malloc(sizeof(u64) * (E0 ? E0 : 1));
    u32 NN = N0 ? N0 : 1;
    u32* w = (u32*)malloc(sizeof(u32) * NN);
    unsigned char* mask = (unsigned char*)malloc(NN);
    u64* considered = (u64*)malloc(sizeof(u64) * NN);
    u64* windex = (u64*)malloc(sizeof(u64) * NN);
    u32* label_map = (u32*)malloc(sizeof(u32) * ((u64)NN + 2));
    u64* col = (u64*)malloc(sizeof(u64) * NN);
    u64* sig = (u64*)malloc(sizeof(u64) * NN);
    trip_t* tmp = (trip_t*)malloc(sizeof(trip_t) * NN);
    u32* comp = (u32*)malloc(sizeof(u32) * NN);
    if (!src || !dst || !wt || !w || !mask || !considered || !windex || !label_map ||
        !col || !sig || !tmp || !comp) {
        free(src); free(dst); free(wt); free(w); free(mask); free(considered);
        free(windex); free(label_map); free(col); free(sig); free(tmp); free(comp);
        return -1;
    }
    for (u64 e = 0; e < E0; ++e) { src[e] = src0[e]; dst[e] = dst0[e]; wt[e] = wt0 ? wt0[e] : 1ULL; }
    for (u32 i = 0; i < N0; ++i) comp[i] = i;
    u32 node_n, new_node_n = N0;
    u64 edge_n, new_edge_n = E0, batches = 1;
    u32 passes = 0;
    const u64 NOT_INIT = UINT64_MAX;
    do {
        node_n = new_node_n;
        edge_n = new_edge_n;
        new_node_n = 0;
        batches = (u64)ceil((double)edge_n / (double)max_batch);
        ++passes;
        for (u32 i = 0; i < node_n; ++i) w[i] = node_n;
        for (u64 batch = 0; batch < batches; ++batch) {
            u64 bs = batch * max_batch;
            u64 be = bs + max_batch < edge_n ? bs + max_batch : edge_n;
            for (u32 i = 0; i < node_n; ++i) mask[i] = batch == 0 ? MB_DANG : MB_OFF;
            u64 nbs = NOT_INIT, nbe = NOT_INIT;
            for (u64 i = bs; i < be; ++i) {
                if ((!i || src[i - 1] != src[i]) && nbs == NOT_INIT) nbs = i;
                if (i && nbs != NOT_INIT && src[i - 1] != src[i]) nbe = i;
                mask[src[i]] = MB_IN;
            }
            if (be == edge_n || src[be] != src[be - 1]) nbe = be;
            bs = nbs; be = nbe;
            if (bs == NOT_INIT || be == NOT_INIT) continue;
            for (u64 i = 0; i < bs; ++i) mask[src[i]] = MB_OFF;
            for (u64 i = be; i < edge_n; ++i) mask[src[i]] = MB_OFF;
            u32 batch_node_n = 0;
            for (u32 i = 0; i < node_n; ++i) if (mask[i]) ++batch_node_n;
            u32 off_n = node_n - batch_node_n;
            u32 K = off_n + 1, Knew = off_n + 1;
            for (u32 i = 0; i < node_n; ++i)
                col[i] = mask[i] == MB_IN ? (u64)node_n + 1 : (mask[i] == MB_DANG ? 0 : (u64)i + 1);
            u64 it = 0;
            do {
                K = Knew;
                for (u32 i = 0; i < node_n; ++i) sig[i] = 0;
                for (u64 e = bs; e < be; ++e)
                    sig[src[e]] += oracle_randomize(col[dst[e]] + it) * wt[e];
                for (u32 i = 0; i < node_n; ++i) {
                    if (mask[i] == MB_OFF) sig[i] = UINT64_MAX;
                    else if (mask[i] == MB_DANG) sig[i] = 0;
                }
                Knew = rank_pairs(node_n, sig, col, tmp);
                ++it;
            } while (K != Knew);
            for (u32 i = 0; i < node_n; ++i) label_map[i] = node_n;
            for (u32 i = 0; i < node_n; ++i) {
                if (mask[i]) {
                    u32 local = (u32)col[i];
                    if (label_map[local] == node_n) label_map[local] = new_node_n++;
                    w[i] = label_map[local];
                }
            }
        }
        if (batches > 1) {
            new_edge_n = 0;
            u64 first_edge_i = 0;
            u32 current_node = 0;
            int counting = 0;
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
        /* the class of every original node: its label in this pass */
        for (u32 i = 0; i < N0; ++i) comp[i] = comp[i] < node_n ? w[comp[i]] : comp[i];
        if (!(batches > 1 && node_n > new_node_n)) break;
        if (aggressive) {
            float q = ceilf((float)new_edge_n / (float)max_batch);
            if (!((float)batches > q)) break;
        }
    } while (1);
    if (nodes_out) *nodes_out = new_node_n;
    if (done_out) *done_out = batches == 1;
    if (passes_out) *passes_out = passes;
    int rc = 0;
    if (partition_out && N0 && oracle_canonicalize(N0, comp, partition_out) == 0) rc = -1;
    free(src); free(dst); free(wt); free(w); free(mask); free(considered);
    free(windex); free(label_map); free(col); free(sig); free(tmp); free(comp);
    return rc;
}
