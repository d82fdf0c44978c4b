/*
 * lump.c -- exact coarsest lumping by splitter refinement.
 *
 * TEST INFRASTRUCTURE ONLY: a second exact refiner beside
 * oracle_exact_refine (rbsc_oracle.c) and a CPU baseline of the kind of the
 * reference's Valmari tool (src/part_ref/MDPmin.cc), which is compiled from
 * the reference's sources only where that tree is present.  Nothing in the
 * product library calls or links it.
 *
 * Result: the coarsest partition refining the initial colouring in which,
 * for every block B, all members of a block send equal total weight into B
 * (the lumping definition of MDP_minimize.cc:55-65; lines `u w v` go from u
 * to v, so weight into B is the sum of w over u's lines whose head v lies in
 * B, as in rbsc.cu:80-87).  Weights add in Z/2^64, as in the hash-free
 * refiner of rbsc_oracle.c.
 *
 * Method: splitter refinement with a worklist (Hopcroft's "all but the
 * largest" rule, as used for Markov chain lumping by Valmari and
 * Franceschinis, "Simple O(m log n) Time Markov Chain Lumping", TACAS 2010,
 * the paper MDP_minimize.cc:31-34 cites).  Blocks are contiguous ranges of
 * an element array.  Every initial block starts in the worklist.  Taking a
 * splitter S, the weight w(u, S) of every source u of a line into S is
 * summed; each block holding such sources is split by that weight (sources
 * whose sum is 0 stay with the block's other members).  When a block B
 * splits, all its parts enter the worklist if B was waiting there;
 * otherwise all but the largest part do -- the largest part's weights
 * follow from B's (already uniform per block, B having been a splitter or
 * being waiting) minus the others'.  So every node lies in O(log n)
 * splitters and the line scans are O(m log n); the parts of a block are
 * found by sorting its touched sources by weight, O(t log t) for t touched
 * sources.  The worklist ends empty exactly when no splitter splits a
 * block: the partition is then stable, and it is the coarsest, since a
 * split only ever separates nodes with different weight into a block of a
 * coarser stable partition's refinement.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint32_t u32;
typedef uint64_t u64;

typedef struct {
    u32 blk;
    u32 node;
    u64 w;
} touch_t;

/* by block, then weight, then node */
static int touch_cmp(const void* a, const void* b) {
    const touch_t* x = (const touch_t*)a;
    const touch_t* y = (const touch_t*)b;
    if (x->blk != y->blk) return x->blk < y->blk ? -1 : 1;
    if (x->w != y->w) return x->w < y->w ? -1 : 1;
    return x->node < y->node ? -1 : (x->node > y->node);
}

static int u32_cmp_l(const void* a, const void* b) {
    u32 x = *(const u32*)a, y = *(const u32*)b;
    return x < y ? -1 : (x > y);
}

/* Refines N nodes with E lines (src[e] -> dst[e], weight wt[e] or 1; any
 * line order) from init (NULL: one block).  labels_out[v] = the block of v,
 * numbered by first occurrence in node order; *K_out = the block count.
 * Returns 0, or -1 when memory runs out. */
int oracle_lump_refine(u32 N, u64 E, const u32* src, const u64* wt, const u32* dst,
                       const u32* init, u32* labels_out, u32* K_out) {
    if (N == 0) {
        if (K_out) *K_out = 0;
        return 0;
    }
    u64* in_off = (u64*)calloc((u64)N + 1, sizeof(u64));
    u32* in_src = (u32*)malloc(sizeof(u32) * (E ? E : 1));
    u64* in_w = (u64*)malloc(sizeof(u64) * (E ? E : 1));
    u32* elems = (u32*)malloc(sizeof(u32) * N);   /* nodes, blocks contiguous */
    u32* loc = (u32*)malloc(sizeof(u32) * N);     /* position of a node in elems */
    u32* blk = (u32*)malloc(sizeof(u32) * N);     /* block of a node */
    u32* first = (u32*)malloc(sizeof(u32) * N);   /* block b: elems[first[b], end[b]) */
    u32* end = (u32*)malloc(sizeof(u32) * N);
    unsigned char* waiting = (unsigned char*)calloc(N, 1);
    u32* work = (u32*)malloc(sizeof(u32) * N);    /* worklist (a stack of blocks) */
    u64* acc = (u64*)calloc(N, sizeof(u64));      /* w(u, S) of the current splitter */
    unsigned char* hit = (unsigned char*)calloc(N, 1);
    u32* touched = (u32*)malloc(sizeof(u32) * N); /* sources of lines into S */
    touch_t* part = (touch_t*)malloc(sizeof(touch_t) * N);
    u32* lab = (u32*)malloc(sizeof(u32) * N);
    int rc = -1;
    if (!in_off || !in_src || !in_w || !elems || !loc || !blk || !first || !end || !waiting ||
        !work || !acc || !hit || !touched || !part || !lab)
        goto out;
    /* lines by head */
    for (u64 e = 0; e < E; ++e) in_off[dst[e] + 1]++;
    for (u32 v = 0; v < N; ++v) in_off[v + 1] += in_off[v];
    {
        u64* pos = (u64*)malloc(sizeof(u64) * ((u64)N + 1));
        if (!pos) goto out;
        memcpy(pos, in_off, sizeof(u64) * ((u64)N + 1));
        for (u64 e = 0; e < E; ++e) {
            const u64 p = pos[dst[e]]++;
            in_src[p] = src[e];
            in_w[p] = wt ? wt[e] : 1ULL;
        }
        free(pos);
    }
    /* initial blocks: nodes ordered by initial colour */
    u32 nb = 0, nw = 0;
    if (init) {
        /* counting sort by colour value through a sorted copy of the colours */
        u32* cols = (u32*)malloc(sizeof(u32) * N);
        if (!cols) goto out;
        memcpy(cols, init, sizeof(u32) * N);
        qsort(cols, N, sizeof(u32), u32_cmp_l);
        u32 distinct = 0;
        for (u32 i = 0; i < N; ++i)
            if (i == 0 || cols[i] != cols[i - 1]) cols[distinct++] = cols[i];
        /* block of a node = rank of its colour among the distinct ones */
        u32* cnt = (u32*)calloc((u64)distinct + 1, sizeof(u32));
        if (!cnt) { free(cols); goto out; }
        for (u32 v = 0; v < N; ++v) {
            u32 lo = 0, hi = distinct - 1;
            while (lo < hi) {
                const u32 mid = lo + (hi - lo) / 2;
                if (cols[mid] < init[v]) lo = mid + 1; else hi = mid;
            }
            blk[v] = lo;
            cnt[lo + 1]++;
        }
        for (u32 b = 0; b < distinct; ++b) cnt[b + 1] += cnt[b];
        for (u32 b = 0; b < distinct; ++b) { first[b] = cnt[b]; end[b] = cnt[b + 1]; }
        for (u32 v = 0; v < N; ++v) {
            const u32 p = cnt[blk[v]]++;
            elems[p] = v;
            loc[v] = p;
        }
        nb = distinct;
        free(cnt);
        free(cols);
    } else {
        for (u32 v = 0; v < N; ++v) { elems[v] = v; loc[v] = v; blk[v] = 0; }
        first[0] = 0;
        end[0] = N;
        nb = 1;
    }
    for (u32 b = 0; b < nb; ++b) { work[nw++] = b; waiting[b] = 1; }

    while (nw) {
        const u32 S = work[--nw];
        waiting[S] = 0;
        /* w(u, S) for every source u of a line into S */
        u32 nt = 0;
        const u32 s0 = first[S], s1 = end[S];
        for (u32 i = s0; i < s1; ++i) {
            const u32 v = elems[i];
            for (u64 p = in_off[v]; p < in_off[v + 1]; ++p) {
                const u32 u = in_src[p];
                if (!hit[u]) { hit[u] = 1; acc[u] = 0; touched[nt++] = u; }
                acc[u] += in_w[p];
            }
        }
        /* touched sources with a nonzero sum, by block and weight */
        u32 k_all = 0;
        for (u32 i = 0; i < nt; ++i) {
            const u32 u = touched[i];
            hit[u] = 0;
            if (acc[u] == 0) continue;  /* same as a node without lines into S */
            part[k_all].blk = blk[u];
            part[k_all].node = u;
            part[k_all].w = acc[u];
            ++k_all;
        }
        qsort(part, k_all, sizeof(touch_t), touch_cmp);
        for (u32 r0 = 0, r1; r0 < k_all; r0 = r1) {
            const u32 B = part[r0].blk;
            r1 = r0;
            while (r1 < k_all && part[r1].blk == B) ++r1;
            const touch_t* pt = part + r0;
            const u32 k = r1 - r0, size = end[B] - first[B];
            /* one part already when every member is touched with one weight */
            if (k == size && pt[0].w == pt[k - 1].w) continue;
            /* move the touched sources to the block's end, in weight order;
             * the untouched stay in front and keep the block's number */
            u32 tail = end[B];
            for (u32 j = k; j-- > 0;) {
                const u32 u = pt[j].node;
                --tail;
                const u32 pu = loc[u], other = elems[tail];
                elems[tail] = u; loc[u] = tail;
                elems[pu] = other; loc[other] = pu;
            }
            /* part boundaries: [first[B], tail) untouched (possibly empty), then
             * one run per weight */
            u32 largest_b = B, largest_n = tail - first[B];
            const u32 old_end = end[B];
            u32 g = 0, start = tail;
            u32 keep_first = first[B];
            int b_used = tail > keep_first;  /* B keeps the untouched part */
            if (b_used) end[B] = tail;
            while (g < k) {
                u32 j = g;
                while (j < k && pt[j].w == pt[g].w) ++j;
                const u32 lo = start + g, hi = start + j;
                u32 nbk;
                if (!b_used) { nbk = B; b_used = 1; first[B] = lo; end[B] = hi; }
                else {
                    nbk = nb++;
                    first[nbk] = lo; end[nbk] = hi;
                    for (u32 i = lo; i < hi; ++i) blk[elems[i]] = nbk;
                }
                if (hi - lo > largest_n) { largest_n = hi - lo; largest_b = nbk; }
                g = j;
            }
            /* worklist: every part if B was waiting, else all but the largest */
            const int was_waiting = waiting[B];
            for (u32 i = keep_first; i < old_end;) {
                const u32 b = blk[elems[i]];
                if ((was_waiting || b != largest_b) && !waiting[b]) {
                    waiting[b] = 1;
                    work[nw++] = b;
                }
                i = end[b];
            }
        }
    }
    /* first-occurrence labels */
    {
        for (u32 b = 0; b < nb; ++b) lab[b] = 0xFFFFFFFFu;
        u32 K = 0;
        for (u32 v = 0; v < N; ++v) {
            if (lab[blk[v]] == 0xFFFFFFFFu) lab[blk[v]] = K++;
            labels_out[v] = lab[blk[v]];
        }
        if (K_out) *K_out = K;
    }
    rc = 0;
out:
    free(in_off); free(in_src); free(in_w); free(elems); free(loc); free(blk); free(first);
    free(end); free(waiting); free(work); free(acc); free(hit); free(touched); free(part);
    free(lab);
    return rc;
}
