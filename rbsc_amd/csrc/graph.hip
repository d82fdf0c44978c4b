// graph.hip -- building the resident graph layout in HBM.
//
// A graph arrives "raw": rows with >= 1 line in input order (row_node_raw[k]
// is the node of row k, lines row_ptr_raw[k] .. row_ptr_raw[k+1]) and line
// heads as original node ids.  dgraph_build turns it into the layout the
// round kernels read:
//
//   * internal node ids: nodes with lines first, by in-degree descending
//     (ties by original id; without initial colours a stored row's node also
//     counts the in-degree of the nodes whose rows repeat it, since its colour
//     is gathered for them), then the nodes without lines in original order.
//     Colours depend only on (signature, colour) pairs, never on node ids, so
//     every round's colour of every node is unchanged by the renumbering; it
//     makes the hashed colours of the most-gathered heads a compact prefix
//     that stays in L2 / Infinity Cache, and lets the node-side kernels skip
//     nodes without lines entirely.
//   * rows in internal order (row j = node j), so signature writes are
//     coalesced and no row -> node indirection is read per round.
//   * a row-end bitmap (bit p = line p is the last of its row): the signature
//     kernel finds row boundaries from one byte per 8 lines instead of
//     searching row offsets.
//   * pi[orig] = internal id, to map colours back.
//   * identical rows (same heads and weights in the same order) are stored
//     once: a signature depends only on a row's lines, so every round gives
//     identical rows identical signatures.  Rows that repeat an earlier
//     stored row get internal ids after the stored rows ([R, U)) and no lines;
//     dup_rep[j - R] names the stored row whose signature they copy.
#include "internal.h"
#include "scan.h"
#include "tiles.h"
#include <algorithm>

namespace {
constexpr int BLK = 256;
inline unsigned grid_for(u64 n) { return (unsigned)((n + BLK - 1) / BLK); }

__global__ void zero_u32(u32* p, u64 n) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void fill_ones_kernel(u64* p, u64 n) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 1;
}

// In-degrees of the lines e = k * stride (every line for stride 1).  They
// only order the internal ids (most-gathered colours first, for cache
// locality and the hot / dense / warm classes of the binned layout); colours
// are dense ranks of (signature, colour) pairs and do not depend on the
// numbering.  With RawGraph::sample_indeg, graphs of 2^28 lines and more
// count every stride-th line, a power of two leaving 2^27 to 2^28 samples
// (every 8th line of the 2^31-line R-MAT: 11 instead of 88 ms of random
// atomics into N words).  Measured on that graph, the sample's order moves
// 5.3 M lines from the hot to the cold class and costs the binned round
// 0.54 GB more gather traffic, so resident graphs count every line.
// Lines k * stride in [lo, hi) (k from k0); heads not below N (malformed
// input, refused by the build afterwards) are skipped.
__global__ void indeg_kernel(const u32* __restrict__ dst, u64 k0, u64 hi, u64 stride, u64 N,
                             u32* __restrict__ indeg) {
    const u64 k = k0 + (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 e = k * stride;
    if (e < hi) {
        const u32 d = dst[e];
        if (d < N) atomicAdd(&indeg[d], 1u);
    }
}
__global__ void dst_range_check_kernel(const u32* __restrict__ dst, u64 lo, u64 hi, u64 N,
                                       u32* __restrict__ err) {
    const u64 e = lo + (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < hi && dst[e] >= N) atomicOr(err, 1u);
}
// Sampled lines: about 2^27 of them (every line below 2^28 lines).
inline u64 indeg_stride(u64 E) {
    u64 s = 1;
    while ((E >> 27) >= 2 * s) s *= 2;
    return s;
}

// keys: (~in-degree, original id) so an ascending sort gives in-degree
// descending; vals: the raw row index.  Also marks nodes that have rows.
// With rep_raw (identical rows found), rows that repeat another go after
// the stored rows: key bit 63 = duplicate, in-degree clamped to 31 bits.
__global__ void order_keys_kernel(const u32* __restrict__ row_node, u64 U,
                                  const u32* __restrict__ indeg, const u32* __restrict__ rep_raw,
                                  u64* __restrict__ keys, u64* __restrict__ vals,
                                  u64* __restrict__ dangling) {
    const u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= U) return;
    const u32 v = row_node[k];
    if (rep_raw) {
        const u64 dup = rep_raw[k] != (u32)k ? 1ULL : 0ULL;
        const u32 d = indeg[v] < 0x7FFFFFFFu ? indeg[v] : 0x7FFFFFFFu;
        keys[k] = (dup << 63) | ((u64)(0x7FFFFFFFu - d) << 32) | v;
    } else {
        keys[k] = ((u64)(0xFFFFFFFFu - indeg[v]) << 32) | v;
    }
    vals[k] = k;
    dangling[v] = 0;
}

__global__ void place_rows_kernel(const u64* __restrict__ keys, const u64* __restrict__ vals,
                                  u64 U, u64 R, const u64* __restrict__ row_ptr_raw,
                                  u32* __restrict__ pi, u64* __restrict__ deg_new,
                                  u32* __restrict__ src_row) {
    const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= U) return;
    const u32 v = (u32)keys[j];
    const u64 k = vals[j];
    pi[v] = (u32)j;
    if (j < R) deg_new[j] = row_ptr_raw[k + 1] - row_ptr_raw[k];
    src_row[j] = (u32)k;
}
// Without initial colours a repeated row's node always has its stored row's
// colour, and the signature layouts gather the stored row for it: the stored
// row's node absorbs the repeat's in-degree, so internal ids order nodes by
// how often their colour is gathered.
__global__ void absorb_indeg_kernel(const u32* __restrict__ row_node,
                                    const u32* __restrict__ rep_raw, u64 U, u32* indeg) {
    const u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= U || rep_raw[k] == (u32)k) return;
    const u32 d = indeg[row_node[k]];  // repeats are never stored rows: read only
    if (d) atomicAdd(&indeg[row_node[rep_raw[k]]], d);
}

// *out += number of rows that repeat another (out zeroed by the caller).
__global__ void count_dups_kernel(const u32* __restrict__ rep_raw, u64 U,
                                  unsigned long long* __restrict__ out) {
    __shared__ unsigned long long c;
    if (threadIdx.x == 0) c = 0;
    __syncthreads();
    unsigned long long m = 0;
    for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < U; k += (u64)gridDim.x * blockDim.x)
        m += rep_raw[k] != (u32)k;
    atomicAdd(&c, m);
    __syncthreads();
    if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

__global__ void place_dangling_kernel(const u64* __restrict__ dangling,
                                      const u64* __restrict__ dscan, u64 N, u64 U,
                                      u32* __restrict__ pi) {
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < N && dangling[v]) pi[v] = (u32)(U + dscan[v]);
}

// Row fingerprints (sums of mixed (position, head, weight) triples, so rows
// with the same lines in another order differ) and lengths; identical rows
// share all three and so sort next to each other.
__device__ __forceinline__ u64 mix_a(u64 x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    return x ^ (x >> 33);
}
// Row fingerprints summed over lines in tiles of LT_LINES lines (rows are
// skewed): each line adds its terms, position-aware through its index in
// its row, to per-row sums in LDS, which go to h1/h2 (zeroed) with one
// atomic per row and tile.  Sums in Z/2^64 do not depend on the order.
__device__ __forceinline__ void fp_terms(u64 d, u64 w, u64 pos, u64* a, u64* b) {
    const u64 j = pos * 0xbf58476d1ce4e5b9ULL;
    *a += mix_a((d << 1) ^ (w * 0x9e3779b97f4a7c15ULL) ^ j);
    *b += mix_a((d + 0x632be59bd9b4e019ULL) ^ (j >> 7)) * (2 * w + 1);
}
template <bool HAS_WT>
__global__ __launch_bounds__(LT_THREADS) void row_hash_tile_kernel(
    const u64* __restrict__ row_ptr, const u32* __restrict__ dst, const u64* __restrict__ wt,
    u64 U, u64 E, const u32* __restrict__ first, unsigned long long* __restrict__ h1,
    unsigned long long* __restrict__ h2, u64 t0) {
    __shared__ LtLds s;
    __shared__ unsigned long long acc[2][LT_LINES + 1];
    const u64 t = t0 + blockIdx.x;
    u32 row[LT_PER];
    const u32 last = lt_rows(row_ptr, U, E, first, t, &s, row);
    const u32 r0 = first[t], nr = last - r0 + 1;
    for (u32 i = threadIdx.x; i < nr; i += LT_THREADS) acc[0][i] = acc[1][i] = 0;
    __syncthreads();
    const u64 b0 = t * LT_LINES + (u64)threadIdx.x * LT_PER;
    u64 a = 0, b = 0;
    u32 cur = row[0];
#pragma unroll
    for (int j = 0; j < LT_PER; ++j) {
        const u64 e = b0 + j;
        if (e < E) {
            if (row[j] != cur) {
                atomicAdd(&acc[0][cur - r0], (unsigned long long)a);
                atomicAdd(&acc[1][cur - r0], (unsigned long long)b);
                a = b = 0;
                cur = row[j];
            }
            fp_terms(dst[e], HAS_WT ? wt[e] : 1, e - s.start[cur - r0], &a, &b);
        }
    }
    if (b0 < E) {
        atomicAdd(&acc[0][cur - r0], (unsigned long long)a);
        atomicAdd(&acc[1][cur - r0], (unsigned long long)b);
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < nr; i += LT_THREADS) {
        atomicAdd(&h1[r0 + i], acc[0][i]);
        atomicAdd(&h2[r0 + i], acc[1][i]);
    }
}
__global__ void iota_u64_kernel(u64* __restrict__ p, u64 n) {
    const u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = k;
}
// After sorting rows by h1: head[i] = 0 if the row at sorted position i has
// exactly the lines (heads and weights, in order) of the row at i - 1.  A
// lane compares a candidate pair of up to 64 lines itself; longer pairs are
// compared by the whole wave in turn, 64 lines a step.
template <bool HAS_WT>
__global__ __launch_bounds__(256) void row_eq_wave_kernel(
    const u64* __restrict__ h1s, const u64* __restrict__ rows, const u64* __restrict__ h2,
    const u64* __restrict__ row_ptr, const u32* __restrict__ dst, const u64* __restrict__ wt,
    u64 U, u64* __restrict__ head) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool cand = false;
    u64 ax = 0, ay = 0, n = 0;
    if (i < U && i > 0 && h1s[i] == h1s[i - 1]) {
        const u64 x = rows[i], y = rows[i - 1];
        ax = row_ptr[x];
        ay = row_ptr[y];
        n = row_ptr[x + 1] - ax;
        cand = h2[x] == h2[y] && n == row_ptr[y + 1] - ay;
    }
    bool eq = cand;
    const bool longrow = cand && n > 64;
    if (cand && !longrow) {
        for (u64 j = 0; j < n && eq; ++j) {
            if (dst[ax + j] != dst[ay + j]) eq = false;
            if (HAS_WT && wt[ax + j] != wt[ay + j]) eq = false;
        }
    }
    u64 todo = __ballot(longrow);
    while (todo) {
        const int l = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        const u64 lx = __shfl(ax, l), ly = __shfl(ay, l), ln = __shfl(n, l);
        bool ok = true;
        for (u64 j = lane; j < ln; j += 64) {
            if (dst[lx + j] != dst[ly + j]) ok = false;
            if (HAS_WT && wt[lx + j] != wt[ly + j]) ok = false;
        }
        const bool same = __ballot(!ok) == 0;
        if (lane == l) eq = same;
    }
    if (i < U) head[i] = eq ? 0 : 1;
}
// Heads of runs of identical rows record their raw row per run.
__global__ void run_leader_kernel(const u64* __restrict__ head, const u64* __restrict__ hscan,
                                  const u64* __restrict__ rows, u64 U, u64* __restrict__ leader) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < U && head[i]) leader[hscan[i]] = rows[i];
}
// rep_raw[row] = raw row it repeats (itself for stored rows).
__global__ void rep_raw_kernel(const u64* __restrict__ head, const u64* __restrict__ hscan,
                               const u64* __restrict__ rows, const u64* __restrict__ leader,
                               u64 U, u32* __restrict__ rep_raw) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < U) rep_raw[rows[i]] = (u32)leader[hscan[i] + head[i] - 1];
}
// Internal ids of the duplicate rows point at their stored row.
__global__ void dup_rep_kernel(const u32* __restrict__ src_row, const u32* __restrict__ rep_raw,
                               const u32* __restrict__ row_node, const u32* __restrict__ pi,
                               u64 R, u64 U, u32* __restrict__ dup_rep) {
    const u64 j = R + (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < U) dup_rep[j - R] = pi[row_node[rep_raw[src_row[j]]]];
}

// The stored rows' lines in their new order, heads renumbered, in tiles of
// LT_LINES output lines: line e of new row r comes from raw line
// e + (row_ptr_raw[src_row[r]] - row_ptr_new[r]).  Each tile also writes its
// 64 words of the row-end bitmap (tiles are word aligned, so no word is
// shared).  With dang_w (no initial colours) it also adds up, per row, the
// weight of the lines into nodes without lines (new ids >= U) into dang_w
// and, with weights, the weight of all lines into row_w (both zeroed; one
// atomic per row and tile).
template <bool HAS_WT>
__global__ __launch_bounds__(LT_THREADS) void copy_rows_tile_kernel(
    const u64* __restrict__ row_ptr_new, const u32* __restrict__ src_row,
    const u64* __restrict__ row_ptr_raw, const u32* __restrict__ dst_raw,
    const u64* __restrict__ wt_raw, const u32* __restrict__ pi, u64 R, u64 E,
    const u32* __restrict__ first, u32* __restrict__ dst_new, u64* __restrict__ wt_new,
    u32* __restrict__ end_bits, u32 U, unsigned long long* __restrict__ dang_w,
    unsigned long long* __restrict__ row_w) {
    __shared__ LtLds s;
    __shared__ u64 base[LT_LINES + 1];
    __shared__ u32 endw[LT_LINES / 32];
    __shared__ unsigned long long acc_d[LT_LINES + 1];
    __shared__ unsigned long long acc_w[HAS_WT ? LT_LINES + 1 : 1];
    const u64 t = blockIdx.x;
    u32 row[LT_PER];
    const u32 last = lt_rows(row_ptr_new, R, E, first, t, &s, row);
    const u32 r0 = first[t], nr = last - r0 + 1;
    for (u32 i = threadIdx.x; i < nr; i += LT_THREADS) {
        base[i] = row_ptr_raw[src_row[r0 + i]] - s.start[i];
        acc_d[i] = 0;
        if (HAS_WT) acc_w[i] = 0;
    }
    for (u32 i = threadIdx.x; i < LT_LINES / 32; i += LT_THREADS) endw[i] = 0;
    __syncthreads();
    const u64 e0 = t * LT_LINES, b0 = e0 + (u64)threadIdx.x * LT_PER;
    u64 d = 0, wsum = 0;  // this thread's run of lines of row cur
    u32 cur = row[0];
#pragma unroll
    for (int j = 0; j < LT_PER; ++j) {
        const u64 e = b0 + j;
        if (e < E) {
            const u32 k = row[j] - r0;
            const u64 src = e + base[k];
            const u32 v = pi[dst_raw[src]];
            dst_new[e] = v;
            const u64 wv = HAS_WT ? wt_raw[src] : 1;
            if (HAS_WT) wt_new[e] = wv;
            if (e + 1 == s.start[k + 1])
                atomicOr(&endw[(e - e0) >> 5], 1u << ((e - e0) & 31));
            if (dang_w) {
                if (row[j] != cur) {
                    if (d) atomicAdd(&acc_d[cur - r0], (unsigned long long)d);
                    if (HAS_WT && wsum) atomicAdd(&acc_w[cur - r0], (unsigned long long)wsum);
                    d = wsum = 0;
                    cur = row[j];
                }
                if (v >= U) d += wv;
                if (HAS_WT) wsum += wv;
            }
        }
    }
    if (dang_w) {
        if (d) atomicAdd(&acc_d[cur - r0], (unsigned long long)d);
        if (HAS_WT && wsum) atomicAdd(&acc_w[cur - r0], (unsigned long long)wsum);
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < LT_LINES / 32; i += LT_THREADS) end_bits[e0 / 32 + i] = endw[i];
    if (dang_w)
        for (u32 i = threadIdx.x; i < nr; i += LT_THREADS) {
            if (acc_d[i]) atomicAdd(&dang_w[r0 + i], acc_d[i]);
            if (HAS_WT && acc_w[i]) atomicAdd(&row_w[r0 + i], acc_w[i]);
        }
}

__global__ void permute_init_kernel(const u32* __restrict__ init_raw, const u32* __restrict__ pi,
                                    u64 N, u32* __restrict__ init_new) {
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < N) init_new[pi[v]] = init_raw[v];
}

__global__ void set_last_kernel(u64* p, u64 idx, u64 v) { p[idx] = v; }

// Tile metadata over internal rows.
__global__ void tile_setup_kernel(const u64* __restrict__ row_ptr, u32 U, u64 E, u64 n_tiles,
                                  u32* __restrict__ tile_row, u32* __restrict__ carry_row) {
    const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > n_tiles) return;
    if (t == n_tiles) {
        tile_row[t] = U - 1;
        return;
    }
    const u64 e0 = t * SIG_TW;
    u32 lo = 0, hi = U - 1;
    while (lo < hi) {
        const u32 mid = lo + (hi - lo + 1) / 2;
        if (row_ptr[mid] <= e0) lo = mid; else hi = mid - 1;
    }
    tile_row[t] = lo;
    const u64 elast = (e0 + SIG_TW < E ? e0 + SIG_TW : E) - 1;
    hi = U - 1;
    while (lo < hi) {
        const u32 mid = lo + (hi - lo + 1) / 2;
        if (row_ptr[mid] <= elast) lo = mid; else hi = mid - 1;
    }
    carry_row[t] = row_ptr[lo + 1] > elast + 1 ? lo : NO_ROW;
}
}  // namespace

void raw_release(RawGraph* r, hipStream_t s) {
    ctx_free(r->row_ptr, s);
    ctx_free(r->row_node, s);
    ctx_free(r->dst, s);
    ctx_free(r->wt, s);
    ctx_free(r->init, s);
    ctx_free(r->dst_err, s);
    *r = RawGraph();
}

int dgraph_build(rbsc_ctx* ctx, RawGraph* raw, rbsc_dgraph* g) {
    hipStream_t s = ctx->stream;
    const u64 N = raw->N, E_in = raw->E, U = raw->U;
    *g = rbsc_dgraph();
    g->N = (u32)N;
    g->E_in = E_in;
    g->U = (u32)U;
    u64 *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr}, *dang = nullptr,
        *deg = nullptr, *tmp = nullptr, *h2 = nullptr, *d_tot = nullptr;
    u32 *indeg = nullptr, *src_row = nullptr, *rep_raw = nullptr, *tfirst = nullptr;
    u64 R = U, E = E_in, h_tot[2] = {0, 0};
    int rc = RBSC_OK;
    bool hashing = false;
#define B_TRY(expr)                                                                  \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            rbsc_set_error("%s failed: %s", #expr, hipGetErrorString(_e));           \
            rc = _e == hipErrorOutOfMemory ? RBSC_ERR_NOMEM : RBSC_ERR_HIP;          \
            goto done;                                                               \
        }                                                                            \
    } while (0)
    B_TRY(ctx_alloc(&g->pi, sizeof(u32) * std::max<u64>(N, 1), s));
    if (raw->init) B_TRY(ctx_alloc(&g->init, sizeof(u32) * std::max<u64>(N, 1), s));
    B_TRY(ctx_alloc(&indeg, sizeof(u32) * std::max<u64>(N, 1), s));
    B_TRY(ctx_alloc(&dang, sizeof(u64) * (2 * N + 1), s));
    B_TRY(ctx_alloc(&deg, sizeof(u64) * (U + 1), s));
    B_TRY(ctx_alloc(&src_row, sizeof(u32) * std::max<u64>(U, 1), s));
    B_TRY(ctx_alloc(&tmp, sizeof(u64) * scan_tmp_words(std::max(N, U) + 1), s));
    B_TRY(ctx_alloc(&d_tot, sizeof(u64) * 2, s));
    B_TRY(ctx_alloc(&tfirst, sizeof(u32) * (lt_tiles(E_in) + 1), s));
    for (int i = 0; i < 2; ++i) {
        B_TRY(ctx_alloc(&keys[i], sizeof(u64) * std::max<u64>(U, 1), s));
        B_TRY(ctx_alloc(&vals[i], sizeof(u64) * std::max<u64>(U, 1), s));
    }
    if (U) {
        rc = sort_ws_init(&ctx->sort, U, ctx->d_err, s);
        if (rc) goto done;
    }
    zero_u32<<<grid_for(N), BLK, 0, s>>>(indeg, N);
    fill_ones_kernel<<<grid_for(N), BLK, 0, s>>>(dang, N);
    // Identical rows: sort rows by fingerprint, compare neighbours exactly,
    // one leader (stored row) per run of identical neighbours.
    hashing = U > 1 && ctx->dedup;
    if (hashing) {
        B_TRY(ctx_alloc(&h2, sizeof(u64) * U, s));
        B_TRY(ctx_alloc(&rep_raw, sizeof(u32) * U, s));
        const u64 nt = lt_tiles(E_in);
        lt_first_row_kernel<<<grid_for(nt + 1), BLK, 0, s>>>(raw->row_ptr, U, E_in, nt, tfirst);
        B_TRY(hipMemsetAsync(keys[0], 0, sizeof(u64) * U, s));
        B_TRY(hipMemsetAsync(h2, 0, sizeof(u64) * U, s));
        iota_u64_kernel<<<grid_for(U), BLK, 0, s>>>(vals[0], U);
    }
    // The kernels that read every line (head check, in-degrees, row
    // fingerprints), chunk by chunk as the lines arrive.
    {
        const int n_ch = raw->n_chunks ? raw->n_chunks : 1;
        const u64 st = raw->sample_indeg ? indeg_stride(E_in) : 1;
        auto* h1p = reinterpret_cast<unsigned long long*>(keys[0]);
        auto* h2p = reinterpret_cast<unsigned long long*>(h2);
        for (int c = 0; c < n_ch && E_in; ++c) {
            const u64 lo = c ? raw->chunk_end[c - 1] : 0;
            const u64 hi = raw->n_chunks ? raw->chunk_end[c] : E_in;
            if (raw->n_chunks) B_TRY(hipStreamWaitEvent(s, raw->chunk_ev[c], 0));
            if (hi <= lo) continue;
            if (raw->dst_err)
                dst_range_check_kernel<<<grid_for(hi - lo), BLK, 0, s>>>(raw->dst, lo, hi, N,
                                                                        raw->dst_err);
            const u64 k0 = (lo + st - 1) / st, k1 = (hi + st - 1) / st;
            if (k1 > k0)
                indeg_kernel<<<grid_for(k1 - k0), BLK, 0, s>>>(raw->dst, k0, hi, st, N, indeg);
            if (hashing) {
                const u64 t0 = lo / LT_LINES, t1 = (hi + LT_LINES - 1) / LT_LINES;
                if (raw->wt)
                    row_hash_tile_kernel<true><<<(unsigned)(t1 - t0), LT_THREADS, 0, s>>>(
                        raw->row_ptr, raw->dst, raw->wt, U, E_in, tfirst, h1p, h2p, t0);
                else
                    row_hash_tile_kernel<false><<<(unsigned)(t1 - t0), LT_THREADS, 0, s>>>(
                        raw->row_ptr, raw->dst, nullptr, U, E_in, tfirst, h1p, h2p, t0);
            }
        }
    }
    if (raw->dst_err) {
        // a head not below N: refused before any kernel indexes by head
        u32 bad = 0;
        B_TRY(hipMemcpyAsync(&bad, raw->dst_err, sizeof(u32), hipMemcpyDeviceToHost, s));
        B_TRY(hipStreamSynchronize(s));
        if (bad) {
            rbsc_set_error("malformed lines: a head is not below n_nodes");
            rc = RBSC_ERR_INPUT;
            goto done;
        }
    }
    if (hashing) {
        int w = 0;
        rc = radix_sort_u64(&ctx->sort, keys, vals, U, 0, 64, &w, s);
        if (rc) goto done;
        u64* head = keys[w ^ 1];
        u64* hscan = vals[w ^ 1];
        if (raw->wt)
            row_eq_wave_kernel<true><<<grid_for(U), BLK, 0, s>>>(keys[w], vals[w], h2, raw->row_ptr,
                                                                 raw->dst, raw->wt, U, head);
        else
            row_eq_wave_kernel<false><<<grid_for(U), BLK, 0, s>>>(keys[w], vals[w], h2,
                                                                  raw->row_ptr, raw->dst, nullptr,
                                                                  U, head);
        rc = scan_excl_u64(head, hscan, U, nullptr, tmp, s);
        if (rc) goto done;
        run_leader_kernel<<<grid_for(U), BLK, 0, s>>>(head, hscan, vals[w], U, h2);  // h2 free now
        rep_raw_kernel<<<grid_for(U), BLK, 0, s>>>(head, hscan, vals[w], h2, U, rep_raw);
        B_TRY(hipMemsetAsync(d_tot, 0, sizeof(u64), s));
        count_dups_kernel<<<(unsigned)std::min<u64>(grid_for(U), 1024), BLK, 0, s>>>(
            rep_raw, U, reinterpret_cast<unsigned long long*>(d_tot));
        B_TRY(hipMemcpyAsync(h_tot, d_tot, sizeof(u64), hipMemcpyDeviceToHost, s));
        B_TRY(hipStreamSynchronize(s));
        R = U - h_tot[0];
        if (R < U && !raw->init)
            absorb_indeg_kernel<<<grid_for(U), BLK, 0, s>>>(raw->row_node, rep_raw, U, indeg);
    }
    if (U) {
        order_keys_kernel<<<grid_for(U), BLK, 0, s>>>(raw->row_node, U, indeg, rep_raw, keys[0],
                                                      vals[0], dang);
        int which = 0;
        rc = radix_sort_u64(&ctx->sort, keys, vals, U, 0, 64, &which, s);
        if (rc) goto done;
        place_rows_kernel<<<grid_for(U), BLK, 0, s>>>(keys[which], vals[which], U, R, raw->row_ptr,
                                                       g->pi, deg, src_row);
    }
    rc = scan_excl_u64(dang, dang + N, N, nullptr, tmp, s);  // scan into the upper half
    if (rc) goto done;
    place_dangling_kernel<<<grid_for(N), BLK, 0, s>>>(dang, dang + N, N, U, g->pi);
    // Row offsets of the stored rows; their total is the stored line count.
    B_TRY(ctx_alloc(&g->row_ptr, sizeof(u64) * (R + 1), s));
    rc = scan_excl_u64(deg, g->row_ptr, R, d_tot + 1, tmp, s);
    if (rc) goto done;
    if (R < U) {
        B_TRY(hipMemcpyAsync(h_tot + 1, d_tot + 1, sizeof(u64), hipMemcpyDeviceToHost, s));
        B_TRY(hipStreamSynchronize(s));
        E = h_tot[1];
    }
    g->R = (u32)R;
    g->E = E;
    g->n_tiles = (E + SIG_TW - 1) / SIG_TW;
    B_TRY(ctx_alloc(&g->dst, sizeof(u32) * std::max<u64>(E, 1) + 32, s));
    if (raw->wt) B_TRY(ctx_alloc(&g->wt, sizeof(u64) * std::max<u64>(E, 1), s));
    g->end_words = (E + 31) / 32 + 64;  // padding covers whole-tile byte reads
    B_TRY(ctx_alloc(&g->end_bits, sizeof(u32) * g->end_words, s));
    B_TRY(ctx_alloc(&g->tile_row, sizeof(u32) * (g->n_tiles + 1), s));
    B_TRY(ctx_alloc(&g->carry_row, sizeof(u32) * std::max<u64>(g->n_tiles, 1), s));
    if (R < U) {
        B_TRY(ctx_alloc(&g->dup_rep, sizeof(u32) * (U - R), s));
        dup_rep_kernel<<<grid_for(U - R), BLK, 0, s>>>(src_row, rep_raw, raw->row_node, g->pi, R, U,
                                                       g->dup_rep);
    }
    if (!raw->init && R) {  // per-row weights of the first round's closed form
        B_TRY(ctx_alloc(&g->dang_w, sizeof(u64) * R, s));
        B_TRY(hipMemsetAsync(g->dang_w, 0, sizeof(u64) * R, s));
        if (raw->wt) {
            B_TRY(ctx_alloc(&g->row_w, sizeof(u64) * R, s));
            B_TRY(hipMemsetAsync(g->row_w, 0, sizeof(u64) * R, s));
        }
    }
    g->bytes = sizeof(u64) * (R + 1) + 4 * N + 4 * E + (raw->wt ? 8 * E : 0) +
               (raw->init ? 4 * N : 0) + 4 * g->end_words + 16 * (g->n_tiles + 1) + 4 * (U - R) +
               (g->dang_w ? 8 * (u64)R : 0) + (g->row_w ? 8 * (u64)R : 0);
    set_last_kernel<<<1, 1, 0, s>>>(g->row_ptr, R, E);
    zero_u32<<<grid_for(g->end_words), BLK, 0, s>>>(g->end_bits, g->end_words);
    if (R) {
        const u64 nt = lt_tiles(E);
        lt_first_row_kernel<<<grid_for(nt + 1), BLK, 0, s>>>(g->row_ptr, R, E, nt, tfirst);
        auto* dw = reinterpret_cast<unsigned long long*>(g->dang_w);
        auto* rw = reinterpret_cast<unsigned long long*>(g->row_w);
        if (raw->wt)
            copy_rows_tile_kernel<true><<<(unsigned)nt, LT_THREADS, 0, s>>>(
                g->row_ptr, src_row, raw->row_ptr, raw->dst, raw->wt, g->pi, R, E, tfirst, g->dst,
                g->wt, g->end_bits, (u32)U, dw, rw);
        else
            copy_rows_tile_kernel<false><<<(unsigned)nt, LT_THREADS, 0, s>>>(
                g->row_ptr, src_row, raw->row_ptr, raw->dst, nullptr, g->pi, R, E, tfirst, g->dst,
                nullptr, g->end_bits, (u32)U, dw, nullptr);
    }
    if (raw->init) permute_init_kernel<<<grid_for(N), BLK, 0, s>>>(raw->init, g->pi, N, g->init);
    if (g->n_tiles)
        tile_setup_kernel<<<grid_for(g->n_tiles + 1), BLK, 0, s>>>(g->row_ptr, (u32)R, E, g->n_tiles,
                                                                   g->tile_row, g->carry_row);
    B_TRY(hipGetLastError());
    B_TRY(hipStreamSynchronize(s));
    rc = check_kernel_errors(ctx);
done:
    for (int i = 0; i < 2; ++i) {
        ctx_free(keys[i], s);
        ctx_free(vals[i], s);
    }
    for (void* p : {(void*)dang, (void*)deg, (void*)tmp, (void*)indeg, (void*)src_row, (void*)h2,
                    (void*)rep_raw, (void*)d_tot, (void*)tfirst})
        ctx_free(p, s);
    return rc;
#undef B_TRY
}

// ------------------------------------------------------------ partitioning
namespace {
// Partition of a line by its head: interleaved (head mod P) or, with a hot
// range, 0 for heads below `hot` or without lines (never gathered) and 1 for
// the other heads with lines.
__device__ __forceinline__ u32 part_of(u32 d, u32 P, u32 hot, u32 U) {
    if (hot) return (d >= hot && d < U) ? 1u : 0u;
    return d & (P - 1);
}
// counts[p * U + j] = lines of row j whose head is in partition p.
__global__ void part_count_kernel(const u64* __restrict__ row_ptr, const u32* __restrict__ dst,
                                  u64 U, u32 P, u32 hot, u32 lined, u64* __restrict__ counts) {
    const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= U) return;
    u64 c[16];
    for (u32 p = 0; p < P; ++p) c[p] = 0;
    for (u64 e = row_ptr[j]; e < row_ptr[j + 1]; ++e) c[part_of(dst[e], P, hot, lined)]++;
    for (u32 p = 0; p < P; ++p) counts[(u64)p * U + j] = c[p];
}
__global__ void flag_nonzero_kernel(const u64* __restrict__ a, u64 n, u64* __restrict__ f) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = a[i] != 0;
}
// Scatter each row's lines into its partition runs; write segment starts,
// segment rows and row-end bits.
template <bool HAS_WT>
__global__ void part_fill_kernel(const u64* __restrict__ row_ptr, const u32* __restrict__ dst,
                                 const u64* __restrict__ wt, u64 U, u32 P, u32 hot, u32 lined,
                                 const u64* __restrict__ line_off, const u64* __restrict__ seg_idx,
                                 const u64* __restrict__ counts, const u64* __restrict__ adj_line,
                                 const u64* __restrict__ adj_seg, u32* __restrict__ ndst,
                                 u64* __restrict__ nwt, u64* __restrict__ seg_ptr,
                                 u32* __restrict__ seg_row, u32* __restrict__ end_bits) {
    const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= U) return;
    u64 pos[16];
    for (u32 p = 0; p < P; ++p) {
        const u64 i = (u64)p * U + j;
        pos[p] = line_off[i] + adj_line[p];
        if (counts[i]) {
            const u64 k = seg_idx[i] + adj_seg[p];
            seg_ptr[k] = pos[p];
            seg_row[k] = (u32)j;
            const u64 last = pos[p] + counts[i] - 1;
            atomicOr(&end_bits[last >> 5], 1u << (last & 31));
        }
    }
    for (u64 e = row_ptr[j]; e < row_ptr[j + 1]; ++e) {
        const u32 d = dst[e];
        const u32 p = part_of(d, P, hot, lined);
        ndst[pos[p]] = d;
        if (HAS_WT) nwt[pos[p]] = wt[e];
        pos[p]++;
    }
}
__global__ void pad_fill_kernel(u32* __restrict__ ndst, u64 a, u64 b) {
    const u64 i = a + (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < b) ndst[i] = PAD_DST;
}
__global__ void pad_seg_kernel(u64* __restrict__ seg_ptr, u32* __restrict__ seg_row, u64 k,
                               u64 start, u32 slot, u32* __restrict__ end_bits, u64 last) {
    seg_ptr[k] = start;
    seg_row[k] = slot;
    atomicOr(&end_bits[last >> 5], 1u << (last & 31));
}
}  // namespace

int dgraph_relayout(rbsc_ctx* ctx, rbsc_dgraph* g, u32 P, u32 hot) {
    if (P == 1 && g->n_parts == 1) return RBSC_OK;
    if (g->n_parts != 1 || P == 0 || P > 16 || (P & (P - 1)) || (hot && P != 2)) {
        rbsc_set_error("relayout: partitions must be a power of two <= 16 (2 with a hot range), "
                       "from the plain layout");
        return RBSC_ERR_INVALID;
    }
    if (!g->dst) {
        rbsc_set_error("relayout: lines are streamed from the host; partitioning needs them resident");
        return RBSC_ERR_INVALID;
    }
    hipStream_t s = ctx->stream;
    // rows: the stored rows; heads below g->U are nodes with lines
    const u64 U = g->R, E = g->E, N = g->N;
    const u32 lined = g->U;
    if (U == 0 || E == 0) return RBSC_OK;
    const u64 PU = (u64)P * U;
    u64 *counts = nullptr, *line_off = nullptr, *flags = nullptr, *tmp = nullptr, *tot = nullptr,
        *adj = nullptr, *seg_ptr = nullptr, *nwt = nullptr;
    u32 *ndst = nullptr, *seg_row = nullptr, *nbits = nullptr, *tile_row = nullptr,
        *carry_row = nullptr;
    std::vector<u64> Ep(P), Rp(P), adj_line(P), adj_seg(P), h(2 * P);
    int rc = RBSC_OK;
    u64 E_pad = 0, R = 0, n_tiles = 0, end_words = 0;
#define R_TRY(expr)                                                                  \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            rbsc_set_error("%s failed: %s", #expr, hipGetErrorString(_e));           \
            rc = _e == hipErrorOutOfMemory ? RBSC_ERR_NOMEM : RBSC_ERR_HIP;          \
            goto done;                                                               \
        }                                                                            \
    } while (0)
    R_TRY(ctx_alloc(&counts, sizeof(u64) * PU, s));
    R_TRY(ctx_alloc(&line_off, sizeof(u64) * PU, s));
    R_TRY(ctx_alloc(&flags, sizeof(u64) * PU, s));
    R_TRY(ctx_alloc(&tmp, sizeof(u64) * scan_tmp_words(PU + 1), s));
    R_TRY(ctx_alloc(&tot, sizeof(u64) * 2 * P, s));
    R_TRY(ctx_alloc(&adj, sizeof(u64) * 2 * P, s));
    part_count_kernel<<<grid_for(U), BLK, 0, s>>>(g->row_ptr, g->dst, U, P, hot, lined, counts);
    flag_nonzero_kernel<<<grid_for(PU), BLK, 0, s>>>(counts, PU, flags);
    // per-partition totals of lines and segments
    for (u32 p = 0; p < P; ++p) {
        rc = scan_excl_u64(counts + (u64)p * U, line_off + (u64)p * U, U, tot + p, tmp, s);
        if (rc) goto done;
    }
    for (u32 p = 0; p < P; ++p) {
        rc = scan_excl_u64(flags + (u64)p * U, flags + (u64)p * U, U, tot + P + p, tmp, s);
        if (rc) goto done;
    }
    R_TRY(hipMemcpyAsync(h.data(), tot, sizeof(u64) * 2 * P, hipMemcpyDeviceToHost, s));
    R_TRY(hipStreamSynchronize(s));
    {
        u64 line_base = 0, seg_base = 0;
        for (u32 p = 0; p < P; ++p) {
            Ep[p] = h[p];
            Rp[p] = h[P + p];
            adj_line[p] = line_base;  // partition p starts at this padded line
            adj_seg[p] = seg_base;
            const u64 padded = (Ep[p] + SIG_TW - 1) / SIG_TW * SIG_TW;
            line_base += padded;
            seg_base += Rp[p] + (padded > Ep[p] ? 1 : 0);
        }
        E_pad = line_base;
        R = seg_base;
    }
    n_tiles = E_pad / SIG_TW;
    end_words = (E_pad + 31) / 32 + 64;
    R_TRY(ctx_alloc(&ndst, sizeof(u32) * E_pad + 32, s));
    if (g->wt) R_TRY(ctx_alloc(&nwt, sizeof(u64) * E_pad, s));
    R_TRY(ctx_alloc(&seg_ptr, sizeof(u64) * (R + 1), s));
    R_TRY(ctx_alloc(&seg_row, sizeof(u32) * R, s));
    R_TRY(ctx_alloc(&nbits, sizeof(u32) * end_words, s));
    R_TRY(ctx_alloc(&tile_row, sizeof(u32) * (n_tiles + 1), s));
    R_TRY(ctx_alloc(&carry_row, sizeof(u32) * n_tiles, s));
    R_TRY(hipMemcpyAsync(adj, adj_line.data(), sizeof(u64) * P, hipMemcpyHostToDevice, s));
    R_TRY(hipMemcpyAsync(adj + P, adj_seg.data(), sizeof(u64) * P, hipMemcpyHostToDevice, s));
    zero_u32<<<grid_for(end_words), BLK, 0, s>>>(nbits, end_words);
    if (nwt) R_TRY(hipMemsetAsync(nwt, 0, sizeof(u64) * E_pad, s));
    if (g->wt)
        part_fill_kernel<true><<<grid_for(U), BLK, 0, s>>>(g->row_ptr, g->dst, g->wt, U, P, hot, lined, line_off,
                                                           flags, counts, adj, adj + P, ndst, nwt,
                                                           seg_ptr, seg_row, nbits);
    else
        part_fill_kernel<false><<<grid_for(U), BLK, 0, s>>>(g->row_ptr, g->dst, nullptr, U, P,
                                                            hot, lined, line_off, flags, counts, adj, adj + P,
                                                            ndst, nullptr, seg_ptr, seg_row, nbits);
    for (u32 p = 0; p < P; ++p) {
        const u64 a = adj_line[p] + Ep[p];
        const u64 b = (p + 1 < P) ? adj_line[p + 1] : E_pad;
        if (b > a) {
            pad_fill_kernel<<<grid_for(b - a), BLK, 0, s>>>(ndst, a, b);
            pad_seg_kernel<<<1, 1, 0, s>>>(seg_ptr, seg_row, adj_seg[p] + Rp[p], a, (u32)N, nbits,
                                           b - 1);
        }
    }
    set_last_kernel<<<1, 1, 0, s>>>(seg_ptr, R, E_pad);
    tile_setup_kernel<<<grid_for(n_tiles + 1), BLK, 0, s>>>(seg_ptr, (u32)R, E_pad, n_tiles,
                                                            tile_row, carry_row);
    R_TRY(hipGetLastError());
    R_TRY(hipStreamSynchronize(s));
    // swap in the partitioned arrays
    ctx_free(g->dst, s);
    ctx_free(g->wt, s);
    ctx_free(g->end_bits, s);
    ctx_free(g->tile_row, s);
    ctx_free(g->carry_row, s);
    g->dst = ndst;
    g->wt = nwt;
    g->end_bits = nbits;
    g->end_words = end_words;
    g->tile_row = tile_row;
    g->carry_row = carry_row;
    g->n_tiles = n_tiles;
    g->seg_row = seg_row;
    g->n_segs = R;
    g->E_pad = E_pad;
    g->n_parts = P;
    g->part_tile.assign(P + 1, 0);
    for (u32 p = 0; p < P; ++p) g->part_tile[p] = adj_line[p] / SIG_TW;
    g->part_tile[P] = n_tiles;
    R_TRY(ctx_alloc(&g->d_part_tile, sizeof(u64) * (P + 1), s));
    R_TRY(hipMemcpy(g->d_part_tile, g->part_tile.data(), sizeof(u64) * (P + 1), hipMemcpyHostToDevice));
    g->bytes += 4 * (E_pad - E) + 4 * R;
    ndst = nullptr;
    nwt = nullptr;
    nbits = nullptr;
    tile_row = carry_row = nullptr;
    seg_row = nullptr;
done:
    for (u64* p : {counts, line_off, flags, tmp, tot, adj, seg_ptr, nwt})
        ctx_free(p, s);
    for (u32* p : {ndst, seg_row, nbits, tile_row, carry_row})
        ctx_free(p, s);
    return rc;
#undef R_TRY
}
