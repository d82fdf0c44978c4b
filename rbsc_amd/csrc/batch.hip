// batch.hip -- many small graphs refined in persistent launches, one per size tier.
//
// Graphs are launched in four tiers by size (above BATCH_NMID nodes with the
// autotuned shape and the full LDS layout, one workgroup per CU; above
// BATCH_NLOW with half the layout, two per CU; above BATCH_NTINY with a
// quarter, four per CU; the rest with an eighth, eight per CU), so small
// graphs do not hold a CU's whole LDS.  In each tier a
// workgroup takes graphs from an atomic queue (largest first) and runs
// the whole refinement of one graph inside LDS: colours, signatures, a
// stable LDS radix sort of the top 32 signature bits over positions kept in
// colour order (the rare runs colliding in those bits ordered afterwards),
// the dense rank and the fixpoint test, then the first-occurrence canonical
// labels.  Every graph gets exactly the colours the single-graph path gives
// it alone (same round function, same dense ranks, same stop rule), so a
// result is bit-identical alone and in a batch.  Graphs with more nodes than
// fit in LDS go through the single-graph path instead (host side below).
#include "internal.h"
#include "batch.h"
#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace {

// Workgroup width (picked by the autotuner for the graphs above BATCH_NMID
// nodes) and node capacity NM (the LDS layout) are template parameters.
constexpr int MID_THREADS = 512;
constexpr int LOW_THREADS = 256;
constexpr int TINY_THREADS = 128;
template <int BT, int NM>
struct Cfg {
    static constexpr int BW = BT / 64;                   // waves
    static constexpr int IPT = (NM + BT - 1) / BT;       // items per thread (blocked)
    static constexpr int IPW = (NM + BW - 1) / BW;       // items per wave chunk
    static constexpr int WIT = (IPW + 63) / 64;          // 64-item iterations per wave
};

template <int BT, int NM>
struct Lds {
    u64 key[2][NM];
    u32 val[2][NM];   // (dense colour << 16) | node
    u32 col[NM];      // colour used in the hash (exact c0 in round 0)
    u32 whist[Cfg<BT, NM>::BW][256];
    u32 wsum[Cfg<BT, NM>::BW];
    u32 scal[8];
};

// Exclusive scan of one u32 per thread over the workgroup; returns the
// exclusive prefix, total in *tot.  Uses s->wsum.
template <int BT, int NM>
__device__ u32 block_scan_excl(u32 v, Lds<BT, NM>* s, u32* tot) {
    constexpr int BW = Cfg<BT, NM>::BW;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    u32 x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const u32 y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) s->wsum[wave] = x;
    __syncthreads();
    u32 off = 0, t = 0;
#pragma unroll
    for (int w = 0; w < BW; ++w) {
        if (w < wave) off += s->wsum[w];
        t += s->wsum[w];
    }
    __syncthreads();
    *tot = t;
    return off + x - v;
}

// One stable LSD pass over n items by digit (key >> shift) & (2^RB - 1)
// (RB = 8 or 4 bits: the autotuner picks per size tier), from buffer `in`
// to buffer 1 - in.  Wave w owns items [w*IPW, (w+1)*IPW).
template <int BT, int NM, int RB>
__device__ void lds_radix_pass(Lds<BT, NM>* s, int in, int n, int shift) {
    constexpr int BW = Cfg<BT, NM>::BW, IPW = Cfg<BT, NM>::IPW, WIT = Cfg<BT, NM>::WIT;
    constexpr int ND = 1 << RB;
    constexpr u32 DM = (u32)ND - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < BW * 256; i += BT) (&s->whist[0][0])[i] = 0;
    __syncthreads();
    const u64 below = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    u32 rank[WIT];
#pragma unroll
    for (int it = 0; it < WIT; ++it) {
        const int idx = wave * IPW + it * 64 + lane;
        const bool valid = idx < n && (it * 64 + lane) < IPW;
        const u32 d = valid ? (u32)(s->key[in][idx] >> shift) & DM : 0u;
        // the group's count is read by every lane, then written by its
        // lowest lane (in-order LDS, as in wave_rank)
        const u64 peers = wave_match<RB>(d, __ballot(valid));
        const u32 bl = (u32)__popcll(peers & below);
        u32 old = 0;
        if (valid) old = s->whist[wave][d];
        if (valid && bl == 0) s->whist[wave][d] = old + (u32)__popcll(peers);
        rank[it] = old + bl;
    }
    __syncthreads();
    // digit-major offsets: for digit d, waves in order; thread t takes the
    // digits [t * DPT, t * DPT + DPT) (one each from ND threads up)
    constexpr int DPT = (ND + BT - 1) / BT;
    u32 tot_d[DPT];
    u32 tsum = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = (int)threadIdx.x * DPT + k;
        tot_d[k] = 0;
        if (d < ND) {
#pragma unroll
            for (int w = 0; w < BW; ++w) {
                const u32 c = s->whist[w][d];
                s->whist[w][d] = tot_d[k];
                tot_d[k] += c;
            }
        }
        tsum += tot_d[k];
    }
    u32 total;
    u32 run = block_scan_excl(tsum, s, &total);
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = (int)threadIdx.x * DPT + k;
        if (d < ND) {
#pragma unroll
            for (int w = 0; w < BW; ++w) s->whist[w][d] += run;
            run += tot_d[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < WIT; ++it) {
        const int idx = wave * IPW + it * 64 + lane;
        if (idx < n && (it * 64 + lane) < IPW) {
            const u64 k = s->key[in][idx];
            const u32 d = (u32)(k >> shift) & DM;
            const u32 pos = s->whist[wave][d] + rank[it];
            s->key[1 - in][pos] = k;
            s->val[1 - in][pos] = s->val[in][idx];
        }
    }
    __syncthreads();
}

template <int BT, int NM, int RB>
__device__ void lds_sort(Lds<BT, NM>* s, int n, int bit_begin, int bit_end) {
    int in = 0;
    for (int sh = bit_begin; sh < bit_end; sh += RB) {
        lds_radix_pass<BT, NM, RB>(s, in, n, sh);
        in ^= 1;
    }
    // callers sort 32 or 16 bits, an even number of passes for either digit
    // width: the result is back in buffer 0
}

template <int BT, int NM, bool HAS_WT, bool HAS_INIT, int RB>
__global__ __launch_bounds__(BT) void batch_refine_kernel(
    const BatchDesc* __restrict__ desc, const u32* __restrict__ order, u32 n_jobs,
    const u32* __restrict__ range, const u32* __restrict__ skip, const u32* __restrict__ rp_all, const unsigned short* __restrict__ dst_all,
    const u64* __restrict__ wt_all, const u32* __restrict__ init_all, u32 max_rounds,
    u32* __restrict__ labels_out, u32* __restrict__ colors_out, u32* __restrict__ sizes_out,
    u32* __restrict__ blocks_out, u32* __restrict__ rounds_out, u32* __restrict__ queue,
    int split_lo) {
    constexpr int IPT = Cfg<BT, NM>::IPT;
    __shared__ Lds<BT, NM> s;
    // Every job runs the whole body: no `continue` back to the thread-0 queue
    // fetch.  A branch that skips the body lets the compiler put the barrier
    // after the fetch at the head of an inner loop that only thread 0 leaves,
    // and the rest of wave 0 then waits there for ever.  Graphs without nodes
    // are therefore answered on the host and never queued.
    // A layout built on the device gives the tier's jobs as a device range
    // [range[0], range[1]) of the order, and a nonzero *skip (a layout
    // error found by the build) cancels the launch.
    // (Read through readfirstlane: the values are the same for every thread,
    // and the loop's exit below must stay a uniform branch.)
    if (skip && __builtin_amdgcn_readfirstlane(*skip)) return;
    if (range) {
        const u32 j0 = __builtin_amdgcn_readfirstlane(range[0]);
        n_jobs = __builtin_amdgcn_readfirstlane(range[1]) - j0;
        order += j0;
    }
    for (;;) {
        if (threadIdx.x == 0) s.scal[0] = atomicAdd(queue, 1u);
        __syncthreads();
        const u32 job = s.scal[0];
        __syncthreads();
        if (job >= n_jobs) break;
        const u32 g = order[job];
        const BatchDesc d = desc[g];
        const int n = (int)d.n;
        const u32* rp = rp_all + d.rp_base;
        const unsigned short* dst = dst_all + d.line_base;
        const u64* wt = HAS_WT ? wt_all + d.line_base : nullptr;

        // ---- round-0 colours, positions sorted by colour
        if (HAS_INIT) {
            const u32* ic = init_all + d.node_base;
            for (int u = threadIdx.x; u < n; u += BT) {
                s.col[u] = ic[u];
                s.key[0][u] = ic[u];
                s.val[0][u] = (u32)u;
            }
            __syncthreads();
            lds_sort<BT, NM, RB>(&s, n, 0, 32);
            // dense rank of the initial colours, in position order
            u32 f[IPT];
            u32 cnt = 0;
            const int b0 = threadIdx.x * IPT;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int i = b0 + j;
                f[j] = (i > 0 && i < n && s.key[0][i] != s.key[0][i - 1]) ? 1u : 0u;
                cnt += f[j];
            }
            u32 tot;
            u32 r = block_scan_excl(cnt, &s, &tot);
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int i = b0 + j;
                r += f[j];
                if (i < n) s.val[0][i] = (r << 16) | (s.val[0][i] & 0xFFFFu);
            }
            __syncthreads();
        } else {
            // nodes without lines first (colour 0), then the rest (colour n+1)
            u32 f[IPT];
            u32 cnt = 0;
            const int b0 = threadIdx.x * IPT;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int u = b0 + j;
                f[j] = (u < n && rp[u + 1] == rp[u]) ? 1u : 0u;
                cnt += f[j];
            }
            u32 ndang;
            const u32 pd = block_scan_excl(cnt, &s, &ndang);
            u32 qd = pd;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int u = b0 + j;
                if (u < n) {
                    const u32 pos = f[j] ? qd : ndang + ((u32)u - qd);
                    qd += f[j];
                    s.val[0][pos] = ((f[j] ? 0u : 1u) << 16) | (u32)u;
                    s.col[u] = f[j] ? 0u : (u32)n + 1u;
                }
            }
            __syncthreads();
        }

        // ---- rounds
        u32 K = 1, Knew = 1, it = 0;
        do {
            K = Knew;
            // signatures by node into key[1]
            for (int u = threadIdx.x; u < n; u += BT) {
                u64 acc = 0;
                for (u32 e = rp[u]; e < rp[u + 1]; ++e) {
                    u64 h = rbsc_randomize((u64)s.col[dst[e]] + it);
                    if (HAS_WT) h *= wt[e];
                    acc += h;
                }
                s.key[1][u] = acc;
            }
            __syncthreads();
            for (int i = threadIdx.x; i < n; i += BT) s.key[0][i] = s.key[1][s.val[0][i] & 0xFFFFu];
            if (threadIdx.x == 0) s.scal[1] = 0;
            __syncthreads();
            // Stable sort by the top 32 signature bits (4 passes; split_lo
            // is 32, or 48 as a verification hook that makes collisions
            // common), then neighbours equal there and different below --
            // distinct signatures colliding in 32 bits, about n^2 / 2^33
            // pairs a round -- are put in order by thread 0 with a stable
            // insertion pass over the full keys (inversions exist only
            // inside such runs).
            lds_sort<BT, NM, RB>(&s, n, split_lo, 64);
            for (int i = threadIdx.x + 1; i < n; i += BT) {
                const u64 a = s.key[0][i - 1], b = s.key[0][i];
                if ((a >> split_lo) == (b >> split_lo) && a != b) s.scal[1] = 1;
            }
            __syncthreads();
            if (s.scal[1]) {  // uniform
                if (threadIdx.x == 0) {
                    for (int i = 1; i < n; ++i) {
                        const u64 k = s.key[0][i];
                        const u32 v = s.val[0][i];
                        int j = i;
                        while (j > 0 && s.key[0][j - 1] > k) {
                            s.key[0][j] = s.key[0][j - 1];
                            s.val[0][j] = s.val[0][j - 1];
                            --j;
                        }
                        s.key[0][j] = k;
                        s.val[0][j] = v;
                    }
                }
                __syncthreads();
            }
            // dense rank of (signature, colour)
            u32 f[IPT];
            u32 cnt = 0;
            const int b0 = threadIdx.x * IPT;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int i = b0 + j;
                f[j] = (i > 0 && i < n &&
                        (s.key[0][i] != s.key[0][i - 1] || (s.val[0][i] >> 16) != (s.val[0][i - 1] >> 16)))
                           ? 1u : 0u;
                cnt += f[j];
            }
            u32 tot;
            u32 r = block_scan_excl(cnt, &s, &tot);
            __syncthreads();
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int i = b0 + j;
                r += f[j];
                if (i < n) {
                    const u32 node = s.val[0][i] & 0xFFFFu;
                    s.val[0][i] = (r << 16) | node;
                    s.col[node] = r;
                }
            }
            Knew = n > 0 ? tot + 1 : 0;
            ++it;
            __syncthreads();
        } while (K != Knew && !(max_rounds && it >= max_rounds));

        // ---- outputs: raw colours, first-occurrence labels, counts
        u32* first = reinterpret_cast<u32*>(&s.key[1][0]);
        u32* lab = first + NM;
        for (int c = threadIdx.x; c < n; c += BT) first[c] = 0xFFFFFFFFu;
        __syncthreads();
        for (int u = threadIdx.x; u < n; u += BT) atomicMin(&first[s.col[u]], (u32)u);
        __syncthreads();
        {
            u32 f[IPT];
            u32 cnt = 0;
            const int b0 = threadIdx.x * IPT;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int u = b0 + j;
                f[j] = (u < n && first[s.col[u]] == (u32)u) ? 1u : 0u;
                cnt += f[j];
            }
            u32 tot;
            u32 r = block_scan_excl(cnt, &s, &tot);
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                const int u = b0 + j;
                if (u < n && f[j]) lab[s.col[u]] = r;
                r += f[j];
            }
        }
        __syncthreads();
        for (int u = threadIdx.x; u < n; u += BT) {
            if (labels_out) labels_out[d.node_base + u] = lab[s.col[u]];
            if (colors_out) colors_out[d.node_base + u] = s.col[u];
        }
        if (threadIdx.x == 0) {
            blocks_out[g] = Knew;
            rounds_out[g] = it;
        }
        __syncthreads();
        if (sizes_out) {
            // Block sizes in descending order, then zeros (n entries), by a
            // counting sort: size[c] nodes per colour c < Knew, then bys[i]
            // blocks of size n - i, scanned, so block j of the descending
            // order has the size n - i for the last i whose scan is <= j.
            u32* size = first;                                  // n entries
            u32* bys = reinterpret_cast<u32*>(&s.key[0][0]);    // n entries
            for (int i = threadIdx.x; i < n; i += BT) {
                size[i] = 0;
                bys[i] = 0;
            }
            __syncthreads();
            for (int u = threadIdx.x; u < n; u += BT) atomicAdd(&size[s.col[u]], 1u);
            __syncthreads();
            for (int c = threadIdx.x; c < (int)Knew; c += BT) atomicAdd(&bys[n - (int)size[c]], 1u);
            __syncthreads();
            u32 v[IPT];
            u32 cnt = 0;
            const int b0 = threadIdx.x * IPT;
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                v[j] = b0 + j < n ? bys[b0 + j] : 0u;
                cnt += v[j];
            }
            u32 tot;
            u32 r = block_scan_excl(cnt, &s, &tot);
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                if (b0 + j < n) bys[b0 + j] = r;
                r += v[j];
            }
            __syncthreads();
            for (int j = threadIdx.x; j < n; j += BT) {
                u32 out = 0;
                if (j < (int)Knew) {
                    int lo = 0, hi = n - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (bys[mid] <= (u32)j) lo = mid; else hi = mid - 1;
                    }
                    out = (u32)(n - lo);
                }
                sizes_out[d.node_base + j] = out;
            }
            __syncthreads();
        }
    }
}
static_assert(sizeof(Lds<1024, BATCH_NMAX>) <= 160 * 1024, "batch LDS layout exceeds 160 KB");
static_assert(sizeof(Lds<MID_THREADS, BATCH_NMID>) <= 80 * 1024,
              "two mid-size layouts must fit one CU's LDS");
static_assert(sizeof(Lds<LOW_THREADS, BATCH_NLOW>) <= 40 * 1024,
              "four small layouts must fit one CU's LDS");
static_assert(sizeof(Lds<TINY_THREADS, BATCH_NTINY>) <= 20 * 1024,
              "eight tiny layouts must fit one CU's LDS");

__global__ void zero_queue(u32* q) {
    q[0] = 0;
    q[1] = 0;
    q[2] = 0;
    q[3] = 0;
}

// Jobs [j0, j1) of the order (largest first) on `grid` workgroups, taken
// from the queue counter q.
template <int BT, int NM>
static void launch_bt(hipStream_t s, unsigned grid, const BatchDevice& b, u32 j0, u32 j1,
                      u32 max_rounds, u32* d_labels, u32* d_colors, u32* d_sizes, u32* d_blocks,
                      u32* d_rounds, u32* q, int split_lo, int rb, const u32* range = nullptr,
                      const u32* skip = nullptr) {
#define BL(W, I, R)                                                                           \
    batch_refine_kernel<BT, NM, W, I, R><<<grid, BT, 0, s>>>(                                   \
        b.desc, b.order + j0, j1 - j0, range, skip, b.rp, b.dst, b.wt, b.init, max_rounds,      \
        d_labels, d_colors, d_sizes, d_blocks, d_rounds, q, split_lo)
#define BLR(W, I)                   \
    do {                            \
        if (rb == 4) BL(W, I, 4);   \
        else BL(W, I, 8);           \
    } while (0)
    if (b.wt) {
        if (b.init) BLR(true, true); else BLR(true, false);
    } else {
        if (b.init) BLR(false, true); else BLR(false, false);
    }
#undef BLR
#undef BL
}

}  // namespace

// Stream of each tier: tiers 1 and 3 run on the context's second stream
// beside tiers 0 and 2 (forked after the stream's earlier work, joined
// before its next).  A tier's persistent workgroups finish their queue
// unevenly, and the other stream's tier fills the CUs they leave: on 50,000
// mixed graphs 6.58 -> 5.94 ms of device time (each tier on a stream of its
// own: 5.99 ms; tiers 1-3 together on one second stream: 6.22 ms).
static hipError_t tier_streams(rbsc_ctx* ctx, bool fork, hipStream_t st[4]) {
    st[0] = st[1] = st[2] = st[3] = ctx->stream;
    if (!fork) return hipSuccess;
    hipError_t e = hipEventRecord(ctx->ev_fork, ctx->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->copy_stream, ctx->ev_fork, 0);
    st[1] = st[3] = ctx->copy_stream;
    return e;
}
static hipError_t tier_join(rbsc_ctx* ctx, bool fork) {
    if (!fork) return hipSuccess;
    hipError_t e = hipEventRecord(ctx->ev_join, ctx->copy_stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0);
    return e;
}

int batch_launch(rbsc_ctx* ctx, const BatchDevice& b, u32 max_rounds, u32* d_labels, u32* d_colors,
                 u32* d_sizes, u32* d_blocks, u32* d_rounds, u32* d_queue, int threads, int n_wg,
                 int split_lo, u32 tiers) {
    hipStream_t s0 = ctx->stream;
    zero_queue<<<1, 1, 0, s0>>>(d_queue);
    if (b.n_small == 0) return RBSC_OK;
    const bool fork = b.n_hi && (tiers & 1u) && b.n_small > b.n_hi && (tiers & 10u);
    hipStream_t st[4];
    HIP_TRY(tier_streams(ctx, fork, st));
    hipStream_t s = st[0];
    // graphs above BATCH_NMID nodes: the tuned shape and the full layout
    if (b.n_hi && (tiers & 1u)) {
        const unsigned grid = (unsigned)std::max(1, std::min<int>(n_wg, (int)b.n_hi));
        switch (threads) {
            case 256: launch_bt<256, BATCH_NMAX>(s, grid, b, 0, b.n_hi, max_rounds, d_labels, d_colors, d_sizes, d_blocks, d_rounds, d_queue, split_lo, b.rb[0]); break;
            case 1024: launch_bt<1024, BATCH_NMAX>(s, grid, b, 0, b.n_hi, max_rounds, d_labels, d_colors, d_sizes, d_blocks, d_rounds, d_queue, split_lo, b.rb[0]); break;
            default: launch_bt<512, BATCH_NMAX>(s, grid, b, 0, b.n_hi, max_rounds, d_labels, d_colors, d_sizes, d_blocks, d_rounds, d_queue, split_lo, b.rb[0]); break;
        }
    }
    s = st[1];
    // graphs above BATCH_NLOW nodes: half the layout, two workgroups per CU
    if (b.n_mid > b.n_hi && (tiers & 2u)) {
        const unsigned grid =
            (unsigned)std::max(1, std::min<int>(2 * b.cus, (int)(b.n_mid - b.n_hi)));
        launch_bt<MID_THREADS, BATCH_NMID>(s, grid, b, b.n_hi, b.n_mid, max_rounds, d_labels,
                                           d_colors, d_sizes, d_blocks, d_rounds, d_queue + 1,
                                           split_lo, b.rb[1]);
    }
    s = st[2];
    // graphs above BATCH_NTINY nodes: a quarter of the layout, four per CU
    if (b.n_low > b.n_mid && (tiers & 4u)) {
        const unsigned grid =
            (unsigned)std::max(1, std::min<int>(4 * b.cus, (int)(b.n_low - b.n_mid)));
        launch_bt<LOW_THREADS, BATCH_NLOW>(s, grid, b, b.n_mid, b.n_low, max_rounds, d_labels,
                                           d_colors, d_sizes, d_blocks, d_rounds, d_queue + 2,
                                           split_lo, b.rb[2]);
    }
    s = st[3];
    // the others: an eighth of the layout, eight workgroups per CU
    if (b.n_small > b.n_low && (tiers & 8u)) {
        const unsigned grid =
            (unsigned)std::max(1, std::min<int>(8 * b.cus, (int)(b.n_small - b.n_low)));
        launch_bt<TINY_THREADS, BATCH_NTINY>(s, grid, b, b.n_low, b.n_small, max_rounds, d_labels,
                                             d_colors, d_sizes, d_blocks, d_rounds, d_queue + 3,
                                             split_lo, b.rb[3]);
    }
    HIP_TRY(tier_join(ctx, fork));
    HIP_TRY(hipGetLastError());
    return RBSC_OK;
}

// ------------------------------------------------------------------ host
struct rbsc_dbatch {
    u32 n_graphs = 0;
    u64 total_nodes = 0, total_lines = 0;
    std::vector<u64> node_off;          // n_graphs + 1
    std::vector<u32> small, big, empty; // graph ids (empty: no nodes, never launched)
    std::vector<rbsc_dgraph*> big_g;    // resident big graphs (same order as big)
    BatchDevice dev;
    u32 *d_labels = nullptr, *d_colors = nullptr, *d_blocks = nullptr, *d_rounds = nullptr,
        *d_queue = nullptr;
    u32* d_sizes = nullptr;             // block sizes, allocated by the first refine asking for them
    int threads = 512, wgs = 0;
    u64 bytes = 0;
};

// The batch's buffers come from ctx_alloc (a serving loop takes and gives
// back one batch per launch while other refinements run).
static void dbatch_release(rbsc_ctx* ctx, rbsc_dbatch* b) {
    for (rbsc_dgraph* g : b->big_g) rbsc_graph_free(ctx, g);
    hipStream_t s = ctx ? ctx->stream : nullptr;
    for (void* p : {(void*)b->dev.desc, (void*)b->dev.order, (void*)b->dev.rp, (void*)b->dev.dst,
                    (void*)b->dev.wt, (void*)b->dev.init, (void*)b->d_labels, (void*)b->d_colors,
                    (void*)b->d_blocks, (void*)b->d_rounds, (void*)b->d_queue, (void*)b->d_sizes})
        ctx_free(p, s);
}

extern "C" int rbsc_batch_upload(rbsc_ctx* ctx, const rbsc_batch* in, rbsc_dbatch** out) {
    if (!ctx || !in || !out || !in->node_offsets || !in->line_offsets) {
        rbsc_set_error("rbsc_batch_upload: null argument");
        return RBSC_ERR_INVALID;
    }
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    const u32 G = in->n_graphs;
    const u64 NT = in->node_offsets[G], LT = in->line_offsets[G];
    if (in->node_offsets[0] != 0 || in->line_offsets[0] != 0 || (LT && (!in->src || !in->dst))) {
        rbsc_set_error("rbsc_batch_upload: offsets must start at 0 and lines need src and dst");
        return RBSC_ERR_INVALID;
    }
    for (u32 g = 0; g < G; ++g)
        if (in->node_offsets[g + 1] < in->node_offsets[g] || in->line_offsets[g + 1] < in->line_offsets[g] ||
            in->node_offsets[g + 1] - in->node_offsets[g] >= 0xFFFFFFFFull) {
            rbsc_set_error("rbsc_batch_upload: offsets of graph %u not monotone", g);
            return RBSC_ERR_INVALID;
        }
    rbsc_dbatch* b = new rbsc_dbatch();
    b->n_graphs = G;
    b->total_nodes = NT;
    b->total_lines = LT;
    b->node_off.assign(in->node_offsets, in->node_offsets + G + 1);
    // Small graphs get compacted row-offset and line arrays (line_base and
    // rp_base count small graphs only); big graphs keep their own uploads.
    std::vector<BatchDesc> desc(G);
    u64 small_rp = 0, small_lines = 0;
    for (u32 g = 0; g < G; ++g) {
        const u64 n = in->node_offsets[g + 1] - in->node_offsets[g];
        desc[g].node_base = in->node_offsets[g];
        desc[g].n = (u32)n;
        if (n == 0) {
            b->empty.push_back(g);  // no nodes: 0 blocks, 0 rounds, set on the host
        } else if (n <= BATCH_NMAX) {
            desc[g].line_base = small_lines;
            desc[g].rp_base = small_rp;
            small_lines += in->line_offsets[g + 1] - in->line_offsets[g];
            small_rp += n + 1;
            b->small.push_back(g);
        } else {
            b->big.push_back(g);
        }
    }
    std::vector<u32> rp(std::max<u64>(small_rp, 1), 0);
    std::vector<unsigned short> dst16(std::max<u64>(small_lines, 1));
    std::vector<u64> wt(in->weight ? small_lines : 0);
    // Validate and build local CSR of every small graph (parallel over graphs).
    std::vector<int> bad(G, 0);
    {
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                std::vector<u32> cnt;
                std::vector<unsigned char> seen;
                for (u64 gi = t; gi < G; gi += nt) {
                    const u32 g = (u32)gi;
                    const u32 n = desc[g].n;
                    if (n > BATCH_NMAX) continue;  // validated by rbsc_graph_upload
                    const u64 a = in->line_offsets[g], z = in->line_offsets[g + 1];
                    cnt.assign(n + 1, 0);
                    seen.assign(n, 0);
                    for (u64 e = a; e < z; ++e) {
                        const u32 sv = in->src[e], dv = in->dst[e];
                        if (sv >= n || dv >= n) { bad[g] = 1; break; }
                        if (e == a || in->src[e - 1] != sv) {
                            if (seen[sv]) { bad[g] = 2; break; }
                            seen[sv] = 1;
                        }
                        cnt[sv + 1]++;
                    }
                    if (bad[g]) continue;
                    u32* r = rp.data() + desc[g].rp_base;
                    r[0] = 0;
                    for (u32 u = 0; u < n; ++u) r[u + 1] = r[u] + cnt[u + 1];
                    const u64 lb = desc[g].line_base;
                    std::vector<u32> pos(r, r + n);
                    for (u64 e = a; e < z; ++e) {
                        const u32 o = pos[in->src[e]]++;
                        dst16[lb + o] = (unsigned short)in->dst[e];
                        if (in->weight) wt[lb + o] = in->weight[e];
                    }
                }
            });
        for (auto& x : th) x.join();
    }
    for (u32 g = 0; g < G; ++g)
        if (bad[g]) {
            rbsc_set_error("graph %u of the batch: %s", g,
                           bad[g] == 1 ? "node id out of range" : "lines of a source are not grouped");
            delete b;
            return RBSC_ERR_INPUT;
        }
    std::sort(b->small.begin(), b->small.end(),
              [&](u32 x, u32 y) { return desc[x].n > desc[y].n || (desc[x].n == desc[y].n && x < y); });
    b->dev.n_hi = 0;
    while (b->dev.n_hi < b->small.size() && desc[b->small[b->dev.n_hi]].n > BATCH_NMID) ++b->dev.n_hi;
    b->dev.n_mid = b->dev.n_hi;
    while (b->dev.n_mid < b->small.size() && desc[b->small[b->dev.n_mid]].n > BATCH_NLOW) ++b->dev.n_mid;
    b->dev.n_low = b->dev.n_mid;
    while (b->dev.n_low < b->small.size() && desc[b->small[b->dev.n_low]].n > BATCH_NTINY) ++b->dev.n_low;
    (void)hipDeviceGetAttribute(&b->dev.cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    // Big graphs go through the single-graph path.
    for (u32 g : b->big) {
        rbsc_graph one{};
        one.n_nodes = desc[g].n;
        one.n_edges = in->line_offsets[g + 1] - in->line_offsets[g];
        one.src = in->src + in->line_offsets[g];
        one.dst = in->dst + in->line_offsets[g];
        one.weight = in->weight ? in->weight + in->line_offsets[g] : nullptr;
        one.init_color = in->init_color ? in->init_color + in->node_offsets[g] : nullptr;
        rbsc_dgraph* dg = nullptr;
        const int rc = rbsc_graph_upload(ctx, &one, &dg);
        if (rc) {
            dbatch_release(ctx, b);
            delete b;
            return rc;
        }
        b->big_g.push_back(dg);
    }
    hipStream_t s = ctx->stream;
    auto up = [&](void** dptr, const void* src, size_t bytes) -> int {
        HIP_TRY(ctx_alloc(dptr, std::max<size_t>(bytes, 8), s));
        if (bytes && src) HIP_TRY(hipMemcpyAsync(*dptr, src, bytes, hipMemcpyHostToDevice, s));
        b->bytes += bytes;
        return RBSC_OK;
    };
    int rc = RBSC_OK;
    if (!rc) rc = up((void**)&b->dev.desc, desc.data(), sizeof(BatchDesc) * G);
    if (!rc) rc = up((void**)&b->dev.order, b->small.data(), sizeof(u32) * b->small.size());
    if (!rc) rc = up((void**)&b->dev.rp, rp.data(), sizeof(u32) * small_rp);
    if (!rc) rc = up((void**)&b->dev.dst, dst16.data(), sizeof(unsigned short) * small_lines);
    if (!rc && in->weight) rc = up((void**)&b->dev.wt, wt.data(), sizeof(u64) * small_lines);
    if (!rc && in->init_color) rc = up((void**)&b->dev.init, in->init_color, sizeof(u32) * NT);
    b->dev.n_small = (u32)b->small.size();
    if (!rc) rc = up((void**)&b->d_labels, nullptr, sizeof(u32) * NT);
    if (!rc) rc = up((void**)&b->d_colors, nullptr, sizeof(u32) * NT);
    if (!rc) rc = up((void**)&b->d_blocks, nullptr, sizeof(u32) * G);
    if (!rc) rc = up((void**)&b->d_rounds, nullptr, sizeof(u32) * G);
    if (!rc) rc = up((void**)&b->d_queue, nullptr, 4 * sizeof(u32));
    if (!rc && hipStreamSynchronize(s) != hipSuccess) {
        rbsc_set_error("batch upload failed");
        rc = RBSC_ERR_HIP;
    }
    if (rc) {
        dbatch_release(ctx, b);
        delete b;
        return rc;
    }
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    b->wgs = cus;
    *out = b;
    return RBSC_OK;
}

extern "C" int rbsc_batch_free(rbsc_ctx* ctx, rbsc_dbatch* b) {
    if (!b) return RBSC_OK;
    if (ctx) (void)hipSetDevice(ctx->device);
    dbatch_release(ctx, b);
    delete b;
    return RBSC_OK;
}

static int batch_run(rbsc_ctx* ctx, rbsc_dbatch* b, const rbsc_opts* opts, u32* blocks_out,
                     u32* rounds_out, u32* labels_out, u32* colors_out, u32* sizes_out,
                     rbsc_result* total, int threads, int wgs) {
    hipStream_t s = ctx->stream;
    const u32 max_rounds = opts ? opts->max_rounds : 0;
    auto t0 = std::chrono::steady_clock::now();
    // Labels (and block sizes) into buffers from rbsc_host_alloc are written
    // by the kernels themselves, so their transfer overlaps the refinement.
    u32* direct = labels_out
        ? static_cast<u32*>(host_device_ptr(labels_out, sizeof(u32) * b->total_nodes))
        : nullptr;
    u32* direct_sizes = sizes_out
        ? static_cast<u32*>(host_device_ptr(sizes_out, sizeof(u32) * b->total_nodes))
        : nullptr;
    if (sizes_out && !direct_sizes && !b->d_sizes)
        HIP_TRY(ctx_alloc(&b->d_sizes, sizeof(u32) * std::max<u64>(b->total_nodes, 1), s));
    u32* d_sizes = sizes_out ? (direct_sizes ? direct_sizes : b->d_sizes) : nullptr;
    HIP_TRY(hipEventRecord(ctx->ev0, s));
    const int split_lo = (opts && (opts->flags & RBSC_FLAG_SPLIT16)) ? 48 : 32;
    int rc = batch_launch(ctx, b->dev, max_rounds, direct ? direct : b->d_labels, b->d_colors,
                          d_sizes, b->d_blocks, b->d_rounds, b->d_queue, threads, wgs, split_lo);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ctx->ev1, s));
    const u32 G = b->n_graphs;
    std::vector<u32> blocks(G, 0), rounds(G, 0);
    HIP_TRY(hipMemcpyAsync(blocks.data(), b->d_blocks, sizeof(u32) * G, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(rounds.data(), b->d_rounds, sizeof(u32) * G, hipMemcpyDeviceToHost, s));
    if (labels_out && !direct)
        HIP_TRY(hipMemcpyAsync(labels_out, b->d_labels, sizeof(u32) * b->total_nodes,
                               hipMemcpyDeviceToHost, s));
    if (colors_out)
        HIP_TRY(hipMemcpyAsync(colors_out, b->d_colors, sizeof(u32) * b->total_nodes,
                               hipMemcpyDeviceToHost, s));
    if (sizes_out && !direct_sizes)
        HIP_TRY(hipMemcpyAsync(sizes_out, b->d_sizes, sizeof(u32) * b->total_nodes,
                               hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    double dev_ms = ms;
    for (u32 g : b->empty) blocks[g] = rounds[g] = 0;
    // graphs too large for LDS: the single-graph path, one by one
    for (size_t i = 0; i < b->big.size(); ++i) {
        const u32 g = b->big[i];
        rbsc_dgraph* dg = b->big_g[i];
        RefineParams prm;
        prm.max_rounds = max_rounds;
        prm.exact_colours = colors_out != nullptr;  // raw colours of the last round
        RefineOut o;
        rc = refine_core(ctx, dg, prm, &o);
        if (rc) return rc;
        dev_ms += o.device_ms;
        blocks[g] = o.K;
        rounds[g] = o.rounds;
        if (sizes_out) {  // before the partition, which reuses the colours' arrays
            rc = device_block_sizes(ctx, dg, o.K, sizes_out + b->node_off[g]);
            if (rc) return rc;
        }
        if (labels_out || colors_out) {
            rc = device_partition(ctx, dg, labels_out ? labels_out + b->node_off[g] : nullptr,
                                  colors_out ? colors_out + b->node_off[g] : nullptr);
            if (rc) return rc;
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (blocks_out) memcpy(blocks_out, blocks.data(), sizeof(u32) * G);
    if (rounds_out) memcpy(rounds_out, rounds.data(), sizeof(u32) * G);
    if (total) {
        rbsc_result r{};
        for (u32 g = 0; g < G; ++g) {
            r.nodes += blocks[g];
            r.rounds = std::max(r.rounds, rounds[g]);
        }
        r.done = 1;
        r.passes = 1;
        r.device_ms = dev_ms;
        r.time_s = std::chrono::duration<double>(t1 - t0).count();
        r.edges_per_round = b->total_lines;
        *total = r;
    }
    return RBSC_OK;
}

extern "C" int rbsc_batch_refine(rbsc_ctx* ctx, rbsc_dbatch* b, const rbsc_opts* opts,
                                 uint32_t* blocks_out, uint32_t* rounds_out,
                                 uint32_t* labels_out, uint32_t* colors_out,
                                 rbsc_result* total) {
    if (!ctx || !b) {
        rbsc_set_error("rbsc_batch_refine: null argument");
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    return batch_run(ctx, b, opts, blocks_out, rounds_out, labels_out, colors_out, nullptr, total,
                     b->threads, b->wgs);
}

extern "C" int rbsc_batch_refine_ex(rbsc_ctx* ctx, rbsc_dbatch* b, const rbsc_opts* opts,
                                    const rbsc_batch_outputs* out, rbsc_result* total) {
    if (!ctx || !b) {
        rbsc_set_error("rbsc_batch_refine_ex: null argument");
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    const rbsc_batch_outputs none{};
    const rbsc_batch_outputs& o = out ? *out : none;
    return batch_run(ctx, b, opts, o.blocks, o.rounds, o.labels, o.colors, o.block_sizes, total,
                     b->threads, b->wgs);
}

// Time one launch of the tiers in `tiers` (second of two runs; the first
// warms the caches).
static int time_tiers(rbsc_ctx* ctx, rbsc_dbatch* b, int threads, int wgs, u32 tiers, double* ms_out) {
    double sum = 0;
    for (int rep = 0; rep < 2; ++rep) {
        HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
        const int rc = batch_launch(ctx, b->dev, 0, b->d_labels, nullptr, nullptr, b->d_blocks,
                                    b->d_rounds, b->d_queue, threads, wgs, 32, tiers);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
        HIP_TRY(hipEventSynchronize(ctx->ev1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        if (rep) sum += ms;
    }
    *ms_out = sum;
    return RBSC_OK;
}

extern "C" int rbsc_batch_autotune(rbsc_ctx* ctx, rbsc_dbatch* b, uint32_t* threads_out,
                                   uint32_t* wgs_out, double* best_ms_out) {
    if (!ctx || !b) {
        rbsc_set_error("rbsc_batch_autotune: null argument");
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    // 1. the largest tier's workgroup width and count, timed with every tier
    //    (the tiers share the device: tiers 1 and 3 run beside 0 and 2)
    double best = 1e30;
    int bt = b->threads, bw = b->wgs;
    for (int threads : {256, 512, 1024}) {
        for (int mult : {1, 2, 4}) {
            double ms = 0;
            const int rc = time_tiers(ctx, b, threads, cus * mult, 0xFu, &ms);
            if (rc) return rc;
            if (ms < best) {
                best = ms;
                bt = threads;
                bw = cus * mult;
            }
        }
    }
    b->threads = bt;
    b->wgs = bw;
    // 2. the LDS sort's digit width (8 or 4 bits) of each tier
    for (int t = 0; t < 4; ++t) {
        double ms_best = 1e30;
        int rb_best = b->dev.rb[t];
        for (int rb : {8, 4}) {
            b->dev.rb[t] = rb;
            double ms = 0;
            const int rc = time_tiers(ctx, b, bt, bw, 1u << t, &ms);
            if (rc) return rc;
            if (ms < ms_best) {
                ms_best = ms;
                rb_best = rb;
            }
        }
        b->dev.rb[t] = rb_best;
    }
    double total = 0;
    const int rc = time_tiers(ctx, b, bt, bw, 0xFu, &total);
    if (rc) return rc;
    if (threads_out) *threads_out = (uint32_t)bt;
    if (wgs_out) *wgs_out = (uint32_t)bw;
    if (best_ms_out) *best_ms_out = total;
    return RBSC_OK;
}

extern "C" int rbsc_batch_launch_info(const rbsc_dbatch* b, uint32_t* info) {
    if (!b || !info) {
        rbsc_set_error("rbsc_batch_launch_info: null argument");
        return RBSC_ERR_INVALID;
    }
    info[0] = (uint32_t)b->threads;
    info[1] = (uint32_t)b->wgs;
    for (int t = 0; t < 4; ++t) info[2 + t] = (uint32_t)b->dev.rb[t];
    const u32 n[4] = {b->dev.n_hi, b->dev.n_mid - b->dev.n_hi, b->dev.n_low - b->dev.n_mid,
                      b->dev.n_small - b->dev.n_low};
    for (int t = 0; t < 4; ++t) info[6 + t] = n[t];
    return RBSC_OK;
}

extern "C" int rbsc_batch_set_digits(rbsc_dbatch* b, const uint32_t* digit_bits) {
    if (!b || !digit_bits) {
        rbsc_set_error("rbsc_batch_set_digits: null argument");
        return RBSC_ERR_INVALID;
    }
    for (int t = 0; t < 4; ++t)
        if (digit_bits[t] != 4 && digit_bits[t] != 8) {
            rbsc_set_error("rbsc_batch_set_digits: digit widths must be 4 or 8 bits");
            return RBSC_ERR_INVALID;
        }
    for (int t = 0; t < 4; ++t) b->dev.rb[t] = (int)digit_bits[t];
    return RBSC_OK;
}

extern "C" int rbsc_batch_set_launch(rbsc_ctx* ctx, rbsc_dbatch* b, uint32_t threads,
                                     uint32_t wgs) {
    if (!ctx || !b || (threads != 256 && threads != 512 && threads != 1024) || wgs == 0) {
        rbsc_set_error("rbsc_batch_set_launch: threads must be 256, 512 or 1024 and wgs > 0");
        return RBSC_ERR_INVALID;
    }
    b->threads = (int)threads;
    b->wgs = (int)wgs;
    return RBSC_OK;
}

extern "C" int rbsc_refine_batch(rbsc_ctx* ctx, const rbsc_batch* batch, const rbsc_opts* opts,
                                 uint32_t* blocks_out, uint32_t* rounds_out,
                                 uint32_t* labels_out, rbsc_result* total) {
    rbsc_dbatch* b = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    int rc = rbsc_batch_upload(ctx, batch, &b);
    if (rc) return rc;
    rc = rbsc_batch_refine(ctx, b, opts, blocks_out, rounds_out, labels_out, nullptr, total);
    rbsc_batch_free(ctx, b);
    if (!rc && total)
        total->time_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

// ------------------------------------------- batches in device memory, no allocation
// rbsc_batch_refine_device refines a batch whose arrays the caller already
// holds in device memory and allocates nothing: the launch layout (graph
// descriptors, the tier order, local row offsets, 16-bit heads) and the
// queue counters are built in `scratch`, a fixed array of this library on
// every device it runs on.  Calls on one device share that array, so calls
// that overlap on one device race (one at a time per device).
namespace {
constexpr u32 ND_MAX_GRAPHS = 1u << 15;
constexpr u64 ND_MAX_NODES = 1ull << 21;
constexpr u64 ND_MAX_LINES = 1ull << 22;
// Byte offsets inside the scratch array.
constexpr u64 ND_CTL = 0;                                      // 64 u32 of counters
constexpr u64 ND_DESC = 256;
constexpr u64 ND_ORDER = ND_DESC + sizeof(BatchDesc) * ND_MAX_GRAPHS;
constexpr u64 ND_RP = ND_ORDER + 4ull * ND_MAX_GRAPHS;
constexpr u64 ND_DST = ND_RP + 4ull * (ND_MAX_NODES + ND_MAX_GRAPHS);
constexpr u64 ND_BYTES = ND_DST + 2ull * ND_MAX_LINES + 256;
// Control words: [0..3] tier queues, [4..7] graphs per tier, [8..15] tier
// ranges (begin, end), [16..19] per-tier fill counters, [20] error flags.
constexpr int NDC_QUEUE = 0, NDC_COUNT = 4, NDC_RANGE = 8, NDC_FILL = 16, NDC_ERR = 20;
constexpr u32 NDE_OFFSETS = 1u, NDE_TOO_BIG = 2u, NDE_ID = 4u, NDE_ORDER = 8u;

static __device__ __align__(256) u64 scratch[ND_BYTES / 8];

std::mutex g_nd_mu;
void* g_nd_ptr[64] = {};  // scratch's device address, per device

__device__ __forceinline__ u32 nd_tier(u32 n) {
    return n > BATCH_NMID ? 0u : n > BATCH_NLOW ? 1u : n > BATCH_NTINY ? 2u : 3u;
}

__global__ void nd_clear_kernel(u32* ctl) {
    if (threadIdx.x < 64) ctl[threadIdx.x] = 0;
}

// One thread per graph: descriptor, tier count; graphs without nodes are
// answered here (0 blocks, 0 rounds) and never queued.
__global__ void nd_desc_kernel(const u64* __restrict__ node_off, const u64* __restrict__ line_off,
                               u32 G, u64 n_nodes, u64 n_lines, BatchDesc* __restrict__ desc,
                               u32* __restrict__ ctl, u32* __restrict__ blocks,
                               u32* __restrict__ rounds) {
    const u32 g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const u64 a = node_off[g], z = node_off[g + 1], la = line_off[g], lz = line_off[g + 1];
    u32 err = 0;
    if (z < a || lz < la || z > n_nodes || lz > n_lines) err |= NDE_OFFSETS;
    if (g == 0 && (a != 0 || la != 0)) err |= NDE_OFFSETS;
    if (g == G - 1 && (z != n_nodes || lz != n_lines)) err |= NDE_OFFSETS;
    const u64 n = err ? 0 : z - a;
    if (n > BATCH_NMAX) err |= NDE_TOO_BIG;
    BatchDesc d;
    d.node_base = a;
    d.line_base = la;
    d.rp_base = a + g;
    d.n = err ? 0u : (u32)n;
    d.pad = 0;
    desc[g] = d;
    if (err) atomicOr(&ctl[NDC_ERR], err);
    if (d.n == 0) {
        blocks[g] = 0;
        rounds[g] = 0;
    } else {
        atomicAdd(&ctl[NDC_COUNT + nd_tier(d.n)], 1u);
    }
}

__global__ void nd_ranges_kernel(u32* ctl) {
    if (threadIdx.x != 0) return;
    u32 b = 0;
    for (int t = 0; t < 4; ++t) {
        ctl[NDC_RANGE + 2 * t] = b;
        b += ctl[NDC_COUNT + t];
        ctl[NDC_RANGE + 2 * t + 1] = b;
    }
}

// Each graph with nodes takes a place in its tier's range (order inside a
// tier does not change any graph's result).
__global__ void nd_order_kernel(const BatchDesc* __restrict__ desc, u32 G, u32* __restrict__ ctl,
                                u32* __restrict__ order) {
    const u32 g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const u32 n = desc[g].n;
    if (n == 0) return;
    const u32 t = nd_tier(n);
    order[ctl[NDC_RANGE + 2 * t] + atomicAdd(&ctl[NDC_FILL + t], 1u)] = g;
}

// One workgroup per graph: local row offsets from lines sorted by source
// (rp[u] = first local line of u, rp[n] = line count), heads as 16-bit
// local ids, and the ids and order validated.  Writes stay inside the
// graph's own rp range and line range whatever the input holds.
__global__ __launch_bounds__(256) void nd_rows_kernel(const BatchDesc* __restrict__ desc,
                                                      const u64* __restrict__ line_off,
                                                      const u32* __restrict__ src,
                                                      const u32* __restrict__ dst,
                                                      u32* __restrict__ rp_all,
                                                      unsigned short* __restrict__ dst16,
                                                      u32* __restrict__ ctl) {
    const u32 g = blockIdx.x;
    const BatchDesc d = desc[g];
    const u32 n = d.n;
    if (n == 0) return;
    const u64 la = d.line_base, L = line_off[g + 1] - la;
    u32* rp = rp_all + d.rp_base;
    u32 err = 0;
    for (u64 i = threadIdx.x; i < L; i += blockDim.x) {
        const u64 e = la + i;
        const u32 sv = src[e], dv = dst[e];
        const long long ps = i ? (long long)src[e - 1] : -1;
        if (sv >= n || dv >= n) err |= NDE_ID;
        if ((long long)sv < ps) err |= NDE_ORDER;
        dst16[e] = (unsigned short)(dv < n ? dv : 0u);
        const long long s_hi = sv < n ? (long long)sv : (long long)n - 1;
        if (i == 0 || (long long)sv != ps)
            for (long long u = ps + 1; u <= s_hi; ++u) rp[u] = (u32)i;
        if (i == L - 1)
            for (long long u = s_hi + 1; u <= (long long)n; ++u) rp[u] = (u32)L;
    }
    if (L == 0)
        for (u32 u = threadIdx.x; u <= n; u += blockDim.x) rp[u] = 0;
    if (err) atomicOr(&ctl[NDC_ERR], err);
}
}  // namespace

int batch_device_scratch_init(int device) {
    if (device < 0 || device >= 64) {
        rbsc_set_error("device %d beyond the library's device table", device);
        return RBSC_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(g_nd_mu);
    if (g_nd_ptr[device]) return RBSC_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        rbsc_set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                       hipGetErrorString(e));
        return RBSC_ERR_HIP;
    }
    g_nd_ptr[device] = p;
    return RBSC_OK;
}

extern "C" int rbsc_batch_refine_device(rbsc_ctx* ctx, const rbsc_batch* batch, uint64_t n_nodes,
                                        uint64_t n_lines, const rbsc_opts* opts,
                                        uint32_t* blocks, uint32_t* rounds, uint32_t* labels,
                                        uint32_t* colors) {
    if (!ctx || !batch || !blocks || !rounds || !batch->node_offsets || !batch->line_offsets ||
        (n_lines && (!batch->src || !batch->dst))) {
        rbsc_set_error("rbsc_batch_refine_device: null argument");
        return RBSC_ERR_INVALID;
    }
    const u32 G = batch->n_graphs;
    if (G == 0) return RBSC_OK;
    if (G > ND_MAX_GRAPHS || n_nodes > ND_MAX_NODES || n_lines > ND_MAX_LINES) {
        rbsc_set_error("rbsc_batch_refine_device: batch of %u graphs, %llu nodes, %llu lines "
                       "exceeds the fixed scratch (at most %u graphs, %llu nodes, %llu lines)",
                       G, (unsigned long long)n_nodes, (unsigned long long)n_lines, ND_MAX_GRAPHS,
                       (unsigned long long)ND_MAX_NODES, (unsigned long long)ND_MAX_LINES);
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_nd_mu);
        if (ctx->device >= 0 && ctx->device < 64) base = g_nd_ptr[ctx->device];
    }
    if (!base) {
        const int rc = batch_device_scratch_init(ctx->device);
        if (rc) return rc;
        std::lock_guard<std::mutex> lk(g_nd_mu);
        base = g_nd_ptr[ctx->device];
    }
    unsigned char* sp = static_cast<unsigned char*>(base);
    u32* ctl = reinterpret_cast<u32*>(sp + ND_CTL);
    BatchDevice b;
    b.desc = reinterpret_cast<BatchDesc*>(sp + ND_DESC);
    b.order = reinterpret_cast<u32*>(sp + ND_ORDER);
    b.rp = reinterpret_cast<u32*>(sp + ND_RP);
    b.dst = reinterpret_cast<unsigned short*>(sp + ND_DST);
    b.wt = const_cast<u64*>(batch->weight);
    b.init = const_cast<u32*>(batch->init_color);
    b.n_small = G;
    (void)hipDeviceGetAttribute(&b.cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    hipStream_t s = ctx->stream;
    const u32 max_rounds = opts ? opts->max_rounds : 0;
    const int split_lo = (opts && (opts->flags & RBSC_FLAG_SPLIT16)) ? 48 : 32;
    const unsigned gg = (G + 255) / 256;
    nd_clear_kernel<<<1, 64, 0, s>>>(ctl);
    nd_desc_kernel<<<gg, 256, 0, s>>>(batch->node_offsets, batch->line_offsets, G, n_nodes, n_lines,
                                      b.desc, ctl, blocks, rounds);
    nd_ranges_kernel<<<1, 64, 0, s>>>(ctl);
    nd_order_kernel<<<gg, 256, 0, s>>>(b.desc, G, ctl, b.order);
    nd_rows_kernel<<<G, 256, 0, s>>>(b.desc, batch->line_offsets, batch->src, batch->dst, b.rp,
                                     b.dst, ctl);
    // every tier launched at its full grid; its jobs come from the device
    // ranges, so nothing is read back before the refinement
    const u32* err = ctl + NDC_ERR;
    u32* q = ctl + NDC_QUEUE;
    const u32* rg = ctl + NDC_RANGE;
    const int cus = b.cus;
    hipStream_t st[4];
    HIP_TRY(tier_streams(ctx, true, st));
    launch_bt<512, BATCH_NMAX>(st[0], cus, b, 0, 0, max_rounds, labels, colors, nullptr, blocks,
                               rounds, q, split_lo, b.rb[0], rg, err);
    launch_bt<MID_THREADS, BATCH_NMID>(st[1], 2 * cus, b, 0, 0, max_rounds, labels, colors,
                                       nullptr, blocks, rounds, q + 1, split_lo, b.rb[1], rg + 2,
                                       err);
    launch_bt<LOW_THREADS, BATCH_NLOW>(st[2], 4 * cus, b, 0, 0, max_rounds, labels, colors,
                                       nullptr, blocks, rounds, q + 2, split_lo, b.rb[2], rg + 4,
                                       err);
    launch_bt<TINY_THREADS, BATCH_NTINY>(st[3], 8 * cus, b, 0, 0, max_rounds, labels, colors,
                                         nullptr, blocks, rounds, q + 3, split_lo, b.rb[3], rg + 6,
                                         err);
    HIP_TRY(tier_join(ctx, true));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ctx->h_scalar + 14, err, sizeof(u32), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const u32 e = ctx->h_scalar[14];
    if (e) {
        rbsc_set_error("rbsc_batch_refine_device: %s",
                       (e & NDE_OFFSETS) ? "offsets not monotone or not ending at the totals"
                       : (e & NDE_TOO_BIG) ? "a graph has more than 5120 nodes"
                       : (e & NDE_ID) ? "a node id is out of its graph's range"
                       : "lines of a graph are not sorted by source");
        return (e & (NDE_OFFSETS | NDE_TOO_BIG)) ? RBSC_ERR_INVALID : RBSC_ERR_INPUT;
    }
    return RBSC_OK;
}
