// sort.hip -- onesweep LSD radix sort for gfx950 (64-wide waves, LDS-staged
// scatter, decoupled look-back with epoch-stamped status words).
//
// Per pass, one workgroup of 256 threads takes a SORT_TILE-key tile through an
// atomic ticket (so every tile it waits on is already resident), ranks its
// keys stably with wave ballots, publishes its per-digit counts, looks back
// over earlier tiles for its per-digit global offset, reorders the tile in
// LDS by digit and writes it out in digit runs (coalesced).
#include "sort.h"
#include "../../include/rbsc.h"
#include <algorithm>

namespace {

constexpr int THREADS = 256;
constexpr int WAVES = THREADS / RBSC_WAVE;
constexpr int ITEMS = RBSC_SORT_ITEMS;
constexpr int TILE = THREADS * ITEMS;  // == SORT_TILE
constexpr int RADIX = 256;
static_assert(TILE == (int)SORT_TILE, "tile size mismatch");
static_assert(THREADS == RADIX, "one thread per digit in the look-back");

constexpr int HIST_ITEMS = 16;

// Histograms of up to SORT_HD 8-bit digits of the sort (positions
// [0, npass) from bit_begin) into ghist.
__global__ __launch_bounds__(THREADS) void hist_kernel(const u64* __restrict__ keys, u64 n,
                                                       int bit_begin, int npass,
                                                       u32* __restrict__ ghist) {
    __shared__ u32 lh[SORT_HWORDS];
    for (int i = threadIdx.x; i < SORT_HWORDS; i += THREADS) lh[i] = 0;
    __syncthreads();
    // HIST_ITEMS keys per thread per step, all loaded before any is counted
    // (one load in flight per thread would leave the pass latency-bound).
    const u64 stride = (u64)gridDim.x * THREADS * HIST_ITEMS;
    for (u64 b = (u64)blockIdx.x * THREADS * HIST_ITEMS; b < n; b += stride) {
        u64 k[HIST_ITEMS];
#pragma unroll
        for (int j = 0; j < HIST_ITEMS; ++j) {
            const u64 i = b + (u64)j * THREADS + threadIdx.x;
            k[j] = i < n ? keys[i] : 0;
        }
#pragma unroll
        for (int j = 0; j < HIST_ITEMS; ++j) {
            if (b + (u64)j * THREADS + threadIdx.x >= n) continue;
            for (int p = 0; p < npass; ++p)
                sort_hist_add(lh, p, (u32)(k[j] >> (bit_begin + 8 * p)) & 255u);
        }
    }
    __syncthreads();
    sort_hist_flush(lh, npass, ghist);
}

// Histograms of the digits of [bit_begin, bit_begin + 8 * npass) into
// ghist[p * 256 + d], SORT_HD positions per launch.
static void launch_hist(const u64* keys, u64 n, int bit_begin, int npass, u32* ghist,
                        hipStream_t s) {
    const unsigned blocks =
        (unsigned)std::min<u64>((n + THREADS * HIST_ITEMS - 1) / (THREADS * HIST_ITEMS), 2048);
    for (int p0 = 0; p0 < npass; p0 += SORT_HD)
        hist_kernel<<<blocks, THREADS, 0, s>>>(keys, n, bit_begin + 8 * p0,
                                               std::min(SORT_HD, npass - p0), ghist + 256 * p0);
}

// Exclusive scan over a 256-thread block, one value per thread.
__device__ __forceinline__ u64 block_excl_scan_256(u64 v, u64* s_wave, u64* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    u64 x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        u64 y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    u64 off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        if (w < wave) off += s_wave[w];
        tot += s_wave[w];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

// Per-pass exclusive digit offsets from the histograms (one block); also
// clears the histograms for the next sort.
__global__ __launch_bounds__(THREADS) void scan_hist_kernel(u32* __restrict__ ghist, int npass,
                                                            u64* __restrict__ base) {
    __shared__ u64 s_wave[WAVES];
    for (int p = 0; p < npass; ++p) {
        u64 tot;
        const u64 v = ghist[p * RADIX + threadIdx.x];
        base[p * RADIX + threadIdx.x] = block_excl_scan_256(v, s_wave, &tot);
        ghist[p * RADIX + threadIdx.x] = 0;
    }
}

__global__ void fill_u64_kernel(u64* p, u64 n, u64 v) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// VM: 0 keys only; 1 values from vin; 2 values made from a u32 array and the
// position, vin_col[i] << 32 | i (the first pass of a sort of nodes in id
// order, values (colour << 32) | node).
template <int VM>
__global__ __launch_bounds__(THREADS, RBSC_SORT_OCC) void onesweep_kernel(
    const u64* __restrict__ kin, const u64* __restrict__ vin, u64* __restrict__ kout,
    u64* __restrict__ vout, u64 n, u32 shift, const u64* __restrict__ digit_base,
    u64* status, u64* ticket, u64 ticket_base, u32 epoch, u32* err) {
    constexpr bool HAS_VALS = VM == 1 || VM == 2;
    __shared__ u64 s_stage[TILE];
    __shared__ unsigned char s_digit[TILE];
    __shared__ u32 s_whist[WAVES][RADIX];
    __shared__ u32 s_tstart[RADIX];
    __shared__ u64 s_gbase[RADIX];
    __shared__ u64 s_wave[WAVES];
    __shared__ u64 s_tile;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd((unsigned long long*)ticket, 1ULL) - ticket_base;
    for (int i = tid; i < WAVES * RADIX; i += THREADS) (&s_whist[0][0])[i] = 0;
    __syncthreads();
    const u64 tile = s_tile;
    const u64 tile_base = tile * TILE;
    const u64 wave_base = tile_base + (u64)wave * (RBSC_WAVE * ITEMS);

    u64 key[ITEMS];
    u32 rank[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const u64 idx = wave_base + (u64)j * RBSC_WAVE + lane;
        key[j] = idx < n ? kin[idx] : ~0ULL;
        // VM 3: packed key, the bits from `shift` up with the index below
        if (VM == 3 && idx < n) key[j] = ((key[j] >> shift) << shift) | idx;
    }
    // Values are loaded now so their latency hides behind ranking and look-back.
    u64 val[HAS_VALS ? ITEMS : 1];
    if (VM == 1) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const u64 idx = wave_base + (u64)j * RBSC_WAVE + lane;
            val[j] = idx < n ? vin[idx] : 0;
        }
    } else if (VM == 2) {
        const u32* col = reinterpret_cast<const u32*>(vin);
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const u64 idx = wave_base + (u64)j * RBSC_WAVE + lane;
            val[j] = idx < n ? ((u64)col[idx] << 32) | idx : 0;
        }
    }
    if (tile_base + TILE <= n) wave_rank<8, ITEMS, true>(key, rank, shift, wave_base, n, s_whist[wave]);
    else wave_rank<8, ITEMS, false>(key, rank, shift, wave_base, n, s_whist[wave]);
    __syncthreads();

    // Per digit (thread == digit): counts across waves, tile-local starts.
    const u32 d = tid;
    u32 total = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        const u32 c = s_whist[w][d];
        s_whist[w][d] = total;
        total += c;
    }
    u64 tile_sum;
    const u32 tstart = (u32)block_excl_scan_256(total, s_wave, &tile_sum);
    s_tstart[d] = tstart;

    // Decoupled look-back for this digit.
    u64* st = status + tile * RADIX + d;
    u64 excl = 0;
    if (tile == 0) {
        lb_store(st, lb_pack(LB_PREFIX, epoch, total));
    } else {
        lb_store(st, lb_pack(LB_AGG, epoch, total));
        long long p = (long long)tile - 1;
        u32 spins = 0;
        while (p >= 0) {
            const u64 s = lb_load(status + (u64)p * RADIX + d);
            const u64 state = s >> LB_STATE_SHIFT;
            const u32 ep = (u32)((s >> LB_EPOCH_SHIFT) & LB_EPOCH_MASK);
            if (state == 0 || ep != (epoch & LB_EPOCH_MASK)) {
                if (++spins > LB_SPIN_LIMIT) {
                    atomicOr(err, KERR_SPIN);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += s & LB_COUNT_MASK;
            if (state == LB_PREFIX) break;
            --p;
        }
        lb_store(st, lb_pack(LB_PREFIX, epoch, excl + total));
    }
    s_gbase[d] = digit_base[d] + excl - tstart;
    __syncthreads();

    // Stage keys in LDS by digit, then write digit runs.
    // rank[j] becomes the item's LDS position (one register array for both)
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const u64 idx = wave_base + (u64)j * RBSC_WAVE + lane;
        if (idx < n) {
            const u32 dj = (u32)(key[j] >> shift) & 255u;
            rank[j] += s_tstart[dj] + s_whist[wave][dj];
            s_stage[rank[j]] = key[j];
            s_digit[rank[j]] = (unsigned char)dj;
        }
    }
    __syncthreads();
    const u64 rem = n - tile_base;
    const u32 tile_n = rem < (u64)TILE ? (u32)rem : (u32)TILE;
    for (u32 i = tid; i < tile_n; i += THREADS) kout[s_gbase[s_digit[i]] + i] = s_stage[i];
    if (HAS_VALS) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const u64 idx = wave_base + (u64)j * RBSC_WAVE + lane;
            if (idx < n) s_stage[rank[j]] = val[j];
        }
        __syncthreads();
        for (u32 i = tid; i < tile_n; i += THREADS) vout[s_gbase[s_digit[i]] + i] = s_stage[i];
    }
}

}  // namespace

u32 lb_next_epoch(SortWorkspace* ws, hipStream_t s) {
    ws->epoch = (ws->epoch + 1) & (u32)LB_EPOCH_MASK;
    if (ws->epoch == 0) {
        // Wrapped: clear every status word so no stale word matches.
        const u64 n = ws->status_tiles * RADIX;
        if (n) fill_u64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(ws->status, n, 0);
        const u64 m = ws->extra_words;
        if (m && ws->extra_status)
            fill_u64_kernel<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(ws->extra_status, m, 0);
        ws->epoch = 1;
    }
    return ws->epoch;
}

int sort_ws_init(SortWorkspace* ws, u64 max_n, u32* err_flag, hipStream_t s) {
    const u64 tiles = (max_n + TILE - 1) / TILE + 1;
    if (ws->status && ws->status_tiles >= tiles) return RBSC_OK;
    ctx_free(ws->status, s);
    ws->status = nullptr;
    if (!ws->hist) {
        HIP_TRY(ctx_alloc(&ws->hist, sizeof(u32) * 8 * RADIX, s));
        HIP_TRY(ctx_alloc(&ws->digit_base, sizeof(u64) * 8 * RADIX, s));
        HIP_TRY(ctx_alloc(&ws->ticket, sizeof(u64), s));
        HIP_TRY(hipMemsetAsync(ws->hist, 0, sizeof(u32) * 8 * RADIX, s));
        HIP_TRY(hipMemsetAsync(ws->ticket, 0, sizeof(u64), s));
        ws->ticket_total = 0;
    }
    HIP_TRY(ctx_alloc(&ws->status, sizeof(u64) * tiles * RADIX, s));
    HIP_TRY(hipMemsetAsync(ws->status, 0, sizeof(u64) * tiles * RADIX, s));
    ws->status_tiles = tiles;
    ws->err = err_flag;
    return RBSC_OK;
}

void sort_ws_free(SortWorkspace* ws, hipStream_t s) {
    ctx_free(ws->hist, s);
    ctx_free(ws->digit_base, s);
    ctx_free(ws->status, s);
    ctx_free(ws->ticket, s);
    *ws = SortWorkspace();
}

int radix_sort_u64(SortWorkspace* ws, u64* keys[2], u64* vals[2], u64 n, int bit_begin,
                   int bit_end, int* which, hipStream_t s, bool hist_ready) {
    *which = 0;
    if (n == 0 || bit_end <= bit_begin) return RBSC_OK;
    if ((bit_end - bit_begin) % 8 != 0 || bit_end > 64 || bit_begin < 0) {
        rbsc_set_error("radix_sort_u64: bit range [%d,%d) not a multiple of 8", bit_begin,
                       bit_end);
        return RBSC_ERR_INVALID;
    }
    const u64 tiles = (n + TILE - 1) / TILE;
    if (tiles + 1 > ws->status_tiles) {
        rbsc_set_error("radix_sort_u64: workspace sized for fewer keys");
        return RBSC_ERR_INVALID;
    }
    const int npass = (bit_end - bit_begin) / 8;
    if (!hist_ready) launch_hist(keys[0], n, bit_begin, npass, ws->hist, s);
    scan_hist_kernel<<<1, THREADS, 0, s>>>(ws->hist, npass, ws->digit_base);
    int cur = 0;
    for (int p = 0; p < npass; ++p) {
        const u32 ep = lb_next_epoch(ws, s);
        const u32 shift = (u32)(bit_begin + 8 * p);
        if (vals) {
            onesweep_kernel<1><<<(unsigned)tiles, THREADS, 0, s>>>(
                keys[cur], vals[cur], keys[cur ^ 1], vals[cur ^ 1], n, shift,
                ws->digit_base + p * RADIX, ws->status, ws->ticket, ws->ticket_total, ep, ws->err);
        } else {
            onesweep_kernel<0><<<(unsigned)tiles, THREADS, 0, s>>>(
                keys[cur], nullptr, keys[cur ^ 1], nullptr, n, shift,
                ws->digit_base + p * RADIX, ws->status, ws->ticket, ws->ticket_total, ep, ws->err);
        }
        ws->ticket_total += tiles;
        cur ^= 1;
    }
    HIP_TRY(hipGetLastError());
    *which = cur;
    return RBSC_OK;
}

int radix_sort_nodes(SortWorkspace* ws, const u64* key0, const u32* col, u64* keys[2],
                     u64* vals[2], u64 n, int bit_begin, int bit_end, hipStream_t s) {
    if (n == 0 || bit_end <= bit_begin) return RBSC_OK;
    if ((bit_end - bit_begin) % 8 != 0 || bit_end > 64 || bit_begin < 0 || n >= (1ull << 32)) {
        rbsc_set_error("radix_sort_nodes: bit range [%d,%d) or %llu keys not supported",
                       bit_begin, bit_end, (unsigned long long)n);
        return RBSC_ERR_INVALID;
    }
    const u64 tiles = (n + TILE - 1) / TILE;
    if (tiles + 1 > ws->status_tiles) {
        rbsc_set_error("radix_sort_nodes: workspace sized for fewer keys");
        return RBSC_ERR_INVALID;
    }
    const int npass = (bit_end - bit_begin) / 8;
    launch_hist(key0, n, bit_begin, npass, ws->hist, s);
    scan_hist_kernel<<<1, THREADS, 0, s>>>(ws->hist, npass, ws->digit_base);
    // pass p writes buffer (npass - 1 - p) & 1, so the last pass writes buffer 0
    for (int p = 0; p < npass; ++p) {
        const u32 ep = lb_next_epoch(ws, s);
        const u32 shift = (u32)(bit_begin + 8 * p);
        const int out = (npass - 1 - p) & 1;
        if (p == 0)
            onesweep_kernel<2><<<(unsigned)tiles, THREADS, 0, s>>>(
                key0, reinterpret_cast<const u64*>(col), keys[out], vals[out], n, shift,
                ws->digit_base, ws->status, ws->ticket, ws->ticket_total, ep, ws->err);
        else
            onesweep_kernel<1><<<(unsigned)tiles, THREADS, 0, s>>>(
                keys[out ^ 1], vals[out ^ 1], keys[out], vals[out], n, shift,
                ws->digit_base + p * RADIX, ws->status, ws->ticket, ws->ticket_total, ep, ws->err);
        ws->ticket_total += tiles;
    }
    HIP_TRY(hipGetLastError());
    return RBSC_OK;
}

int radix_sort_nodes_packed(SortWorkspace* ws, const u64* key0, u64* keys[2], u64 n, int bit_begin,
                            int bit_end, hipStream_t s) {
    if (n == 0 || bit_end <= bit_begin) return RBSC_OK;
    if ((bit_end - bit_begin) % 8 != 0 || bit_end > 64 || bit_begin < 0 ||
        bit_begin > 63 || (n >> bit_begin) != 0) {
        rbsc_set_error("radix_sort_nodes_packed: bit range [%d,%d) or %llu keys not supported",
                       bit_begin, bit_end, (unsigned long long)n);
        return RBSC_ERR_INVALID;
    }
    const u64 tiles = (n + TILE - 1) / TILE;
    if (tiles + 1 > ws->status_tiles) {
        rbsc_set_error("radix_sort_nodes_packed: workspace sized for fewer keys");
        return RBSC_ERR_INVALID;
    }
    const int npass = (bit_end - bit_begin) / 8;
    launch_hist(key0, n, bit_begin, npass, ws->hist, s);
    scan_hist_kernel<<<1, THREADS, 0, s>>>(ws->hist, npass, ws->digit_base);
    // pass p writes buffer (npass - 1 - p) & 1, so the last pass writes buffer 0
    for (int p = 0; p < npass; ++p) {
        const u32 ep = lb_next_epoch(ws, s);
        const u32 shift = (u32)(bit_begin + 8 * p);
        const int out = (npass - 1 - p) & 1;
        if (p == 0)
            onesweep_kernel<3><<<(unsigned)tiles, THREADS, 0, s>>>(
                key0, nullptr, keys[out], nullptr, n, shift, ws->digit_base, ws->status, ws->ticket,
                ws->ticket_total, ep, ws->err);
        else
            onesweep_kernel<0><<<(unsigned)tiles, THREADS, 0, s>>>(
                keys[out ^ 1], nullptr, keys[out], nullptr, n, shift, ws->digit_base + p * RADIX,
                ws->status, ws->ticket, ws->ticket_total, ep, ws->err);
        ws->ticket_total += tiles;
    }
    HIP_TRY(hipGetLastError());
    return RBSC_OK;
}
