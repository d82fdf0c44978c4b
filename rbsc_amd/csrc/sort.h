// sort.h -- LSD onesweep radix sort of 64-bit keys (optionally with 64-bit
// values) for gfx950: one histogram launch, one scan launch, then one launch
// per 8-bit digit with decoupled look-back.  Replaces the reference's
// cub::DeviceRadixSort::SortPairs / SortKeys calls (rbsc.cu:386, 390, 409).
#pragma once
#include "common.h"

struct SortWorkspace {
    u32* hist = nullptr;        // [8][256] digit counts
    u64* digit_base = nullptr;  // [8][256] exclusive global offsets
    u64* status = nullptr;      // [tiles][256] look-back words
    u64 status_tiles = 0;
    u64* ticket = nullptr;      // monotone tile ticket counter
    u64 ticket_total = 0;       // tickets handed out so far (host mirror)
    u32 epoch = 0;              // look-back epoch (host mirror)
    u32* err = nullptr;         // kernel error flags
    // Another look-back status array stamped with this epoch counter (the
    // dense-rank pass); cleared with the sort's own on epoch wrap-around.
    u64* extra_status = nullptr;
    u64 extra_words = 0;
};

// Keys per thread and per onesweep tile (256 threads).  RBSC_SORT_ITEMS and
// RBSC_SORT_OCC (workgroups per CU asked of the compiler) are build-time
// tuning knobs.
// 32 x 256 = 8192-key tiles, two workgroups per CU (79 KB of LDS each, 245
// VGPRs with values, 181 without, no spills): the cost of a pass is mostly
// per tile (ranking, look-back, the staged write-out), so the largest tile
// two workgroups fit is fastest.  Measured on 33 M random keys, 4 passes:
// 0.97 ms keys only and 1.35 ms with values, against 1.09 / 1.44 ms for
// 6144-key tiles and 1.24 / 1.64 ms for 10240-key tiles at one workgroup
// per CU; 134 M keys: 3.52 / 4.98 against 3.98 / 5.38 ms.
#ifndef RBSC_SORT_ITEMS
#define RBSC_SORT_ITEMS 32
#endif
#ifndef RBSC_SORT_OCC
#define RBSC_SORT_OCC 2
#endif
constexpr u64 SORT_TILE = 256 * RBSC_SORT_ITEMS;

int sort_ws_init(SortWorkspace* ws, u64 max_n, u32* err_flag, hipStream_t s);
void sort_ws_free(SortWorkspace* ws, hipStream_t s);

// Sort (keys[0], vals[0]) by key bits [bit_begin, bit_end) (multiple of 8 wide),
// stably, ping-ponging with (keys[1], vals[1]).  vals may be null (keys only).
// Returns in *which the buffer index (0 or 1) that holds the result.
// hist_ready: ws->hist already holds the digit counts of keys[0] for this
// bit range (accumulated with sort_hist_add by the kernel that wrote them).
int radix_sort_u64(SortWorkspace* ws, u64* keys[2], u64* vals[2], u64 n, int bit_begin,
                   int bit_end, int* which, hipStream_t s, bool hist_ready = false);

// Stable sort of n < 2^32 nodes in id order by key0 bits [bit_begin,
// bit_end): the first pass reads key0[i] and makes the value
// (col[i] << 32) | i; the result is in (keys[0], vals[0]).
int radix_sort_nodes(SortWorkspace* ws, const u64* key0, const u32* col, u64* keys[2],
                     u64* vals[2], u64 n, int bit_begin, int bit_end, hipStream_t s);

// Stable sort of n < 2^bit_begin nodes in id order by key0 bits
// [bit_begin, bit_end), keys only: element i is (key0[i] with the bits below
// bit_begin replaced by i); the result is in keys[0].
int radix_sort_nodes_packed(SortWorkspace* ws, const u64* key0, u64* keys[2], u64 n, int bit_begin,
                            int bit_end, hipStream_t s);

// Digit histograms in LDS: SORT_HC copies per digit position, copy = lane
// mod SORT_HC, each padded to 257 words so equal digits of neighbouring lanes
// fall into different banks; many equal keys (the early rounds of a
// refinement) then contend on SORT_HC addresses instead of one.  Layout
// lh[p][copy][257] for up to SORT_HD digit positions per launch.
constexpr int SORT_HC = 8;
constexpr int SORT_HD = 4;
constexpr int SORT_HSTRIDE = 257;
constexpr int SORT_HWORDS = SORT_HD * SORT_HC * SORT_HSTRIDE;
__device__ __forceinline__ void sort_hist_add(u32* lh, int p, u32 d) {
    atomicAdd(&lh[(p * SORT_HC + (int)(threadIdx.x % SORT_HC)) * SORT_HSTRIDE + d], 1u);
}
// Adds the copies of digit positions [0, npass) to ghist[p * 256 + d].
__device__ __forceinline__ void sort_hist_flush(const u32* lh, int npass, u32* ghist) {
    for (int i = threadIdx.x; i < npass * 256; i += blockDim.x) {
        const int p = i >> 8, d = i & 255;
        u32 v = 0;
#pragma unroll
        for (int c = 0; c < SORT_HC; ++c) v += lh[(p * SORT_HC + c) * SORT_HSTRIDE + d];
        if (v) atomicAdd(&ghist[i], v);
    }
}

// Lanes of the wave (among those set in `among`, a ballot) whose low BITS
// digit bits equal this lane's: one ballot per bit, the mask kept as two
// 32-bit halves and narrowed by the ballot or its complement, chosen by the
// lane's own bit as a 0 / -1 mask from a signed bit extract of ~d (six VALU
// operations per bit).
template <int BITS>
__device__ __forceinline__ u64 wave_match(u32 d, u64 among) {
    u32 lo = (u32)among, hi = (u32)(among >> 32);
    const int nd = (int)~d;
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        const u32 nm = (u32)__builtin_amdgcn_sbfe(nd, b, 1);  // 0 if bit b is set, else -1
        const u64 bal = __ballot(nm == 0);
        lo &= (u32)bal ^ nm;
        hi &= (u32)(bal >> 32) ^ nm;
    }
    return ((u64)hi << 32) | lo;
}

// Stable wave ranks of N keys per lane by BITS-bit digit (key >> shift):
// rank[j] = earlier keys of this wave with the same digit (lower j, or the
// same j and a lower lane) plus that digit's count in whist before this
// call; whist is the wave's own LDS histogram, updated.  Every lane reads
// its digit's count in one LDS instruction, then the lowest lane of each
// digit group writes the count plus the group's size; a wave's LDS
// instructions execute in order, so every lane sees the count from before
// the group's update and the next key sees it after.  valid(j) says which
// keys take part (all of them in a full tile).
template <int BITS, int N, bool FULL>
__device__ __forceinline__ void wave_rank(const u64 (&key)[N], u32 (&rank)[N], u32 shift,
                                          u64 first, u64 n, u32* whist) {
    const int lane = threadIdx.x & 63;
    const u64 below_mask = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const bool valid = FULL || first + (u64)j * RBSC_WAVE + lane < n;
        const u32 d = (u32)(key[j] >> shift) & ((1u << BITS) - 1);
        const u64 peers = wave_match<BITS>(d, FULL ? ~0ULL : __ballot(valid));
        const u32 below = (u32)__popcll(peers & below_mask);
        u32 old = 0;
        if (valid) old = whist[d];
        if (valid && below == 0) whist[d] = old + (u32)__popcll(peers);
        rank[j] = old + below;
    }
}

// Epoch/ticket bookkeeping shared with the other look-back kernels.
u32 lb_next_epoch(SortWorkspace* ws, hipStream_t s);
