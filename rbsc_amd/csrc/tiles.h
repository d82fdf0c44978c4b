// tiles.h -- line tiles over a CSR row array, balanced by lines.
//
// Build passes that visit every line of a graph run one workgroup per tile
// of LT_LINES consecutive lines instead of one thread per row: R-MAT rows
// are skewed (at scale 27 the longest holds about a million lines), and a
// thread walking such a row alone bounds the whole launch.  Inside a tile
// each thread holds LT_PER consecutive lines (blocked layout); the row of
// every line comes from the tile's row starts, flagged in LDS and scanned.
// Rows must be nonempty (row_ptr strictly increasing), as the stored rows
// of a graph and the rows of a raw upload are.
#pragma once
#include "common.h"

constexpr int LT_THREADS = 256;
constexpr int LT_PER = 8;
constexpr u64 LT_LINES = (u64)LT_THREADS * LT_PER;  // 2048

inline u64 lt_tiles(u64 E) { return (E + LT_LINES - 1) / LT_LINES; }

// first[t] = row holding line t * LT_LINES (t < n_tiles), first[n_tiles] =
// the row holding the last line.  row_ptr has R + 1 entries, row_ptr[R] = E;
// one thread per tile, a binary search each.
[[maybe_unused]] static __global__ void lt_first_row_kernel(const u64* __restrict__ row_ptr, u64 R, u64 E,
                                                            u64 n_tiles, u32* __restrict__ first) {
    const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > n_tiles) return;
    const u64 e = t < n_tiles ? t * LT_LINES : E - 1;
    u64 lo = 0, hi = R - 1;  // last row whose start is <= e
    while (lo < hi) {
        const u64 mid = lo + (hi - lo + 1) / 2;
        if (row_ptr[mid] <= e) lo = mid; else hi = mid - 1;
    }
    first[t] = (u32)lo;
}

struct LtLds {
    u32 flag[LT_LINES];       // 1 where a row starts (index = line - e0)
    u64 start[LT_LINES + 2];  // row_ptr[r0 + i], i <= r1 - r0 + 1
    u32 wsum[LT_THREADS / 64];
    u32 tot;
};

// Rows of this thread's lines: row[j] for line e0 + threadIdx.x * LT_PER + j
// (lines at or beyond e1 get the last row).  Also fills s->start with the
// starts of rows r0 .. r1 + 1 (entries past row R read as E).  Returns r1,
// the row of the tile's last line.  Call with all threads of the workgroup.
__device__ __forceinline__ u32 lt_rows(const u64* __restrict__ row_ptr, u64 R, u64 E,
                                       const u32* __restrict__ first, u64 t, LtLds* s,
                                       u32 row[LT_PER]) {
    const u64 e0 = t * LT_LINES;
    const u64 e1 = e0 + LT_LINES < E ? e0 + LT_LINES : E;
    const u32 r0 = first[t];
    const u32 r1 = first[t + 1] < R ? first[t + 1] : (u32)(R - 1);
    for (u32 i = threadIdx.x; i < LT_LINES; i += LT_THREADS) s->flag[i] = 0;
    __syncthreads();
    for (u64 r = (u64)r0 + threadIdx.x; r <= (u64)r1 + 1; r += LT_THREADS) {
        const u64 st = r <= R ? row_ptr[r] : E;
        s->start[r - r0] = st;
        if (r > r0 && st > e0 && st < e1) s->flag[st - e0] = 1;
    }
    __syncthreads();
    // inclusive count of row starts up to each line
    const int b = threadIdx.x * LT_PER;
    u32 f[LT_PER], c = 0;
#pragma unroll
    for (int j = 0; j < LT_PER; ++j) {
        f[j] = s->flag[b + j];
        c += f[j];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    u32 x = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const u32 y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) s->wsum[wave] = x;
    __syncthreads();
    u32 off = 0;
#pragma unroll
    for (int w = 0; w < LT_THREADS / 64; ++w)
        if (w < wave) off += s->wsum[w];
    u32 run = off + x - c;
    // the row of the tile's last line: r0 plus every row start in the tile
    u32 starts = 0;
#pragma unroll
    for (int w = 0; w < LT_THREADS / 64; ++w) starts += s->wsum[w];
    const u32 last = r0 + starts;
#pragma unroll
    for (int j = 0; j < LT_PER; ++j) {
        run += f[j];
        const u32 rr = r0 + run;
        row[j] = rr < last ? rr : last;
    }
    __syncthreads();
    return last;
}
