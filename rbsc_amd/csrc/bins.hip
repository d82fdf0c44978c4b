// bins.hip -- binned signature layout of a resident graph.
//
// The plain signature kernel gathers one colour per line at a random head.
// On a large graph every such gather is its own L2 request, and gathers of
// the many rarely-used ("cold") heads miss to the fabric, moving a 128-byte
// line for 4 useful bytes.  This layout computes the same signatures
//   sig(u) = sum over u's lines (u, v, w) of randomize(c(v) + r) * w  (mod 2^64)
// (rbsc.cu:80-87 set_values + rbsc.cu:366-377 ReduceByKey) over three
// classes of gathered heads (internal ids are in gathered in-degree order):
// dense (head < D), warm (D <= head < H) and cold (H <= head < gather
// limit).  Each round:
//
//   warm_gather_kernel   copies the colours of the warm lines into the
//                        refinement's buffer.  The warm heads are cut into
//                        segments of equal line counts, a few per XCD; each
//                        XCD gathers only from its own segments, whose
//                        colours its L2 holds.
//   cold_stage_kernel    writes each cold line's colour into the buffer.
//                        Bins are split into 8 groups, one per XCD; a
//                        group's cold lines, in head order, are cut into
//                        tiles of a fixed number of lines, one workgroup per
//                        tile (tile c of group g is workgroup 8c + g:
//                        workgroups are dealt to the XCDs round robin).  The
//                        workgroup reads the colours in head order into LDS
//                        and writes them out in slot order; a bin's cold
//                        slots are in head order, so a tile's lines for one
//                        bin fill one run of consecutive positions, and a
//                        wave's stores cover a few runs.  (RBSC_BIN_SLOT_ORDER
//                        selects cold_scatter_kernel instead: records kept in
//                        slot order, colour reads out of order, no LDS.)
//   bin_sig_kernel       one workgroup per chunk of a bin of 2^log_s rows:
//                        row sums in LDS, dense slots gathered in head order
//                        (a wave's 64 gathers fall into few cache lines),
//                        warm and cold slots read from the buffer, lines
//                        into heads that are never gathered folded in per
//                        row as dang_h * dsum[row].
//
// The layout is read-only during refinement: the per-round cold colours
// live in the refinement's context, so several contexts may refine one
// binned graph at the same time.
//
// Only the order of the additions differs from the plain kernel, and they are
// exact in Z/2^64: signatures are identical.
#include "internal.h"
#include "scan.h"
#include "tiles.h"
#include <algorithm>
#include <vector>

namespace {
constexpr int BLK = 256;
inline unsigned grid_for(u64 n) { return (unsigned)((n + BLK - 1) / BLK); }
inline int bitlen(u64 v) {
    int b = 0;
    while (b < 64 && (v >> b)) ++b;
    return b;
}
constexpr int SIG_THREADS = 512;
// Zeroed entries past the end of the slot arrays and scatter records (the
// last aligned group's loads stay in bounds).
constexpr u64 SLOT_RUN = 512;
// LDS of the staged scatter: one colour per line of a tile.
constexpr u64 STAGE_BYTES = 65536;
constexpr int STAGE_THREADS = 512;

// ---------------------------------------------------------------- build
// Per stored row: lines into hot heads, into cold heads, and the total
// weight of lines into heads that are never gathered.  Entry R is the zero
// that makes the exclusive scans end with the totals.
// Head of a line as gathered: with `rep` (no initial colours), a repeated
// row's node [R, U) always has its stored row's colour, so lines into it
// gather the stored row instead.
__device__ __forceinline__ u32 gathered_head(u32 v, const u32* __restrict__ rep, u32 R, u32 U) {
    return (rep && v >= R && v < U) ? rep[v - R] : v;
}

// Per stored row, in tiles of LT_LINES lines (rows are skewed): lines into
// hot heads, into cold heads, and the total weight of lines into heads that
// are never gathered, added to n_hot / n_cold / dsum (zeroed) with one
// atomic per row and tile; and per tile its hot and cold line counts, whose
// scans place each tile's records of a class (records go in line order).
template <bool HAS_WT>
__global__ __launch_bounds__(LT_THREADS) void bin_count_tile_kernel(
    const u64* __restrict__ row_ptr, const u32* __restrict__ dst, const u64* __restrict__ wt,
    u64 R, u64 E, u32 H, u32 GL, const u32* __restrict__ rep, u32 U,
    const u32* __restrict__ first, unsigned long long* __restrict__ n_hot,
    unsigned long long* __restrict__ n_cold, unsigned long long* __restrict__ dsum,
    u64* __restrict__ tile_hot, u64* __restrict__ tile_cold) {
    __shared__ LtLds s;
    __shared__ u32 acc_h[LT_LINES + 1], acc_c[LT_LINES + 1];
    __shared__ unsigned long long acc_d[LT_LINES + 1];
    __shared__ u32 tot[2];
    const u64 t = blockIdx.x;
    u32 row[LT_PER];
    const u32 last = lt_rows(row_ptr, R, E, first, t, &s, row);
    const u32 r0 = first[t], nr = last - r0 + 1;
    for (u32 i = threadIdx.x; i < nr; i += LT_THREADS) {
        acc_h[i] = acc_c[i] = 0;
        acc_d[i] = 0;
    }
    if (threadIdx.x < 2) tot[threadIdx.x] = 0;
    __syncthreads();
    const u64 b0 = t * LT_LINES + (u64)threadIdx.x * LT_PER;
    u32 h = 0, c = 0, th = 0, tc = 0;
    u64 d = 0;
    u32 cur = row[0];
#pragma unroll
    for (int j = 0; j < LT_PER; ++j) {
        const u64 e = b0 + j;
        if (e < E) {
            if (row[j] != cur) {
                if (h) atomicAdd(&acc_h[cur - r0], h);
                if (c) atomicAdd(&acc_c[cur - r0], c);
                if (d) atomicAdd(&acc_d[cur - r0], (unsigned long long)d);
                h = c = 0;
                d = 0;
                cur = row[j];
            }
            const u32 v = gathered_head(dst[e], rep, (u32)R, U);
            if (v < H) {
                ++h;
                ++th;
            } else if (v < GL) {
                ++c;
                ++tc;
            } else {
                d += HAS_WT ? wt[e] : 1;
            }
        }
    }
    if (h) atomicAdd(&acc_h[cur - r0], h);
    if (c) atomicAdd(&acc_c[cur - r0], c);
    if (d) atomicAdd(&acc_d[cur - r0], (unsigned long long)d);
    if (th) atomicAdd(&tot[0], th);
    if (tc) atomicAdd(&tot[1], tc);
    __syncthreads();
    for (u32 i = threadIdx.x; i < nr; i += LT_THREADS) {
        if (acc_h[i]) atomicAdd(&n_hot[r0 + i], (unsigned long long)acc_h[i]);
        if (acc_c[i]) atomicAdd(&n_cold[r0 + i], (unsigned long long)acc_c[i]);
        if (acc_d[i]) atomicAdd(&dsum[r0 + i], acc_d[i]);
    }
    if (threadIdx.x == 0) {
        tile_hot[t] = tot[0];
        tile_cold[t] = tot[1];
    }
}

// Line records of one class in line (row-major) order: key (bin, head),
// value (line << 16 | row in bin); a tile's records start at toff[t] and
// follow in line order inside it (a workgroup prefix count).
template <bool HOT>
__global__ __launch_bounds__(LT_THREADS) void bin_fill_tile_kernel(
    const u64* __restrict__ row_ptr, const u32* __restrict__ dst, u64 R, u64 E, u32 H, u32 GL,
    const u32* __restrict__ rep, u32 U, u32 log_s, int hb, const u32* __restrict__ first,
    const u64* __restrict__ toff, u64* __restrict__ keys, u64* __restrict__ vals) {
    __shared__ LtLds s;
    const u64 t = blockIdx.x;
    u32 row[LT_PER];
    lt_rows(row_ptr, R, E, first, t, &s, row);
    const u64 b0 = t * LT_LINES + (u64)threadIdx.x * LT_PER;
    u32 take[LT_PER], vv[LT_PER], cnt = 0;
#pragma unroll
    for (int j = 0; j < LT_PER; ++j) {
        const u64 e = b0 + j;
        take[j] = 0;
        vv[j] = 0;
        if (e < E) {
            const u32 v = gathered_head(dst[e], rep, (u32)R, U);
            vv[j] = v;
            take[j] = (HOT ? v < H : (v >= H && v < GL)) ? 1u : 0u;
        }
        cnt += take[j];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    u32 x = cnt;
#pragma unroll
    for (int dd = 1; dd < 64; dd <<= 1) {
        const u32 y = __shfl_up(x, dd);
        if (lane >= dd) x += y;
    }
    if (lane == 63) s.wsum[wave] = x;
    __syncthreads();
    u32 off = 0;
#pragma unroll
    for (int w = 0; w < LT_THREADS / 64; ++w)
        if (w < wave) off += s.wsum[w];
    u64 p = toff[t] + off + x - cnt;
    const u64 mask = (1ull << log_s) - 1;
#pragma unroll
    for (int j = 0; j < LT_PER; ++j) {
        if (take[j]) {
            const u64 e = b0 + j;
            const u64 u = row[j];
            keys[p] = ((u >> log_s) << hb) | vv[j];
            vals[p] = (e << 16) | (u & mask);
            ++p;
        }
    }
}

// Per bin: first hot / cold record in row-major order, slot count, hot count.
__global__ void bin_sizes_kernel(const u64* __restrict__ hot_off, const u64* __restrict__ cold_off,
                                 u64 R, u32 log_s, u64 n_bins, u64* __restrict__ hot_start,
                                 u64* __restrict__ cold_start, u64* __restrict__ bin_n,
                                 u64* __restrict__ bin_hot) {
    const u64 b = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > n_bins) return;
    if (b == n_bins) {
        bin_n[b] = 0;
        return;
    }
    const u64 r0 = b << log_s;
    const u64 r1 = std::min<u64>(r0 + (1ull << log_s), R);
    hot_start[b] = hot_off[r0];
    cold_start[b] = cold_off[r0];
    bin_hot[b] = hot_off[r1] - hot_off[r0];
    bin_n[b] = bin_hot[b] + (cold_off[r1] - cold_off[r0]);
}

// Sorted hot records into their slots (a bin's hot slots in head order).
template <bool HAS_WT>
__global__ void place_hot_kernel(const u64* __restrict__ keys, const u64* __restrict__ vals,
                                 u64 n, int hb, const u64* __restrict__ bin_slot,
                                 const u64* __restrict__ hot_start, const u64* __restrict__ wt,
                                 u32* __restrict__ slot_key, unsigned short* __restrict__ slot_row,
                                 u64* __restrict__ slot_wt) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 k = keys[i], v = vals[i];
    const u64 b = k >> hb;
    const u64 s = bin_slot[b] + (i - hot_start[b]);
    slot_key[s] = (u32)(k & ((1ull << hb) - 1));
    slot_row[s] = (unsigned short)(v & 0xFFFF);
    if (HAS_WT) slot_wt[s] = wt[v >> 16];
}

// Sorted cold records into their slots (after the bin's hot slots), and the
// scatter records: key (group, head), value slot << 32 | head.
// A record's target is its index in a refinement's cold-colour buffer
// (cold_pos[bin] + slot - the bin's first buffered (warm or cold) slot
// rounded down to 8), or its slot when cold_pos is null.
template <bool HAS_WT>
__global__ void place_cold_kernel(const u64* __restrict__ keys, const u64* __restrict__ vals,
                                  u64 n, int gb, const u64* __restrict__ bin_slot,
                                  const u64* __restrict__ hot_all,
                                  const u64* __restrict__ bin_dense,
                                  const u64* __restrict__ cold_start,
                                  const u32* __restrict__ grp_of_bin,
                                  const u64* __restrict__ cold_pos, const u64* __restrict__ wt,
                                  u32* __restrict__ slot_key,
                                  unsigned short* __restrict__ slot_row,
                                  u64* __restrict__ slot_wt, u64* __restrict__ k2,
                                  u64* __restrict__ v2) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 k = keys[i], v = vals[i];
    const u64 b = k >> gb;
    const u64 head = k & ((1ull << gb) - 1);
    const u64 s = bin_slot[b] + hot_all[b] + (i - cold_start[b]);
    const u64 c0 = bin_slot[b] + bin_dense[b];
    slot_key[s] = 0;
    slot_row[s] = (unsigned short)(v & 0xFFFF);
    if (HAS_WT) slot_wt[s] = wt[v >> 16];
    const u64 target = cold_pos ? cold_pos[b] + (s - (c0 & ~7ull)) : s;
    k2[i] = ((u64)grp_of_bin[b] << gb) | head;
    v2[i] = (target << 32) | head;
}

// Records in (group, head) order -> key (tile, slot) or (tile, position), where
// tile = 8 * (position in the group / chunk) + group.
__global__ void chunk_key_kernel(const u64* __restrict__ k2, const u64* __restrict__ v2, u64 n,
                                 int gb, const u64* __restrict__ grp_first, u64 chunk,
                                 int by_slot, u64* __restrict__ k3) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 g = k2[i] >> gb;
    const u64 t = ((i - grp_first[g]) / chunk) * BIN_GROUPS + g;
    k3[i] = (t << 32) | (by_slot ? (v2[i] >> 32) : (i & 0xFFFFFFFFull));
}

// Records sorted by tile -> compact arrays and the first record of every
// tile (tile_line[n_tiles] = n).
__global__ void csc_fill_kernel(const u64* __restrict__ k3, const u64* __restrict__ v3, u64 n,
                                u64 n_tiles, u32* __restrict__ csc_head,
                                u32* __restrict__ csc_slot, u64* __restrict__ tile_line) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 t = k3[i] >> 32;
    csc_head[i] = (u32)(v3[i] & 0xFFFFFFFFull);
    csc_slot[i] = (u32)(v3[i] >> 32);
    const u64 first = i == 0 ? 0 : (k3[i - 1] >> 32) + 1;
    for (u64 u = first; u <= t; ++u) tile_line[u] = i;
    if (i == n - 1)
        for (u64 u = t + 1; u <= n_tiles; ++u) tile_line[u] = n;
}

// Lines into each warm head (D <= head < H), from the sorted hot records.
__global__ void warm_hist_kernel(const u64* __restrict__ keys, u64 n, int hb, u32 D, u32 H,
                                 u32* __restrict__ cnt) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u32 head = (u32)(keys[i] & ((1ull << hb) - 1));
    if (head >= D && head < H) atomicAdd(cnt + (head - D), 1u);
}

// Per bin and segment boundary d[j] (j = 0..WARM_SEGS): the first of the
// bin's hot slots (sorted by head) whose head is >= d[j]; bin_dense = the
// slots below d[0] = D.
__global__ void warm_off_kernel(const u32* __restrict__ slot_key, const u64* __restrict__ bin_slot,
                                const u64* __restrict__ hot_all, u64 n_bins,
                                const u32* __restrict__ bound, u64* __restrict__ warm_off,
                                u64* __restrict__ bin_dense) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_bins * (WARM_SEGS + 1)) return;
    const u64 b = i / (WARM_SEGS + 1), j = i % (WARM_SEGS + 1);
    u64 lo = bin_slot[b], hi = lo + hot_all[b];
    const u32 d = bound[j];
    while (lo < hi) {
        const u64 mid = lo + (hi - lo) / 2;
        if (slot_key[mid] < d) lo = mid + 1;
        else hi = mid;
    }
    warm_off[i] = lo;
    if (j == 0) bin_dense[b] = lo - bin_slot[b];
}

// Staged scatter: records of every tile in head order (sorted) -> key
// (tile, target), value head-order index, to be sorted into slot order.
__global__ void stage_key_kernel(const u64* __restrict__ k3, const u64* __restrict__ v3, u64 n,
                                 u64* __restrict__ k4, u64* __restrict__ v4) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    k4[i] = ((k3[i] >> 32) << 32) | (v3[i] >> 32);
    v4[i] = i;
}

// Records sorted by (tile, target): targets in slot order, and each
// head-order record's rank inside its tile's slot order.
__global__ void stage_fill_kernel(const u64* __restrict__ k4, const u64* __restrict__ v4, u64 n,
                                  const u64* __restrict__ tile_line, u32* __restrict__ csc_slot,
                                  unsigned short* __restrict__ csc_perm) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 t = k4[i] >> 32;
    csc_slot[i] = (u32)(k4[i] & 0xFFFFFFFFull);
    csc_perm[v4[i]] = (unsigned short)(i - tile_line[t]);
}

// ---------------------------------------------------------------- rounds
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef u32 u32x4 __attribute__((ext_vector_type(4)));
typedef u64 u64x2 __attribute__((ext_vector_type(2)));

// Slot-order scatter (RBSC_BIN_SLOT_ORDER): writes the current colour of every
// cold line's head into the line's slot, one workgroup per tile (records
// [tile_line[t], tile_line[t + 1]), in slot order).  Records
// are read 8 at a time with 16-byte loads: thread i takes the aligned groups
// [8q, 8q + 8) for q = lo / 8 + i, + BLK, ...; records outside the tile (a
// neighbour's, or the zeroed padding past the end) are loaded, and their
// heads (valid ids or 0) read, but never stored.  The target is the line's
// index in the refinement's cold-colour buffer; the colour is the u32 raw
// colour on the raw layout, the u64 hashed colour on the hashed layout.
template <typename C>
__global__ __launch_bounds__(BLK) void cold_scatter_kernel(const u32* __restrict__ csc_head,
                                                           const u32* __restrict__ csc_slot,
                                                           const u64* __restrict__ tile_line,
                                                           const C* __restrict__ col,
                                                           C* __restrict__ out) {
    const u64 t = blockIdx.x;
    const u64 lo = tile_line[t], hi = tile_line[t + 1];
    const u64 q1 = (hi + 7) / 8;
    for (u64 q = lo / 8 + threadIdx.x; q < q1; q += BLK) {
        const u32x4 ha = *reinterpret_cast<const u32x4*>(csc_head + 8 * q);
        const u32x4 hb = *reinterpret_cast<const u32x4*>(csc_head + 8 * q + 4);
        const u32x4 sa = *reinterpret_cast<const u32x4*>(csc_slot + 8 * q);
        const u32x4 sb = *reinterpret_cast<const u32x4*>(csc_slot + 8 * q + 4);
        const u32 hd[8] = {ha.x, ha.y, ha.z, ha.w, hb.x, hb.y, hb.z, hb.w};
        const u32 sl[8] = {sa.x, sa.y, sa.z, sa.w, sb.x, sb.y, sb.z, sb.w};
        C c[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) c[k] = col[hd[k]];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const u64 j = 8 * q + k;
            if (j >= lo && j < hi) out[sl[k]] = c[k];
        }
    }
}

// Warm gather: workgroup i runs on XCD x = i mod 8 (workgroups are dealt to
// the XCDs round robin); its waves take x's work items (runs of warm slots
// whose heads lie in x's segments) one wave per item, so the XCD's gathers
// fall into its own segments' colours, which its L2 holds.  Keys are read
// and colours stored in runs; each lane keeps 4 slots in flight per whole
// step.
constexpr int WARM_THREADS = 512;
constexpr int WARM_WAVES = WARM_THREADS / 64;
template <class C>
__global__ __launch_bounds__(WARM_THREADS) void warm_gather_kernel(
    const WarmItem* __restrict__ items, WarmFirst first, const u32* __restrict__ slot_key,
    const C* __restrict__ col, C* __restrict__ out, u32 per_xcd) {
    const u32 x = blockIdx.x % BIN_GROUPS;
    const u32 lane = threadIdx.x & 63;
    const u32 i1 = first.f[x + 1];
    const u32 stride = per_xcd * WARM_WAVES;
    for (u32 i = first.f[x] + (blockIdx.x / BIN_GROUPS) * WARM_WAVES + threadIdx.x / 64; i < i1;
         i += stride) {
        const WarmItem it = items[i];
        const u32* key = slot_key + it.lo;
        C* o = out + it.dst;
        u32 base = lane;
        for (; base + 3 * 64 < it.n; base += 4 * 64) {  // whole steps: straight-line
            u32 k4[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) k4[k] = key[base + k * 64];
            C c[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) c[k] = col[k4[k]];
#pragma unroll
            for (int k = 0; k < 4; ++k) o[base + k * 64] = c[k];
        }
        for (; base < it.n; base += 64) o[base] = col[key[base]];
    }
}

// Staged scatter, one workgroup per tile of at most STAGE_BYTES / sizeof(C)
// lines: the
// tile's colours are read in head order (a wave's 64 reads fall into a few
// cache lines) into LDS at each line's rank in slot order, then written out
// in slot order.  A tile's lines for one bin are one run of consecutive
// targets (a bin's cold slots are in head order and a tile is a head range),
// so a wave's 64 stores cover a few runs instead of 64 scattered lines.
// Each thread keeps 4 lines in flight per whole step.
template <class C>
__global__ __launch_bounds__(STAGE_THREADS) void cold_stage_kernel(
    const u32* __restrict__ csc_head, const unsigned short* __restrict__ csc_perm,
    const u32* __restrict__ csc_slot, const u64* __restrict__ tile_line,
    const C* __restrict__ col, C* __restrict__ out) {
    __shared__ C stage[STAGE_BYTES / sizeof(C)];
    const u64 lo = tile_line[blockIdx.x];
    const u32 n = (u32)(tile_line[blockIdx.x + 1] - lo);
    const u32* head = csc_head + lo;
    const unsigned short* perm = csc_perm + lo;
    const u32* slot = csc_slot + lo;
    constexpr u32 STEP = 4 * STAGE_THREADS;
    for (u32 base = threadIdx.x; base < n; base += STEP) {
        if (base + 3 * STAGE_THREADS < n) {  // whole step: straight-line
            u32 h[4];
            unsigned short p[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                h[k] = head[base + k * STAGE_THREADS];
                p[k] = perm[base + k * STAGE_THREADS];
            }
            C c[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) c[k] = col[h[k]];
#pragma unroll
            for (int k = 0; k < 4; ++k) stage[p[k]] = c[k];
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const u32 j = base + k * STAGE_THREADS;
                if (j < n) stage[perm[j]] = col[head[j]];
            }
        }
    }
    __syncthreads();
    for (u32 base = threadIdx.x; base < n; base += STEP) {
        if (base + 3 * STAGE_THREADS < n) {
            u32 t[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) t[k] = slot[base + k * STAGE_THREADS];
#pragma unroll
            for (int k = 0; k < 4; ++k) out[t[k]] = stage[base + k * STAGE_THREADS];
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const u32 j = base + k * STAGE_THREADS;
                if (j < n) out[slot[j]] = stage[j];
            }
        }
    }
}

// Rows of bins served by several chunks start from zero (chunks add).
__global__ void bin_zero_kernel(const u32* __restrict__ zero_bins, u64 n_zero, u32 log_s, u64 R,
                                u64* __restrict__ sig) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 k = i >> log_s;
    if (k >= n_zero) return;
    const u64 row = ((u64)zero_bins[k] << log_s) + (i & ((1ull << log_s) - 1));
    if (row < R) sig[row] = 0;
}

// Adds the hashed colours of slots [a, e) into the LDS row sums: slots below
// cold0 (dense) hold heads whose colours are gathered, the others (warm and
// cold) take their colours from the refinement's buffer (index s + cbase;
// aligned groups of 8 stay aligned), which the warm gather and the scatter
// fill every round; the graph's own arrays are only read, so several
// refinements of one graph may run at once.  Slots are read 8 at a time with
// 16-byte loads (aligned groups, slots outside [a, e) loaded but not added).
// (Measured no faster: lanes taking 64 consecutive slots per load
// instruction, two groups in flight per thread, dense slots packed to 4
// bytes.)
// Raw layout (HASHED false): col and the buffer hold u32 raw colours, and
// every line hashes randomize(c + salt).  Hashed layout: col and the buffer
// hold the round's u64 hashed colours, and no line hashes.
template <bool HAS_WT, bool HASHED>
__device__ __forceinline__ void bin_accumulate(u64 a, u64 e, u64 cold0, u64 cbase,
                                               const u32* __restrict__ slot_key,
                                               const unsigned short* __restrict__ slot_row,
                                               const u64* __restrict__ slot_wt,
                                               const void* __restrict__ colv,
                                               const void* __restrict__ cold_buf, u64 salt,
                                               u64* sums) {
    const u64 q1 = (e + 7) / 8;
    for (u64 q = a / 8 + threadIdx.x; q < q1; q += SIG_THREADS) {
        const u16x8 r8 = *reinterpret_cast<const u16x8*>(slot_row + 8 * q);
        const u32 row[8] = {r8.s0, r8.s1, r8.s2, r8.s3, r8.s4, r8.s5, r8.s6, r8.s7};
        u64 h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const u64 g0 = 8 * q;
        if (HASHED) {
            const u64* col = static_cast<const u64*>(colv);
            if (g0 < cold0) {  // the group holds dense slots
                const u32x4 ka = *reinterpret_cast<const u32x4*>(slot_key + g0);
                const u32x4 kb = *reinterpret_cast<const u32x4*>(slot_key + g0 + 4);
                const u32 key[8] = {ka.x, ka.y, ka.z, ka.w, kb.x, kb.y, kb.z, kb.w};
#pragma unroll
                for (int k = 0; k < 8; ++k) h[k] = col[g0 + k < cold0 ? key[k] : 0u];
            }
            // Cold slots of this chunk in the group: [max(g0, cold0), min(g0 + 8, e)).
            const u64 c_lo = g0 > cold0 ? g0 : cold0;
            if (c_lo < e && c_lo < g0 + 8) {
                const u64x2* cv =
                    reinterpret_cast<const u64x2*>(static_cast<const u64*>(cold_buf) + (g0 + cbase));
                const u64x2 v0 = cv[0], v1 = cv[1], v2 = cv[2], v3 = cv[3];
                const u64 c8[8] = {v0.x, v0.y, v1.x, v1.y, v2.x, v2.y, v3.x, v3.y};
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (g0 + k >= cold0) h[k] = c8[k];
            }
        } else {
            const u32* col = static_cast<const u32*>(colv);
            u32 key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (g0 < cold0) {  // keys only for groups that hold gathered slots
                const u32x4 ka = *reinterpret_cast<const u32x4*>(slot_key + g0);
                const u32x4 kb = *reinterpret_cast<const u32x4*>(slot_key + g0 + 4);
                key[0] = ka.x; key[1] = ka.y; key[2] = ka.z; key[3] = ka.w;
                key[4] = kb.x; key[5] = kb.y; key[6] = kb.z; key[7] = kb.w;
            }
            // Cold slots of this chunk in the group: [max(g0, cold0), min(g0 + 8, e)),
            // colours from the refinement's buffer (index s + cbase).
            const u64 c_lo = g0 > cold0 ? g0 : cold0;
            if (c_lo < e && c_lo < g0 + 8) {
                const u32* cc = static_cast<const u32*>(cold_buf) + (g0 + cbase);
                const u32x4 ca = *reinterpret_cast<const u32x4*>(cc);
                const u32x4 cb = *reinterpret_cast<const u32x4*>(cc + 4);
                const u32 c8[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (g0 + k >= cold0) key[k] = c8[k];
            }
            if (g0 < cold0) {  // the group holds dense slots
                u32 g[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) g[k] = col[g0 + k < cold0 ? key[k] : 0u];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (g0 + k < cold0) key[k] = g[k];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) h[k] = (u64)key[k];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const u64 s = g0 + k;
            if (s >= a && s < e) {
                u64 v = HASHED ? h[k] : rbsc_randomize(h[k] + salt);
                if (HAS_WT) v *= slot_wt[s];
                atomicAdd((unsigned long long*)&sums[row[k]], (unsigned long long)v);
            }
        }
    }
}

template <int LOG_S, bool HAS_WT, bool HASHED>
__global__ __launch_bounds__(SIG_THREADS) void bin_sig_kernel(
    const BinChunk* __restrict__ chunks, const u64* __restrict__ bin_slot,
    const u64* __restrict__ bin_hot, const u64* __restrict__ cold_pos,
    const u32* __restrict__ slot_key, const unsigned short* __restrict__ slot_row,
    const u64* __restrict__ slot_wt, const void* __restrict__ cold_buf,
    const u64* __restrict__ dsum, const void* __restrict__ col, u64 salt, u64 dang_h, u64 R,
    u64* __restrict__ sig) {
    constexpr u32 S = 1u << LOG_S;
    __shared__ u64 sums[S];
    const BinChunk ch = chunks[blockIdx.x];
    const u64 row0 = (u64)ch.bin << LOG_S;
    const u32 nrows = (u32)(R - row0 < S ? R - row0 : S);
    const bool first = (ch.flags & CHUNK_FIRST) != 0;
    for (u32 i = threadIdx.x; i < S; i += SIG_THREADS)
        sums[i] = (first && i < nrows) ? dang_h * dsum[row0 + i] : 0ull;
    __syncthreads();
    const u64 cold0 = bin_slot[ch.bin] + bin_hot[ch.bin];
    // cold-buffer index of slot s: cold_pos + s - (cold0 rounded down to 8),
    // in Z/2^64 (only ever applied to the bin's cold groups).
    const u64 cbase = cold_pos[ch.bin] - (cold0 & ~7ull);
    bin_accumulate<HAS_WT, HASHED>(ch.lo, ch.hi, cold0, cbase, slot_key, slot_row, slot_wt, col,
                                   cold_buf, salt, sums);
    __syncthreads();
    const bool only = (ch.flags & CHUNK_ONLY) != 0;
    for (u32 i = threadIdx.x; i < nrows; i += SIG_THREADS) {
        if (only) sig[row0 + i] = sums[i];
        else if (sums[i])  // adding 0 changes nothing (rows without lines in this chunk)
            atomicAdd((unsigned long long*)&sig[row0 + i], (unsigned long long)sums[i]);
    }
}

template <int LOG_S, bool HASHED>
void launch_bin_sig2(hipStream_t s, const rbsc_bins* b, const void* col, const void* cold_buf,
                     u64 salt, u64 dang_h, u64 R, u64* sig) {
    if (b->slot_wt)
        bin_sig_kernel<LOG_S, true, HASHED><<<(unsigned)b->n_chunks, SIG_THREADS, 0, s>>>(
            b->chunks, b->bin_slot, b->bin_hot, b->cold_pos, b->slot_key, b->slot_row, b->slot_wt,
            cold_buf, b->dsum, col, salt, dang_h, R, sig);
    else
        bin_sig_kernel<LOG_S, false, HASHED><<<(unsigned)b->n_chunks, SIG_THREADS, 0, s>>>(
            b->chunks, b->bin_slot, b->bin_hot, b->cold_pos, b->slot_key, b->slot_row, nullptr,
            cold_buf, b->dsum, col, salt, dang_h, R, sig);
}
template <int LOG_S>
void launch_bin_sig(hipStream_t s, const rbsc_bins* b, const void* col, const void* cold_buf,
                    u64 salt, u64 dang_h, u64 R, u64* sig) {
    if (b->hashed) launch_bin_sig2<LOG_S, true>(s, b, col, cold_buf, salt, dang_h, R, sig);
    else launch_bin_sig2<LOG_S, false>(s, b, col, cold_buf, salt, dang_h, R, sig);
}
}  // namespace

u64 bins_cold_bytes(const rbsc_bins* b) {
    return (b->n_cold_pad + SLOT_RUN) * (b->hashed ? sizeof(u64) : sizeof(u32));
}

void bins_scatter(rbsc_ctx* ctx, const rbsc_dgraph* g, const void* col, void* cold_buf) {
    const rbsc_bins* b = g->bins;
    if (b->n_warm) {
        // per XCD: 32 CUs, 4 workgroups of 512 threads each
        const u32 per_xcd = 128;
        const unsigned grid = per_xcd * (unsigned)BIN_GROUPS;
        if (b->hashed)
            warm_gather_kernel<u64><<<grid, WARM_THREADS, 0, ctx->stream>>>(
                b->warm_items, b->warm_first, b->slot_key, static_cast<const u64*>(col),
                static_cast<u64*>(cold_buf), per_xcd);
        else
            warm_gather_kernel<u32><<<grid, WARM_THREADS, 0, ctx->stream>>>(
                b->warm_items, b->warm_first, b->slot_key, static_cast<const u32*>(col),
                static_cast<u32*>(cold_buf), per_xcd);
    }
    if (!b->n_cold) return;
    if (b->head_order) {
        if (b->hashed)
            cold_stage_kernel<u64><<<(unsigned)b->n_tiles, STAGE_THREADS, 0, ctx->stream>>>(
                b->csc_head, b->csc_perm, b->csc_slot, b->tile_line, static_cast<const u64*>(col),
                static_cast<u64*>(cold_buf));
        else
            cold_stage_kernel<u32><<<(unsigned)b->n_tiles, STAGE_THREADS, 0, ctx->stream>>>(
                b->csc_head, b->csc_perm, b->csc_slot, b->tile_line, static_cast<const u32*>(col),
                static_cast<u32*>(cold_buf));
        return;
    }
    if (b->hashed)
        cold_scatter_kernel<u64><<<(unsigned)b->n_tiles, BLK, 0, ctx->stream>>>(
            b->csc_head, b->csc_slot, b->tile_line, static_cast<const u64*>(col),
            static_cast<u64*>(cold_buf));
    else
        cold_scatter_kernel<u32><<<(unsigned)b->n_tiles, BLK, 0, ctx->stream>>>(
            b->csc_head, b->csc_slot, b->tile_line, static_cast<const u32*>(col),
            static_cast<u32*>(cold_buf));
}

void bins_sig(rbsc_ctx* ctx, const rbsc_dgraph* g, const void* col, const void* cold_buf,
              u64 salt, u64 dang_h, u64* sig) {
    const rbsc_bins* b = g->bins;
    hipStream_t s = ctx->stream;
    if (b->n_zero)
        bin_zero_kernel<<<grid_for(b->n_zero << b->log_s), BLK, 0, s>>>(b->zero_bins, b->n_zero,
                                                                     b->log_s, g->R, sig);
    switch (b->log_s) {
        case 8: launch_bin_sig<8>(s, b, col, cold_buf, salt, dang_h, g->R, sig); break;
        case 9: launch_bin_sig<9>(s, b, col, cold_buf, salt, dang_h, g->R, sig); break;
        case 10: launch_bin_sig<10>(s, b, col, cold_buf, salt, dang_h, g->R, sig); break;
        case 11: launch_bin_sig<11>(s, b, col, cold_buf, salt, dang_h, g->R, sig); break;
        case 12: launch_bin_sig<12>(s, b, col, cold_buf, salt, dang_h, g->R, sig); break;
        case 13: launch_bin_sig<13>(s, b, col, cold_buf, salt, dang_h, g->R, sig); break;
        default: launch_bin_sig<14>(s, b, col, cold_buf, salt, dang_h, g->R, sig); break;
    }
}

void bins_release(rbsc_bins* b, hipStream_t s) {
    if (!b) return;
    for (void* p : {(void*)b->bin_slot, (void*)b->bin_hot, (void*)b->slot_key, (void*)b->slot_row,
                    (void*)b->slot_wt, (void*)b->dsum, (void*)b->csc_head, (void*)b->csc_slot,
                    (void*)b->tile_line, (void*)b->chunks, (void*)b->zero_bins,
                    (void*)b->cold_pos, (void*)b->csc_perm, (void*)b->warm_items})
        ctx_free(p, s);
    delete b;
}

int dgraph_bin(rbsc_ctx* ctx, rbsc_dgraph* g, u32 H, u32 D, u32 log_s, u64 chunk_slots,
               u64 scatter_chunk, bool head_order, bool hashed) {
    if (!g->dst || g->n_parts != 1) {
        rbsc_set_error("bin: needs the plain layout with lines resident on the device");
        return RBSC_ERR_INVALID;
    }
    if (log_s < 8 || log_s > 14 || scatter_chunk == 0) {
        rbsc_set_error("bin: rows per bin must be a power of two from 256 to 16384, scatter "
                       "chunks nonempty");
        return RBSC_ERR_INVALID;
    }
    hipStream_t s = ctx->stream;
    const u64 R = g->R;
    // Heads at or beyond GL are never gathered: nodes without lines, when no
    // initial colours are given (their colour is 0 every round).
    const u32 GL = g->init ? g->N : g->U;
    H = std::min(H, GL);
    D = std::min(D, H);
    // Without initial colours, lines into repeated rows gather their stored
    // rows (refinement then keeps no colours for repeated rows).
    const u32* rep = (!g->init && g->U > g->R) ? g->dup_rep : nullptr;
    const u64 n_bins = (R + (1ull << log_s) - 1) >> log_s;
    if (chunk_slots == 0) chunk_slots = 1ull << 19;
    // the staged scatter holds a tile's colours in LDS
    if (head_order)
        scatter_chunk = std::min<u64>(scatter_chunk, STAGE_BYTES / (hashed ? 8 : 4));
    rbsc_bins* b = new rbsc_bins();
    b->log_s = log_s;
    b->H = H;
    b->D = D;
    b->GL = GL;
    b->n_bins = n_bins;
    b->remap = rep != nullptr;
    b->scatter_chunk = scatter_chunk;
    b->head_order = head_order;
    b->hashed = hashed;
    u64 *n_hot = nullptr, *n_cold = nullptr, *tmp = nullptr, *tot = nullptr, *hot_start = nullptr,
        *cold_start = nullptr, *bin_n = nullptr, *keys[2] = {nullptr, nullptr},
        *vals[2] = {nullptr, nullptr}, *hot_all = nullptr;
    u32* grp_of_bin = nullptr;
    u32 *warm_cnt = nullptr, *warm_bound = nullptr;
    u64* warm_off = nullptr;
    WarmMap xcd_segs{};
    u64* grp_first = nullptr;
    // line tiles of the stored rows (every stored row has a line)
    const u64 n_lt = lt_tiles(g->E);
    u32* lt_first = nullptr;
    u64 *lt_hot = nullptr, *lt_cold = nullptr;
    std::vector<u64> h_bin_n, h_bin_hot, h_bin_slot;
    std::vector<u32> h_grp(n_bins + 1, 0), h_zero;
    std::vector<BinChunk> h_chunks;
    u64 h_tot[2] = {0, 0};
    int rc = RBSC_OK;
    const int hb = bitlen(H ? H - 1 : 0), gb = bitlen(GL ? GL - 1 : 0), bb = bitlen(n_bins ? n_bins - 1 : 0);
#define B_TRY(expr)                                                                  \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            rbsc_set_error("%s failed: %s", #expr, hipGetErrorString(_e));           \
            rc = _e == hipErrorOutOfMemory ? RBSC_ERR_NOMEM : RBSC_ERR_HIP;          \
            goto done;                                                               \
        }                                                                            \
    } while (0)
    if (hb + bb > 64 || gb + bb > 64 || gb + 3 > 64) {
        rbsc_set_error("bin: too many bins or heads for 64-bit sort keys");
        rc = RBSC_ERR_INVALID;
        goto done;
    }
    B_TRY(ctx_alloc(&n_hot, sizeof(u64) * (R + 1), s));
    B_TRY(ctx_alloc(&n_cold, sizeof(u64) * (R + 1), s));
    B_TRY(ctx_alloc(&b->dsum, sizeof(u64) * std::max<u64>(R, 1), s));
    B_TRY(ctx_alloc(&tmp, sizeof(u64) *
                              scan_tmp_words(std::max<u64>(std::max<u64>(R, n_bins), n_lt) + 1), s));
    B_TRY(ctx_alloc(&tot, sizeof(u64) * 2, s));
    B_TRY(ctx_alloc(&lt_first, sizeof(u32) * (n_lt + 1), s));
    B_TRY(ctx_alloc(&lt_hot, sizeof(u64) * (n_lt + 1), s));
    B_TRY(ctx_alloc(&lt_cold, sizeof(u64) * (n_lt + 1), s));
    B_TRY(hipMemsetAsync(n_hot, 0, sizeof(u64) * (R + 1), s));
    B_TRY(hipMemsetAsync(n_cold, 0, sizeof(u64) * (R + 1), s));
    B_TRY(hipMemsetAsync(b->dsum, 0, sizeof(u64) * std::max<u64>(R, 1), s));
    B_TRY(hipMemsetAsync(lt_hot + n_lt, 0, sizeof(u64), s));
    B_TRY(hipMemsetAsync(lt_cold + n_lt, 0, sizeof(u64), s));
    lt_first_row_kernel<<<grid_for(n_lt + 1), BLK, 0, s>>>(g->row_ptr, R, g->E, n_lt, lt_first);
    {
        auto* nh = reinterpret_cast<unsigned long long*>(n_hot);
        auto* nc = reinterpret_cast<unsigned long long*>(n_cold);
        auto* ds = reinterpret_cast<unsigned long long*>(b->dsum);
        if (g->wt)
            bin_count_tile_kernel<true><<<(unsigned)n_lt, LT_THREADS, 0, s>>>(
                g->row_ptr, g->dst, g->wt, R, g->E, H, GL, rep, g->U, lt_first, nh, nc, ds, lt_hot,
                lt_cold);
        else
            bin_count_tile_kernel<false><<<(unsigned)n_lt, LT_THREADS, 0, s>>>(
                g->row_ptr, g->dst, nullptr, R, g->E, H, GL, rep, g->U, lt_first, nh, nc, ds,
                lt_hot, lt_cold);
    }
    if ((rc = scan_excl_u64(n_hot, n_hot, R + 1, tot, tmp, s))) goto done;
    if ((rc = scan_excl_u64(n_cold, n_cold, R + 1, tot + 1, tmp, s))) goto done;
    if ((rc = scan_excl_u64(lt_hot, lt_hot, n_lt + 1, nullptr, tmp, s))) goto done;
    if ((rc = scan_excl_u64(lt_cold, lt_cold, n_lt + 1, nullptr, tmp, s))) goto done;
    B_TRY(hipMemcpyAsync(h_tot, tot, sizeof(u64) * 2, hipMemcpyDeviceToHost, s));
    B_TRY(hipStreamSynchronize(s));
    b->n_hot = h_tot[0];
    b->n_cold = h_tot[1];
    b->n_slots = b->n_hot + b->n_cold;
    if (b->n_slots >= (1ull << 32)) {
        rbsc_set_error("bin: %llu gathered lines exceed 2^32 slots", (unsigned long long)b->n_slots);
        rc = RBSC_ERR_INVALID;
        goto done;
    }
    // per-bin sizes and slot offsets
    B_TRY(ctx_alloc(&hot_start, sizeof(u64) * (n_bins + 1), s));
    B_TRY(ctx_alloc(&cold_start, sizeof(u64) * (n_bins + 1), s));
    B_TRY(ctx_alloc(&bin_n, sizeof(u64) * (n_bins + 1), s));
    B_TRY(ctx_alloc(&hot_all, sizeof(u64) * (n_bins + 1), s));
    B_TRY(ctx_alloc(&b->bin_hot, sizeof(u64) * (n_bins + 1), s));
    B_TRY(ctx_alloc(&b->bin_slot, sizeof(u64) * (n_bins + 1), s));
    bin_sizes_kernel<<<grid_for(n_bins + 1), BLK, 0, s>>>(n_hot, n_cold, R, log_s, n_bins, hot_start,
                                                          cold_start, bin_n, hot_all);
    h_bin_n.resize(n_bins + 1);
    h_bin_hot.resize(n_bins + 1);
    B_TRY(hipMemcpyAsync(h_bin_n.data(), bin_n, sizeof(u64) * (n_bins + 1), hipMemcpyDeviceToHost, s));
    B_TRY(hipMemcpyAsync(h_bin_hot.data(), hot_all, sizeof(u64) * n_bins, hipMemcpyDeviceToHost, s));
    if ((rc = scan_excl_u64(bin_n, b->bin_slot, n_bins + 1, nullptr, tmp, s))) goto done;
    B_TRY(hipStreamSynchronize(s));
    h_bin_slot.assign(n_bins + 1, 0);
    for (u64 k = 0; k < n_bins; ++k) h_bin_slot[k + 1] = h_bin_slot[k] + h_bin_n[k];
    // Scatter groups: contiguous bin ranges with about equal cold lines.
    {
        u64 seen = 0;
        for (u64 k = 0; k < n_bins; ++k) {
            const u64 cold = h_bin_n[k] - h_bin_hot[k];
            u32 grp = (u32)std::min<u64>(b->n_cold ? seen * BIN_GROUPS / b->n_cold : 0,
                                         BIN_GROUPS - 1);
            h_grp[k] = grp;
            seen += cold;
        }
    }
    // Chunks: a bin in one chunk when its slots fit, else split (those bins'
    // rows are zeroed each round and their chunks add atomically).
    for (u64 k = 0; k < n_bins; ++k) {
        const u64 n = h_bin_n[k], lo = h_bin_slot[k];
        if (n <= chunk_slots) {
            h_chunks.push_back({lo, lo + n, (u32)k, CHUNK_FIRST | CHUNK_ONLY});
        } else {
            const u64 parts = (n + chunk_slots - 1) / chunk_slots;
            const u64 step = (n + parts - 1) / parts;
            for (u64 j = 0; j < parts; ++j)
                h_chunks.push_back({lo + j * step, std::min(lo + n, lo + (j + 1) * step), (u32)k,
                                    j == 0 ? CHUNK_FIRST : 0u});
            h_zero.push_back((u32)k);
        }
    }
    // Larger chunks first: the long ones start early.
    std::stable_sort(h_chunks.begin(), h_chunks.end(), [](const BinChunk& a, const BinChunk& c) {
        return a.hi - a.lo > c.hi - c.lo;
    });
    b->n_chunks = h_chunks.size();
    b->n_zero = h_zero.size();
    B_TRY(ctx_alloc(&b->chunks, sizeof(BinChunk) * std::max<u64>(b->n_chunks, 1), s));
    B_TRY(hipMemcpyAsync(b->chunks, h_chunks.data(), sizeof(BinChunk) * b->n_chunks,
                         hipMemcpyHostToDevice, s));
    if (b->n_zero) {
        B_TRY(ctx_alloc(&b->zero_bins, sizeof(u32) * b->n_zero, s));
        B_TRY(hipMemcpyAsync(b->zero_bins, h_zero.data(), sizeof(u32) * b->n_zero,
                             hipMemcpyHostToDevice, s));
    }
    // slot arrays
    // + SLOT_RUN zeroed entries: the loads of the last run stay in bounds
    B_TRY(ctx_alloc(&b->slot_key, sizeof(u32) * (b->n_slots + SLOT_RUN), s));
    B_TRY(ctx_alloc(&b->slot_row, sizeof(unsigned short) * (b->n_slots + SLOT_RUN), s));
    B_TRY(hipMemsetAsync(b->slot_key, 0, sizeof(u32) * (b->n_slots + SLOT_RUN), s));
    B_TRY(hipMemsetAsync(b->slot_row, 0, sizeof(unsigned short) * (b->n_slots + SLOT_RUN), s));
    if (g->wt) {
        B_TRY(ctx_alloc(&b->slot_wt, sizeof(u64) * (b->n_slots + SLOT_RUN), s));
        B_TRY(hipMemsetAsync(b->slot_wt + b->n_slots, 0, sizeof(u64) * SLOT_RUN, s));
    }
    if ((rc = sort_ws_init(&ctx->sort, std::max(b->n_hot, b->n_cold), ctx->d_err, s))) goto done;
    // hot slots: records sorted by (bin, head)
    if (b->n_hot) {
        for (int i = 0; i < 2; ++i) {
            B_TRY(ctx_alloc(&keys[i], sizeof(u64) * b->n_hot, s));
            B_TRY(ctx_alloc(&vals[i], sizeof(u64) * b->n_hot, s));
        }
        bin_fill_tile_kernel<true><<<(unsigned)n_lt, LT_THREADS, 0, s>>>(
            g->row_ptr, g->dst, R, g->E, H, GL, rep, g->U, log_s, hb, lt_first, lt_hot, keys[0],
            vals[0]);
        int w = 0;
        const int bits = (hb + bb + 7) / 8 * 8;
        if ((rc = radix_sort_u64(&ctx->sort, keys, vals, b->n_hot, 0, bits, &w, s))) goto done;
        if (g->wt)
            place_hot_kernel<true><<<grid_for(b->n_hot), BLK, 0, s>>>(
                keys[w], vals[w], b->n_hot, hb, b->bin_slot, hot_start, g->wt, b->slot_key,
                b->slot_row, b->slot_wt);
        else
            place_hot_kernel<false><<<grid_for(b->n_hot), BLK, 0, s>>>(
                keys[w], vals[w], b->n_hot, hb, b->bin_slot, hot_start, nullptr, b->slot_key,
                b->slot_row, nullptr);
        if (D < H) {
            // Warm segments: boundaries at about equal shares of the warm lines.
            const u64 nw = H - D;
            B_TRY(ctx_alloc(&warm_cnt, sizeof(u32) * nw, s));
            B_TRY(hipMemsetAsync(warm_cnt, 0, sizeof(u32) * nw, s));
            warm_hist_kernel<<<grid_for(b->n_hot), BLK, 0, s>>>(keys[w], b->n_hot, hb, D, H,
                                                               warm_cnt);
            std::vector<u32> h_cnt(nw);
            B_TRY(hipMemcpyAsync(h_cnt.data(), warm_cnt, sizeof(u32) * nw, hipMemcpyDeviceToHost, s));
            B_TRY(hipStreamSynchronize(s));
            u64 total = 0;
            for (u32 c : h_cnt) total += c;
            u32 h_bound[WARM_SEGS + 1];
            h_bound[0] = D;
            h_bound[WARM_SEGS] = H;
            u64 cum = 0, h = 0;
            for (int j = 1; j < WARM_SEGS; ++j) {
                const u64 want = total * (u64)j / WARM_SEGS;
                while (h < nw && cum + h_cnt[h] <= want) cum += h_cnt[h++];
                h_bound[j] = (u32)(D + h);
            }
            // Segments by width (heads), widest first, dealt to the XCDs in
            // snake order: each XCD gets WARM_PER_XCD segments of equal line
            // shares and about equal colour footprints.
            int order[WARM_SEGS];
            for (int j = 0; j < WARM_SEGS; ++j) order[j] = j;
            std::stable_sort(order, order + WARM_SEGS, [&](int a, int c) {
                return h_bound[a + 1] - h_bound[a] > h_bound[c + 1] - h_bound[c];
            });
            for (int r = 0; r < WARM_SEGS; ++r) {
                const int round = r / (int)BIN_GROUPS, pos = r % (int)BIN_GROUPS;
                const int x = (round & 1) ? (int)BIN_GROUPS - 1 - pos : pos;
                xcd_segs.seg[x][round] = (unsigned char)order[r];
            }
            B_TRY(ctx_alloc(&warm_bound, sizeof(u32) * (WARM_SEGS + 1), s));
            B_TRY(hipMemcpyAsync(warm_bound, h_bound, sizeof(u32) * (WARM_SEGS + 1),
                                 hipMemcpyHostToDevice, s));
            B_TRY(ctx_alloc(&warm_off, sizeof(u64) * n_bins * (WARM_SEGS + 1), s));
            warm_off_kernel<<<grid_for(n_bins * (WARM_SEGS + 1)), BLK, 0, s>>>(
                b->slot_key, b->bin_slot, hot_all, n_bins, warm_bound, warm_off, b->bin_hot);
        } else {
            B_TRY(hipMemcpyAsync(b->bin_hot, hot_all, sizeof(u64) * n_bins,
                                 hipMemcpyDeviceToDevice, s));
        }
        B_TRY(hipStreamSynchronize(s));
        for (int i = 0; i < 2; ++i) {
            ctx_free(keys[i], s);
            ctx_free(vals[i], s);
            keys[i] = vals[i] = nullptr;
        }
    }
    if (!b->n_hot)
        B_TRY(hipMemsetAsync(b->bin_hot, 0, sizeof(u64) * n_bins, s));
    {
        // Each bin's buffered (warm and cold) slots [c0, c1) as whole aligned
        // groups of 8 in a refinement's cold-colour buffer.
        std::vector<u64> h_dense(n_bins, 0), h_pos(n_bins, 0);
        B_TRY(hipMemcpyAsync(h_dense.data(), b->bin_hot, sizeof(u64) * n_bins,
                             hipMemcpyDeviceToHost, s));
        B_TRY(hipStreamSynchronize(s));
        u64 pos = 0, dense = 0;
        for (u64 k = 0; k < n_bins; ++k) {
            h_pos[k] = pos;
            dense += h_dense[k];
            const u64 c0 = h_bin_slot[k] + h_dense[k], c1 = h_bin_slot[k + 1];
            if (c1 > c0) pos += ((c1 + 7) & ~7ull) - (c0 & ~7ull);
        }
        b->n_warm = b->n_hot - dense;
        b->n_cold_pad = pos;
        if (b->n_warm) {
            // work items: per XCD, its segments' run in every bin, cut into
            // pieces of at most WARM_ITEM slots
            std::vector<u64> h_off(n_bins * (WARM_SEGS + 1));
            B_TRY(hipMemcpyAsync(h_off.data(), warm_off, sizeof(u64) * h_off.size(),
                                 hipMemcpyDeviceToHost, s));
            B_TRY(hipStreamSynchronize(s));
            std::vector<WarmItem> items;
            for (u64 x = 0; x < BIN_GROUPS; ++x) {
                b->warm_first.f[x] = (u32)items.size();
                for (u64 k = 0; k < n_bins; ++k) {
                    const u64 c0 = (h_bin_slot[k] + h_dense[k]) & ~7ull;
                    for (int r = 0; r < WARM_PER_XCD; ++r) {
                        const int sg = xcd_segs.seg[x][r];
                        const u64 lo = h_off[k * (WARM_SEGS + 1) + sg];
                        const u64 hi = h_off[k * (WARM_SEGS + 1) + sg + 1];
                        for (u64 a = lo; a < hi; a += WARM_ITEM) {
                            const u32 n = (u32)std::min<u64>(WARM_ITEM, hi - a);
                            items.push_back({a, h_pos[k] + (a - c0), n, 0});
                        }
                    }
                }
            }
            b->warm_first.f[BIN_GROUPS] = (u32)items.size();
            B_TRY(ctx_alloc(&b->warm_items, sizeof(WarmItem) * items.size(), s));
            B_TRY(hipMemcpyAsync(b->warm_items, items.data(), sizeof(WarmItem) * items.size(),
                                 hipMemcpyHostToDevice, s));
            b->bytes_warm = sizeof(WarmItem) * items.size();
        }
        if (pos + SLOT_RUN >= (1ull << 32)) {
            rbsc_set_error("bin: %llu buffered slots exceed 2^32", (unsigned long long)pos);
            rc = RBSC_ERR_INVALID;
            goto done;
        }
        B_TRY(ctx_alloc(&b->cold_pos, sizeof(u64) * std::max<u64>(n_bins, 1), s));
        B_TRY(hipMemcpyAsync(b->cold_pos, h_pos.data(), sizeof(u64) * n_bins,
                             hipMemcpyHostToDevice, s));
    }
    // cold slots and the scatter records
    if (b->n_cold) {
        for (int i = 0; i < 2; ++i) {
            B_TRY(ctx_alloc(&keys[i], sizeof(u64) * b->n_cold, s));
            B_TRY(ctx_alloc(&vals[i], sizeof(u64) * b->n_cold, s));
        }
        B_TRY(ctx_alloc(&grp_of_bin, sizeof(u32) * (n_bins + 1), s));
        B_TRY(hipMemcpyAsync(grp_of_bin, h_grp.data(), sizeof(u32) * (n_bins + 1),
                             hipMemcpyHostToDevice, s));
        bin_fill_tile_kernel<false><<<(unsigned)n_lt, LT_THREADS, 0, s>>>(
            g->row_ptr, g->dst, R, g->E, H, GL, rep, g->U, log_s, gb, lt_first, lt_cold, keys[0],
            vals[0]);
        int w = 0;
        int bits = (gb + bb + 7) / 8 * 8;
        if ((rc = radix_sort_u64(&ctx->sort, keys, vals, b->n_cold, 0, bits, &w, s))) goto done;
        if (g->wt)
            place_cold_kernel<true><<<grid_for(b->n_cold), BLK, 0, s>>>(
                keys[w], vals[w], b->n_cold, gb, b->bin_slot, hot_all, b->bin_hot, cold_start, grp_of_bin,
                b->cold_pos, g->wt, b->slot_key, b->slot_row, b->slot_wt, keys[w ^ 1],
                vals[w ^ 1]);
        else
            place_cold_kernel<false><<<grid_for(b->n_cold), BLK, 0, s>>>(
                keys[w], vals[w], b->n_cold, gb, b->bin_slot, hot_all, b->bin_hot, cold_start, grp_of_bin,
                b->cold_pos, nullptr, b->slot_key, b->slot_row, nullptr, keys[w ^ 1],
                vals[w ^ 1]);
        u64* k2[2] = {keys[w ^ 1], keys[w]};
        u64* v2[2] = {vals[w ^ 1], vals[w]};
        int w2 = 0;
        bits = (gb + 3 + 7) / 8 * 8;
        if ((rc = radix_sort_u64(&ctx->sort, k2, v2, b->n_cold, 0, bits, &w2, s))) goto done;
        // tiles: chunks of each group's lines in head order, interleaved by group
        u64 h_first[BIN_GROUPS + 1] = {0};
        u64 most = 0;
        for (u64 k = 0; k < n_bins; ++k) h_first[h_grp[k] + 1] += h_bin_n[k] - h_bin_hot[k];
        for (u64 k = 0; k < BIN_GROUPS; ++k) {
            most = std::max(most, h_first[k + 1]);
            h_first[k + 1] += h_first[k];
        }
        b->n_tiles = (most + scatter_chunk - 1) / scatter_chunk * BIN_GROUPS;
        B_TRY(ctx_alloc(&grp_first, sizeof(u64) * (BIN_GROUPS + 1), s));
        B_TRY(hipMemcpyAsync(grp_first, h_first, sizeof(u64) * (BIN_GROUPS + 1),
                             hipMemcpyHostToDevice, s));
        u64* k3[2] = {k2[w2 ^ 1], k2[w2]};
        u64* v3[2] = {v2[w2], v2[w2 ^ 1]};
        chunk_key_kernel<<<grid_for(b->n_cold), BLK, 0, s>>>(k2[w2], v2[w2], b->n_cold, gb,
                                                            grp_first, scatter_chunk,
                                                            head_order ? 0 : 1, k3[0]);
        int w3 = 0;
        bits = (32 + bitlen(b->n_tiles) + 7) / 8 * 8;
        if ((rc = radix_sort_u64(&ctx->sort, k3, v3, b->n_cold, 0, bits, &w3, s))) goto done;
        B_TRY(ctx_alloc(&b->csc_head, sizeof(u32) * (b->n_cold + SLOT_RUN), s));
        B_TRY(ctx_alloc(&b->csc_slot, sizeof(u32) * (b->n_cold + SLOT_RUN), s));
        B_TRY(hipMemsetAsync(b->csc_head, 0, sizeof(u32) * (b->n_cold + SLOT_RUN), s));
        B_TRY(hipMemsetAsync(b->csc_slot, 0, sizeof(u32) * (b->n_cold + SLOT_RUN), s));
        B_TRY(ctx_alloc(&b->tile_line, sizeof(u64) * (b->n_tiles + 1), s));
        csc_fill_kernel<<<grid_for(b->n_cold), BLK, 0, s>>>(k3[w3], v3[w3], b->n_cold, b->n_tiles,
                                                           b->csc_head, b->csc_slot, b->tile_line);
        if (head_order) {
            u64* k4[2] = {k3[w3 ^ 1], k3[w3]};
            u64* v4[2] = {v3[w3 ^ 1], v3[w3]};
            stage_key_kernel<<<grid_for(b->n_cold), BLK, 0, s>>>(k3[w3], v3[w3], b->n_cold, k4[0],
                                                                v4[0]);
            int w4 = 0;
            if ((rc = radix_sort_u64(&ctx->sort, k4, v4, b->n_cold, 0, bits, &w4, s))) goto done;
            B_TRY(ctx_alloc(&b->csc_perm, sizeof(unsigned short) * b->n_cold, s));
            stage_fill_kernel<<<grid_for(b->n_cold), BLK, 0, s>>>(
                k4[w4], v4[w4], b->n_cold, b->tile_line, b->csc_slot, b->csc_perm);
        }
    }
    B_TRY(hipGetLastError());
    B_TRY(hipStreamSynchronize(s));
    b->bytes = 16 * (n_bins + 1) + 6 * b->n_slots + (g->wt ? 8 * b->n_slots : 0) + 8 * R +
               (head_order ? 10 : 8) * b->n_cold + 8 * (b->n_tiles + 1) +
               b->bytes_warm + sizeof(BinChunk) * b->n_chunks +
               4 * b->n_zero + 8 * n_bins;
    if (g->bins) {
        g->bytes -= g->bins->bytes;
        bins_release(g->bins, s);
    }
    g->bins = b;
    g->bytes += b->bytes;
    b = nullptr;
done:
    for (u64* p : {n_hot, n_cold, tmp, tot, hot_start, cold_start, bin_n, keys[0], keys[1],
                   vals[0], vals[1], grp_first, hot_all, warm_off, lt_hot, lt_cold})
        ctx_free(p, s);
    for (u32* p : {grp_of_bin, warm_cnt, warm_bound, lt_first})
        ctx_free(p, s);
    if (b) bins_release(b, s);
    return rc;
#undef B_TRY
}
