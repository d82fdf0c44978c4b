// refine.hip -- the RBSC refinement round on gfx950 and the C ABI around it.
//
// One round (reference rbsc.cu:354-425), colour c_r, salt r:
//   sig_r(u) = sum over u's lines of randomize(c_r(v) + r) * w      (set_values
//              80-87 + ReduceByKey 366-377 + mend_partition 104-116)
//   c_{r+1}(u) = dense rank of (sig_r(u), c_r(u)), ascending        (384-396)
//   stop when the number of distinct pairs repeats                  (398-425)
//
// gfx950 layout of one round:
//   sig_kernel    edge-balanced wave tiles over the compressed CSR; each lane
//                 folds 8 lines, rows are closed with a 64-lane segmented
//                 scan; lines read coalesced, hashed colours gathered.
//   sig_fixup     adds the partial sums of rows that span wave tiles.
//   gather_keys   keys[i] = sig[node of position i], positions kept sorted
//                 by the current colour from the previous round.
//   radix sort    stable onesweep sort by the 64-bit signature; stability
//                 turns colour order into (sig, colour) order, so the
//                 reference's first sort by colour (rbsc.cu:386) is free.
//   rank_kernel   head flags + decoupled look-back scan = dense rank; writes
//                 the packed (rank, node) order for the next round and the
//                 next round's hashed colour randomize(rank + r + 1) of every
//                 node, so the signature kernel never re-hashes per line.
// The reference's collision loop (rbsc.cu:404-419) always exits after one
// pass because randomize is a bijection; it is not repeated here.
#include "internal.h"
#include "scan.h"
#include "batch.h"
#include "tiles.h"
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>

// ---------------------------------------------------------------- errors
static thread_local char g_err[512] = "";
extern "C" void rbsc_set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
extern "C" const char* rbsc_last_error(void) { return g_err; }
extern "C" const char* rbsc_version(void) { return "rbsc-gfx950 0.1"; }

// Host buffers from rbsc_host_alloc: coherent page-locked memory that
// kernels may write directly (base address -> bytes).
namespace {
std::mutex g_host_mu;
std::map<uintptr_t, u64> g_host_allocs;
}  // namespace

extern "C" int rbsc_host_alloc(uint64_t bytes, void** out) {
    if (!out) {
        rbsc_set_error("rbsc_host_alloc: null argument");
        return RBSC_ERR_INVALID;
    }
    *out = nullptr;
    const u64 n = bytes ? bytes : 1;
    const hipError_t e = hipHostMalloc(out, n, hipHostMallocCoherent);
    if (e != hipSuccess) {
        *out = nullptr;
        rbsc_set_error("hipHostMalloc of %llu bytes failed: %s", (unsigned long long)bytes,
                       hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? RBSC_ERR_NOMEM : RBSC_ERR_HIP;
    }
    std::lock_guard<std::mutex> lk(g_host_mu);
    g_host_allocs[(uintptr_t)*out] = n;
    return RBSC_OK;
}

extern "C" int rbsc_host_free(void* p) {
    if (!p) return RBSC_OK;
    {
        std::lock_guard<std::mutex> lk(g_host_mu);
        g_host_allocs.erase((uintptr_t)p);
    }
    HIP_TRY(hipHostFree(p));
    return RBSC_OK;
}

void* host_device_ptr(const void* p, u64 bytes) {
    if (!p) return nullptr;
    uintptr_t base = 0;
    {
        std::lock_guard<std::mutex> lk(g_host_mu);
        auto it = g_host_allocs.upper_bound((uintptr_t)p);
        if (it == g_host_allocs.begin()) return nullptr;
        --it;
        if ((uintptr_t)p + bytes > it->first + it->second) return nullptr;
        base = it->first;
    }
    void* d = nullptr;
    if (hipHostGetDevicePointer(&d, (void*)base, 0) != hipSuccess || !d) return nullptr;
    return static_cast<char*>(d) + ((uintptr_t)p - base);
}

namespace {

constexpr int BLK = 256;
inline unsigned grid_for(u64 n, int blk = BLK) { return (unsigned)((n + blk - 1) / blk); }

// ---------------------------------------------------------------- kernels
// Signatures of the first round from a uniform start (rbsc.cu:51-63): every
// node with a line has colour N+1 (hot_h = randomize(N + 1 + 0)) and every
// other node colour 0 (dang_h = randomize(0 + 0)), so a stored row's
// signature, the sum over its lines of randomize(c(v)) * w
// (rbsc.cu:80-87, 366-377), is hot_h times the weight of its lines into
// nodes with lines plus dang_h times the weight of the others, in Z/2^64.
// row_w: the rows' total weights, or null without weights (the line count).
__global__ void first_round_sig_kernel(const u64* __restrict__ row_ptr,
                                       const u64* __restrict__ row_w,
                                       const u64* __restrict__ dang_w, u64 R, u64 hot_h,
                                       u64 dang_h, u64* __restrict__ sig) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    const u64 d = dang_w[i];
    const u64 w = row_w ? row_w[i] : row_ptr[i + 1] - row_ptr[i];
    sig[i] = hot_h * (w - d) + dang_h * d;
}

__global__ void fill_u64(u64* p, u64 n, u64 v) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Signature of every row over the internal layout: sig[j] = sum over row
// j's lines of hc[dst] * w (hc[v] = randomize(c(v) + r)).  Heads at or
// beyond gather_limit are nodes without lines whose hashed colour is the
// per-round constant dang_h, so they are never gathered.
// One wave per tile of 512 lines, 8 lines per lane.  Row ends come from the
// row-end bitmap (one byte per lane); the lane's first row is the tile's
// first row plus the row ends before it (wave prefix sum).  Rows spanning
// lanes close with a segmented scan; rows spanning tiles get the tile's
// partial sum in carry_val, added by sig_fixup_kernel.
// With PART (destination-partitioned layout) the runs are segments: the
// partial sum of segment k is added to sig[seg_row[k]] (one segment per row
// per partition, partitions in separate launches, so no two lanes add to the
// same row concurrently).
// With C32 the per-node array holds raw u32 colours and each line hashes
// randomize(colour + salt) itself (half the gathered bytes, four 64-bit
// multiplies per line); otherwise it holds the hashed colours.
// XCD (PART == 2): one launch covers all partitions; workgroup b serves
// partition b mod 8 (workgroups are dealt round-robin to the 8 XCDs, so each
// XCD's L2 caches a disjoint eighth of the colours) and partial sums are
// added atomically because partitions now run concurrently.
template <bool HAS_WT, int PART, bool C32>
__global__ __launch_bounds__(256) void sig_kernel(
    const u32* __restrict__ dst, const u64* __restrict__ wt, u64 E,
    const unsigned char* __restrict__ end_bytes, const u64* __restrict__ hc, u32 gather_limit,
    u64 dang_h, u64 salt, u64* __restrict__ sig, const u32* __restrict__ seg_row,
    const u32* __restrict__ tile_row, u64* __restrict__ carry_val, u64 t_begin, u64 t_end,
    u64 line_base, const u64* __restrict__ part_tile) {
    const u32* __restrict__ col = reinterpret_cast<const u32*>(hc);
    // Tiles [t_begin, t_end); dst/wt hold lines [line_base, ...) (a streamed
    // slice, or the whole array when line_base == 0).
    u64 t;
    if (PART == 2) {
        const u32 p = blockIdx.x & 7;
        t = part_tile[p] + (u64)(blockIdx.x >> 3) * (blockDim.x / 64) + (threadIdx.x >> 6);
        if (t >= part_tile[p + 1]) return;  // uniform per wave
    } else {
        t = t_begin + (u64)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
        if (t >= t_end) return;  // uniform per wave
    }
    const int lane = threadIdx.x & 63;
    const u64 e0 = t * SIG_TW;
    const u64 e1 = e0 + SIG_TW < E ? e0 + SIG_TW : E;
    const u64 es = e0 + (u64)lane * SIG_V;
    const bool active = es < e1;
    const u32 nv = active ? (u32)(e1 - es < SIG_V ? e1 - es : SIG_V) : 0;
    // Row-end flags of this lane's lines (bits beyond E are zero).
    const u32 flags = active ? (u32)end_bytes[t * 64 + lane] : 0u;
    // Row of the lane's first line.
    u32 cnt = (u32)__popc(flags);
    u32 x = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const u32 y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    u32 k = tile_row[t] + (x - cnt);
    // Does the lane start inside a row that began earlier?
    const u32 pflags = __shfl_up(flags, 1);
    u32 prev_end;
    if (lane > 0) prev_end = (pflags >> 7) & 1u;
    else prev_end = t == 0 ? 1u : ((u32)end_bytes[t * 64 - 1] >> 7) & 1u;
    bool cont = prev_end == 0;

    u64 hv[SIG_V];
    if (nv == SIG_V) {
        typedef u32 u32x4 __attribute__((ext_vector_type(4)));
        const u32x4* p = reinterpret_cast<const u32x4*>(dst + (es - line_base));
        const u32x4 a = __builtin_nontemporal_load(p);
        const u32x4 b = __builtin_nontemporal_load(p + 1);
        const u32 dv[SIG_V] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        if (C32) {
            u32 cv[SIG_V];
#pragma unroll
            for (int j = 0; j < SIG_V; ++j) cv[j] = dv[j] < gather_limit ? col[dv[j]] : 0u;
#pragma unroll
            for (int j = 0; j < SIG_V; ++j)
                hv[j] = dv[j] < gather_limit ? rbsc_randomize((u64)cv[j] + salt) : dang_h;
        } else {
#pragma unroll
            for (int j = 0; j < SIG_V; ++j) hv[j] = dv[j] < gather_limit ? hc[dv[j]] : dang_h;
        }
        if (HAS_WT) {
#pragma unroll
            for (int j = 0; j < SIG_V; ++j) hv[j] *= wt[es - line_base + j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < SIG_V; ++j) {
            hv[j] = 0;
            if ((u32)j < nv) {
                const u32 d = dst[es - line_base + j];
                if (C32) hv[j] = d < gather_limit ? rbsc_randomize((u64)col[d] + salt) : dang_h;
                else hv[j] = d < gather_limit ? hc[d] : dang_h;
                if (HAS_WT) hv[j] *= wt[es - line_base + j];
            }
        }
    }
    u64 acc = 0, head_val = 0;
    bool has_head = false;
    u32 head_k = 0;
#pragma unroll
    for (int j = 0; j < SIG_V; ++j) {
        if ((u32)j < nv) {
            acc += hv[j];
            if ((flags >> j) & 1u) {
                if (cont) {
                    has_head = true;
                    head_val = acc;
                    head_k = k;
                    cont = false;
                } else if (PART == 2) {
                    atomicAdd((unsigned long long*)&sig[seg_row[k]], (unsigned long long)acc);
                } else if (PART) {
                    sig[seg_row[k]] += acc;
                } else {
                    sig[k] = acc;
                }
                acc = 0;
                ++k;
            }
        }
    }
    const bool tail_open = nv > 0 && !((flags >> (nv - 1)) & 1u);
    u64 key = active ? (tail_open ? 2 * (u64)k : 2 * (u64)k - 1)
                     : (1ULL << 40) + 2 * (u64)lane + 1;
    // Segmented inclusive scan of open-row partials across the wave.
    u64 S = tail_open ? acc : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const u64 oS = __shfl_up(S, d);
        const u64 oK = __shfl_up(key, d);
        if (lane >= d && oK == key) S += oS;
    }
    const u64 pS = __shfl_up(S, 1);
    const u64 pK = __shfl_up(key, 1);
    if (has_head) {
        if (lane > 0 && pK == 2 * (u64)head_k) head_val += pS;
        if (PART == 2)
            atomicAdd((unsigned long long*)&sig[seg_row[head_k]], (unsigned long long)head_val);
        else if (PART) sig[seg_row[head_k]] += head_val;
        else sig[head_k] = head_val;
    }
    const int last_lane = (int)((e1 - e0 - 1) / SIG_V);
    if (lane == last_lane && tail_open) carry_val[t] = S;
}

__global__ void sig_fixup_kernel(const u32* __restrict__ carry_row,
                                 const u64* __restrict__ carry_val,
                                 const u32* __restrict__ seg_row, u64* sig, u64 n_tiles) {
    const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles) return;
    const u32 r = carry_row[t];
    if (r != NO_ROW)
        atomicAdd((unsigned long long*)&sig[seg_row ? seg_row[r] : r],
                  (unsigned long long)carry_val[t]);
}

// Rows that repeat a stored row take its signature.
__global__ void dup_sig_kernel(const u32* __restrict__ dup_rep, u64 R, u64 n_dup, u64* sig) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_dup) sig[R + i] = sig[dup_rep[i]];
}

// Colours of repeated rows from their stored rows (u32 colours, or u64 hashed
// colours when wide).
__global__ void dup_colors_kernel(const u32* __restrict__ dup_rep, u64 R, u64 n_dup, void* col,
                                  int wide) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_dup) return;
    if (wide) {
        u64* c = static_cast<u64*>(col);
        c[R + i] = c[dup_rep[i]];
    } else {
        u32* c = static_cast<u32*>(col);
        c[R + i] = c[dup_rep[i]];
    }
}

// Round-0 state when every node with a line has colour N+1 (rbsc.cu:51-63)
// and nodes without lines are left out of the node side: positions are
// already sorted by colour.
__global__ void init_uniform_kernel(u64* __restrict__ P, u64* __restrict__ hc, u64 U, u64 c,
                                    int c32) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < U) {
        P[u] = (c << 32) | u;
        if (c32) reinterpret_cast<u32*>(hc)[u] = (u32)c;
        else hc[u] = rbsc_randomize(c + 0);
    }
}
// Round-0 state from explicit initial colours (internal order); sorted by
// colour afterwards.
__global__ void init_given_kernel(const u32* __restrict__ init, u64* __restrict__ keys,
                                  u64* __restrict__ P, u64* __restrict__ hc, u64 N, int c32) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < N) {
        const u64 c = init[u];
        keys[u] = c;
        P[u] = (c << 32) | u;
        if (c32) reinterpret_cast<u32*>(hc)[u] = (u32)c;
        else hc[u] = rbsc_randomize(c + 0);
    }
}
__global__ void permute_fill_kernel(const u64* __restrict__ fill, const u32* __restrict__ pi,
                                    u64 N, u64* __restrict__ sig) {
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < N) sig[pi[v]] = fill[v];
}

// Sort keys in colour order: keys[i] = sig[node of P[i]].
__global__ void gather_keys_kernel(const u64* __restrict__ P, const u64* __restrict__ sig,
                                   u64* __restrict__ keys, u64 n) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = sig[(u32)P[i]];
}
// The same, accumulating the
// histograms of the npass (<= SORT_HD) 8-bit digits from bit_begin into the
// sort's counters (the sort then skips its own pass over the keys).
constexpr unsigned GK_BLOCKS = 2048;
__global__ __launch_bounds__(256) void gather_keys_hist_kernel(
    const u64* __restrict__ P, const u64* __restrict__ sig, u64* __restrict__ keys, u64 n,
    int bit_begin, int npass, u32* __restrict__ ghist) {
    __shared__ u32 lh[SORT_HWORDS];
    for (int i = threadIdx.x; i < SORT_HWORDS; i += 256) lh[i] = 0;
    __syncthreads();
    constexpr int IT = 8;  // keys per thread per step, all gathered before any is counted
    const u64 stride = (u64)gridDim.x * 256 * IT;
    for (u64 b = (u64)blockIdx.x * 256 * IT; b < n; b += stride) {
        u64 k[IT];
#pragma unroll
        for (int j = 0; j < IT; ++j) {
            const u64 i = b + (u64)j * 256 + threadIdx.x;
            k[j] = i < n ? sig[(u32)P[i]] : 0;
        }
#pragma unroll
        for (int j = 0; j < IT; ++j) {
            const u64 i = b + (u64)j * 256 + threadIdx.x;
            if (i >= n) continue;
            keys[i] = k[j];
            for (int p = 0; p < npass; ++p)
                sort_hist_add(lh, p, (u32)(k[j] >> (bit_begin + 8 * p)) & 255u);
        }
    }
    __syncthreads();
    sort_hist_flush(lh, npass, ghist);
}

// After a stable sort by the signature bits at and above `lo` (colour order
// kept inside equal keys), every run of equal top bits must be ordered by
// the low bits, stably.  Runs of equal signatures are already in order; only
// runs holding a "break" (neighbours with equal top bits and different low
// bits, from distinct signatures colliding in the top bits) need work.
// tie_find_kernel finds breaks in one coalesced pass and lists each such run
// once (bounds by galloping and bisection on the sorted top bits; the first
// break to stamp the run's start with this call's epoch records it), in one
// of three lists by length:
//   small (<= 64 members): tie_small_kernel, one wave per run, a bitonic sort
//     of (low bits, position in run) across the wave's lanes;
//   mid (<= TIE_MID members): tie_mid_kernel, one workgroup per run, a
//     bitonic sort of the same keys in LDS;
//   long: gathered into a compact array and sorted there with the onesweep
//     sort (by low bits, then by run), driven by the host.
// Sorting by (low bits, position) is stable.  A list that overflows raises
// *overflow and the caller redoes the round's sort on all 64 bits.  `lo` is
// 32 normally; 48 is a verification hook that makes ties frequent.
// Packed keys (sig not null): a key holds the signature's top bits and, below
// `lo`, the node; the low bits of the signature are read from sig[node], only
// where neighbours' top bits are equal.
constexpr int TIE_BLK = 256;
constexpr u32 TIE_LIST = 512;  // small runs a workgroup collects in LDS before spilling
constexpr unsigned TIE_GRID = 2048;
constexpr int TF_ITEMS = 16;
constexpr u64 TF_TILE = 256 * TF_ITEMS;
constexpr u32 TIE_SMALL = 64;
constexpr u32 TIE_MID = 4096;
constexpr u64 TIE_MID_CAP = 65536;   // mid runs listed per round
constexpr u64 TIE_LONG_CAP = 4096;   // long runs listed per round
// counts[0..2]: small, mid, long runs listed.
__device__ __forceinline__ u64 tie_low(u64 k, const u64* __restrict__ sig, u64 lowmask) {
    return (sig ? sig[k & lowmask] : k) & lowmask;
}

__global__ __launch_bounds__(TIE_BLK) void tie_find_kernel(
    const u64* __restrict__ keys, const u64* __restrict__ sig, u64 n, u32 lo,
    u32* __restrict__ mark, u32 epoch,
    u64* __restrict__ runs_s, u64 cap_s, u64* __restrict__ runs_m, u64* __restrict__ runs_l,
    u32* __restrict__ counts, u32* __restrict__ overflow) {
    // Small runs are gathered per workgroup in LDS and reserved in the global
    // list with one atomic per workgroup.
    __shared__ u64 ls[2 * TIE_LIST];
    __shared__ u32 lcount, lbase;
    if (threadIdx.x == 0) lcount = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const u64 lowmask = (1ULL << lo) - 1;
    for (u64 t = blockIdx.x; t * TF_TILE < n; t += gridDim.x) {  // uniform per workgroup
        const u64 wbase = t * TF_TILE + (u64)wave * (64 * TF_ITEMS);
        u64 k[TF_ITEMS];
#pragma unroll
        for (int j = 0; j < TF_ITEMS; ++j) {
            const u64 i = wbase + 64 * j + lane;
            k[j] = i < n ? keys[i] : 0;
        }
        u64 lk = (wbase > 0 && wbase < n) ? keys[wbase - 1] : 0;
#pragma unroll
        for (int j = 0; j < TF_ITEMS; ++j) {
            u64 pk = __shfl_up(k[j], 1);
            if (lane == 0) pk = lk;
            lk = __shfl(k[j], 63);
            const u64 i = wbase + 64 * j + lane;
            const u64 top = k[j] >> lo;
            if (!(i < n && i > 0 && top == (pk >> lo))) continue;
            if (sig ? tie_low(k[j], sig, lowmask) == tie_low(pk, sig, lowmask) : k[j] == pk)
                continue;  // not a break
            // Run bounds by galloping outward from the break, then bisection,
            // so a run of length L costs O(log L) probes.
            u64 b = i - 1, step = 1;  // keys[b] is in the run
            u64 a = b;
            while (a > 0) {
                const u64 m = a > step ? a - step : 0;
                if ((keys[m] >> lo) != top) { a = m; break; }
                b = m;
                a = m;
                step <<= 1;
            }
            if ((keys[a] >> lo) != top) {  // first index of the run lies in (a, b]
                ++a;
                while (a < b) {
                    const u64 m = a + (b - a) / 2;
                    if ((keys[m] >> lo) < top) a = m + 1; else b = m;
                }
            }
            u64 c = i, d;  // keys[c] is in the run; find the first index past it
            step = 1;
            for (;;) {
                const u64 m = c + step;
                if (m >= n) { d = n; break; }
                if ((keys[m] >> lo) != top) { d = m; break; }
                c = m;
                step <<= 1;
            }
            ++c;
            while (c < d) {
                const u64 m = c + (d - c) / 2;
                if ((keys[m] >> lo) <= top) c = m + 1; else d = m;
            }
            if (atomicExch(&mark[a], epoch) == epoch) continue;  // run already listed
            const u64 L = c - a;
            if (L <= TIE_SMALL) {
                const u32 ls_slot = atomicAdd(&lcount, 1u);
                if (ls_slot < TIE_LIST) {
                    ls[2 * ls_slot] = a;
                    ls[2 * ls_slot + 1] = c;
                } else {  // LDS list full: take a global slot directly
                    const u32 slot = atomicAdd(&counts[0], 1u);
                    if (slot >= cap_s) {
                        atomicOr(overflow, 1u);
                    } else {
                        runs_s[2 * (u64)slot] = a;
                        runs_s[2 * (u64)slot + 1] = c;
                    }
                }
            } else {
                const bool mid = L <= TIE_MID;
                const u32 slot = atomicAdd(&counts[mid ? 1 : 2], 1u);
                u64* runs = mid ? runs_m : runs_l;
                if (slot >= (mid ? TIE_MID_CAP : TIE_LONG_CAP)) {
                    atomicOr(overflow, 1u);
                } else {
                    runs[2 * (u64)slot] = a;
                    runs[2 * (u64)slot + 1] = c;
                }
            }
        }
    }
    __syncthreads();
    const u32 m = lcount < TIE_LIST ? lcount : TIE_LIST;
    if (threadIdx.x == 0) lbase = m ? atomicAdd(&counts[0], m) : 0u;
    __syncthreads();
    for (u32 q = threadIdx.x; q < m; q += TIE_BLK) {
        const u64 slot = (u64)lbase + q;
        if (slot >= cap_s) {
            atomicOr(overflow, 1u);
        } else {
            runs_s[2 * slot] = ls[2 * q];
            runs_s[2 * slot + 1] = ls[2 * q + 1];
        }
    }
}

// Small runs: wave w sorts runs w, w + waves, ...; lane l holds member l with
// the key (low bits << 6) | l (lanes past the run: all ones), a bitonic
// network over the next power of two sorts the keys, and lane m writes the
// member whose position the m-th key names.
__global__ __launch_bounds__(256) void tie_small_kernel(u64* __restrict__ keys,
                                                        u64* __restrict__ P,
                                                        const u64* __restrict__ sig,
                                                        const u64* __restrict__ runs,
                                                        const u32* __restrict__ count, u64 cap,
                                                        u32 lo) {
    const u32 lane = threadIdx.x & 63;
    const u64 nr = *count < cap ? *count : cap;
    const u64 lowmask = (1ULL << lo) - 1;
    const u64 nw = (u64)gridDim.x * (256 / 64);
    for (u64 r = ((u64)blockIdx.x * 256 + threadIdx.x) >> 6; r < nr; r += nw) {  // uniform
        const u64 a = runs[2 * r], c = runs[2 * r + 1];
        const u32 L = (u32)(c - a);
        u64 kv = 0, pv = 0, comp = ~0ULL;
        if (lane < L) {
            kv = keys[a + lane];
            pv = P ? P[a + lane] : 0;
            comp = (tie_low(kv, sig, lowmask) << 6) | lane;
        }
        u32 K = 2;
        while (K < L) K <<= 1;
        for (u32 kk = 2; kk <= K; kk <<= 1) {
            for (u32 j = kk >> 1; j > 0; j >>= 1) {
                const u64 other = __shfl_xor(comp, (int)j);
                const bool keep_min = ((lane & kk) == 0) == ((lane & j) == 0);
                const u64 mn = other < comp ? other : comp;
                const u64 mx = other < comp ? comp : other;
                comp = keep_min ? mn : mx;
            }
        }
        const int src = (int)(comp & 63);
        const u64 ks = __shfl(kv, src), ps = __shfl(pv, src);
        if (lane < L) {
            keys[a + lane] = ks;
            if (P) P[a + lane] = ps;
        }
    }
}

// Mid runs: workgroup g sorts runs g, g + grid, ... in LDS; keys
// (low bits << 12) | position, a bitonic network over the next power of two.
__global__ __launch_bounds__(256) void tie_mid_kernel(u64* __restrict__ keys, u64* __restrict__ P,
                                                      const u64* __restrict__ sig,
                                                      const u64* __restrict__ runs,
                                                      const u32* __restrict__ count, u32 lo) {
    __shared__ u64 s_comp[TIE_MID], s_k[TIE_MID], s_p[TIE_MID];
    const u64 nr = *count < TIE_MID_CAP ? *count : TIE_MID_CAP;
    const u64 lowmask = (1ULL << lo) - 1;
    for (u64 r = blockIdx.x; r < nr; r += gridDim.x) {  // uniform per workgroup
        const u64 a = runs[2 * r], c = runs[2 * r + 1];
        const u32 L = (u32)(c - a);
        u32 K = 2;
        while (K < L) K <<= 1;
        for (u32 i = threadIdx.x; i < K; i += 256) {
            if (i < L) {
                const u64 kv = keys[a + i];
                s_k[i] = kv;
                s_p[i] = P ? P[a + i] : 0;
                s_comp[i] = (tie_low(kv, sig, lowmask) << 12) | i;
            } else {
                s_comp[i] = ~0ULL;
            }
        }
        __syncthreads();
        for (u32 kk = 2; kk <= K; kk <<= 1) {
            for (u32 j = kk >> 1; j > 0; j >>= 1) {
                for (u32 i = threadIdx.x; i < K / 2; i += 256) {
                    const u32 l = ((i & ~(j - 1)) << 1) | (i & (j - 1));
                    const u32 m = l + j;
                    const u64 x = s_comp[l], y = s_comp[m];
                    if ((x > y) == ((l & kk) == 0)) {
                        s_comp[l] = y;
                        s_comp[m] = x;
                    }
                }
                __syncthreads();
            }
        }
        for (u32 i = threadIdx.x; i < L; i += 256) {
            const u32 src = (u32)(s_comp[i] & (TIE_MID - 1));
            keys[a + i] = s_k[src];
            if (P) P[a + i] = s_p[src];
        }
        __syncthreads();
    }
}
__global__ void tie_len_kernel(const u64* __restrict__ runs, const u32* __restrict__ count, u64 cap,
                               u64* __restrict__ len) {
    const u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= cap) return;
    len[r] = r < *count ? runs[2 * r + 1] - runs[2 * r] : 0;
}
// Compact element j belongs to run r (largest r with pre[r] <= j).
__global__ void tie_gather_kernel(const u64* __restrict__ keys, const u64* __restrict__ P,
                                  const u64* __restrict__ sig,
                                  const u64* __restrict__ runs, const u64* __restrict__ pre, u32 nr,
                                  u64 total, u32 lo, u64* __restrict__ cl, u64* __restrict__ cv,
                                  u64* __restrict__ crun, u64* __restrict__ cfull,
                                  u64* __restrict__ cP, u64* __restrict__ corig) {
    const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= total) return;
    u32 a = 0, b = nr - 1;
    while (a < b) {
        const u32 m = a + (b - a + 1) / 2;
        if (pre[m] <= j) a = m; else b = m - 1;
    }
    const u64 i = runs[2 * (u64)a] + (j - pre[a]);
    const u64 k = keys[i];
    cl[j] = tie_low(k, sig, (1ULL << lo) - 1);
    cv[j] = j;
    crun[j] = a;
    cfull[j] = k;
    cP[j] = P ? P[i] : 0;
    corig[j] = i;
}
__global__ void tie_key2_kernel(const u64* __restrict__ cv, const u64* __restrict__ crun, u64 total,
                                u64* __restrict__ ck) {
    const u64 m = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (m < total) ck[m] = crun[cv[m]];
}
__global__ void tie_scatter_kernel(const u64* __restrict__ cv, const u64* __restrict__ cfull,
                                   const u64* __restrict__ cP, const u64* __restrict__ corig,
                                   u64 total, u64* __restrict__ keys, u64* __restrict__ P) {
    const u64 m = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= total) return;
    const u64 src = cv[m];
    keys[corig[m]] = cfull[src];
    if (P) P[corig[m]] = cP[src];
}

#ifndef RBSC_RANK_ITEMS
#define RBSC_RANK_ITEMS 16
#endif
constexpr int RANK_ITEMS = RBSC_RANK_ITEMS;  // keys per thread of a rank tile
constexpr u64 RANK_TILE = 256 * RANK_ITEMS;

// Dense rank of (sig, colour) over the sorted sequence (rbsc.cu:392-396):
// flag = position starts a new distinct pair, rank = inclusive scan of flags.
// Writes Pout[i] = (rank << 32) | node (unless Pout is null) and
// hc[node] = randomize(rank + salt) (or the raw rank with c32).  A position
// whose key equals its predecessor's but whose colour is smaller (the
// sequence is not in (sig, colour) order there) sets *disorder.
//
// A workgroup takes a tile of RANK_TILE positions by ticket; wave w reads
// positions [base + 1024 w, + 1024) in 16 coalesced steps of 64 (lane l of
// step j holds position 64 j + l), takes each position's predecessor from the
// neighbouring lane (lane 63 of the step before for lane 0), counts flags
// with ballots, and ranks with the wave's running count plus the flags of
// lower lanes in the step.  The tile's count is published by decoupled
// look-back over earlier tiles.
__global__ __launch_bounds__(256) void rank_kernel(
    const u64* __restrict__ keys, const u64* __restrict__ Pin, u64* __restrict__ Pout,
    u64* __restrict__ hc, u64 n, u64 rank_offset, u64 salt_next, int c32, u64* status,
    u64* ticket, u64 ticket_base, u32 epoch, u32* K_out, u32* disorder, u32* err) {
    __shared__ u64 s_wave[4];
    __shared__ u64 s_tile, s_prefix;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd((unsigned long long*)ticket, 1ULL) - ticket_base;
    __syncthreads();
    const u64 tile = s_tile;
    const u64 wbase = tile * RANK_TILE + (u64)wave * (64 * RANK_ITEMS);
    u64 k[RANK_ITEMS], p[RANK_ITEMS];
#pragma unroll
    for (int j = 0; j < RANK_ITEMS; ++j) {
        const u64 i = wbase + 64 * j + lane;
        k[j] = i < n ? keys[i] : 0;
        p[j] = i < n ? Pin[i] : 0;
    }
    // predecessor of the wave's first position
    u64 lk = 0, lc = 0;
    if (wbase > 0 && wbase < n) {
        lk = keys[wbase - 1];
        lc = Pin[wbase - 1] >> 32;
    }
    u64 bal[RANK_ITEMS];
    u64 wcount = 0;
#pragma unroll
    for (int j = 0; j < RANK_ITEMS; ++j) {
        const u64 i = wbase + 64 * j + lane;
        const u64 c = p[j] >> 32;
        u64 pk = __shfl_up(k[j], 1);
        u64 pc = __shfl_up(c, 1);
        if (lane == 0) {
            pk = lk;
            pc = lc;
        }
        const bool f = i < n && i > 0 && (k[j] != pk || c != pc);
        if (i < n && i > 0 && k[j] == pk && c < pc) atomicOr(disorder, 1u);
        bal[j] = __ballot(f);
        wcount += (u64)__popcll(bal[j]);
        lk = __shfl(k[j], 63);  // predecessor of the next step's lane 0
        lc = __shfl(c, 63);
    }
    if (lane == 0) s_wave[wave] = wcount;
    __syncthreads();
    if (tid == 0) {
        const u64 agg = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        u64* st = status + tile;
        u64 excl = 0;
        if (tile == 0) {
            lb_store(st, lb_pack(LB_PREFIX, epoch, agg));
        } else {
            lb_store(st, lb_pack(LB_AGG, epoch, agg));
            long long q = (long long)tile - 1;
            u32 spins = 0;
            while (q >= 0) {
                const u64 sw = lb_load(status + q);
                const u64 state = sw >> LB_STATE_SHIFT;
                const u32 ep = (u32)((sw >> LB_EPOCH_SHIFT) & LB_EPOCH_MASK);
                if (state == 0 || ep != (epoch & LB_EPOCH_MASK)) {
                    if (++spins > LB_SPIN_LIMIT) {
                        atomicOr(err, KERR_SPIN);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += sw & LB_COUNT_MASK;
                if (state == LB_PREFIX) break;
                --q;
            }
            lb_store(st, lb_pack(LB_PREFIX, epoch, excl + agg));
        }
        s_prefix = excl;
    }
    __syncthreads();
    u64 r = s_prefix + rank_offset;
    for (int w = 0; w < wave; ++w) r += s_wave[w];
    const u64 below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    const u64 upto = below | (1ULL << lane);
#pragma unroll
    for (int j = 0; j < RANK_ITEMS; ++j) {
        const u64 i = wbase + 64 * j + lane;
        const u64 rj = r + (u64)__popcll(bal[j] & upto);
        r += (u64)__popcll(bal[j]);
        if (i < n) {
            const u32 node = (u32)p[j];
            if (Pout) Pout[i] = (rj << 32) | node;
            if (c32) reinterpret_cast<u32*>(hc)[node] = (u32)rj;
            else hc[node] = rbsc_randomize(rj + salt_next);
            if (i == n - 1) *K_out = (u32)(rj + 1);
        }
    }
}

// Dense rank over packed keys (node order, raw u32 colours): a key holds
// the signature's top bits and, below `lo`, the node (the sort moved 8 bytes
// per node).  A position starts a new (sig, colour) pair when its top bits
// differ from its predecessor's; only where they are equal are the low bits
// of both signatures (sig[node]) and both old colours (col[node]) read and
// compared (and a smaller colour after an equal signature sets *disorder).
// Writes col_out[node] = rank.  Tiles, waves and the decoupled look-back are
// those of rank_kernel.  With `members` (else null): every position whose
// pair equals its predecessor's appends its node and the predecessor's to
// members[0, *members_n) (wave-aggregated), so the list holds every node of
// every block of two or more nodes, some twice -- what the next round
// needs to see whether any block split (confirm_check_kernel).
__global__ __launch_bounds__(256) void rank_packed_kernel(
    const u64* __restrict__ keys, const u64* __restrict__ sig, const u32* __restrict__ col,
    u32* __restrict__ col_out, u64 n, u32 lo, u64 rank_offset, u64* status, u64* ticket,
    u64 ticket_base, u32 epoch, u32* K_out, u32* disorder, u32* err, u32* __restrict__ members,
    u32* __restrict__ members_n) {
    __shared__ u64 s_wave[4];
    __shared__ u64 s_tile, s_prefix;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd((unsigned long long*)ticket, 1ULL) - ticket_base;
    __syncthreads();
    const u64 tile = s_tile;
    const u64 wbase = tile * RANK_TILE + (u64)wave * (64 * RANK_ITEMS);
    const u64 lowmask = (1ULL << lo) - 1;
    u64 k[RANK_ITEMS];
#pragma unroll
    for (int j = 0; j < RANK_ITEMS; ++j) {
        const u64 i = wbase + 64 * j + lane;
        k[j] = i < n ? keys[i] : 0;
    }
    // neighbours of the wave's first and last positions
    const u64 wend = wbase + 64 * RANK_ITEMS;
    u64 lk = (wbase > 0 && wbase < n) ? keys[wbase - 1] : 0;
    const u64 lk0 = lk;
    const u64 nk_last = wend < n ? keys[wend] : 0;
    // low signature bits and colour of the first position's predecessor,
    // read only when it shares the first position's top bits
    u64 ls = 0;
    u32 lc = 0;
    if (wbase > 0 && wbase < n && (lk >> lo) == (k[0] >> lo)) {
        ls = sig[lk & lowmask] & lowmask;
        lc = col[lk & lowmask];
    }
    u64 bal[RANK_ITEMS];
    u64 wcount = 0;
#pragma unroll
    for (int j = 0; j < RANK_ITEMS; ++j) {
        const u64 i = wbase + 64 * j + lane;
        const u64 top = k[j] >> lo;
        u64 pk = __shfl_up(k[j], 1);
        if (lane == 0) pk = lk;
        u64 nk = __shfl_down(k[j], 1);
        if (j + 1 < RANK_ITEMS) {
            const u64 nx = __shfl(k[j + 1 < RANK_ITEMS ? j + 1 : j], 0);
            if (lane == 63) nk = nx;
        } else if (lane == 63) {
            nk = nk_last;
        }
        // a position in a run of equal top bits reads its own low bits and
        // colour once; its predecessor's come from the neighbouring lane
        const bool with_pred = i < n && i > 0 && top == (pk >> lo);
        const bool with_next = i + 1 < n && top == (nk >> lo);
        u64 so = 0;
        u32 co = 0;
        if (with_pred || with_next) {
            so = sig[k[j] & lowmask] & lowmask;
            co = col[k[j] & lowmask];
        }
        u64 sp = __shfl_up(so, 1);
        u32 cp = __shfl_up(co, 1);
        if (lane == 0) {
            sp = ls;
            cp = lc;
        }
        bool f = false;
        if (i < n && i > 0) {
            if (!with_pred) {
                f = true;
            } else {
                f = so != sp || co != cp;
                if (so == sp && co < cp) atomicOr(disorder, 1u);
            }
        }
        bal[j] = __ballot(f);
        wcount += (u64)__popcll(bal[j]);
        lk = __shfl(k[j], 63);  // predecessor of the next step's lane 0
        ls = __shfl(so, 63);
        lc = __shfl(co, 63);
    }
    if (lane == 0) s_wave[wave] = wcount;
    __syncthreads();
    if (tid == 0) {
        const u64 agg = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        u64* st = status + tile;
        u64 excl = 0;
        if (tile == 0) {
            lb_store(st, lb_pack(LB_PREFIX, epoch, agg));
        } else {
            lb_store(st, lb_pack(LB_AGG, epoch, agg));
            long long q = (long long)tile - 1;
            u32 spins = 0;
            while (q >= 0) {
                const u64 sw = lb_load(status + q);
                const u64 state = sw >> LB_STATE_SHIFT;
                const u32 ep = (u32)((sw >> LB_EPOCH_SHIFT) & LB_EPOCH_MASK);
                if (state == 0 || ep != (epoch & LB_EPOCH_MASK)) {
                    if (++spins > LB_SPIN_LIMIT) {
                        atomicOr(err, KERR_SPIN);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += sw & LB_COUNT_MASK;
                if (state == LB_PREFIX) break;
                --q;
            }
            lb_store(st, lb_pack(LB_PREFIX, epoch, excl + agg));
        }
        s_prefix = excl;
    }
    __syncthreads();
    u64 r = s_prefix + rank_offset;
    for (int w = 0; w < wave; ++w) r += s_wave[w];
    const u64 below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    const u64 upto = below | (1ULL << lane);
#pragma unroll
    for (int j = 0; j < RANK_ITEMS; ++j) {
        const u64 i = wbase + 64 * j + lane;
        const u64 rj = r + (u64)__popcll(bal[j] & upto);
        r += (u64)__popcll(bal[j]);
        if (i < n) {
            col_out[k[j] & lowmask] = (u32)rj;
            if (i == n - 1) *K_out = (u32)(rj + 1);
        }
    }
    if (members) {
        // positions sharing their predecessor's pair: counted over the whole
        // wave first, so the wave reserves its output with one atomic
        u32 total = 0;
#pragma unroll
        for (int j = 0; j < RANK_ITEMS; ++j) {
            const u64 i = wbase + 64 * j + lane;
            total += (u32)__popcll(__ballot(i < n && i > 0 && !((bal[j] >> lane) & 1ULL)));
        }
        if (total) {
            u32 at = 0;
            if (lane == 0) at = atomicAdd(members_n, 2u * total);
            at = __shfl(at, 0);
            u64 pk = lk0;  // predecessor of the step's lane 0
#pragma unroll
            for (int j = 0; j < RANK_ITEMS; ++j) {
                const u64 i = wbase + 64 * j + lane;
                u64 pred = __shfl_up(k[j], 1);
                if (lane == 0) pred = pk;
                pk = __shfl(k[j], 63);
                const bool same = i < n && i > 0 && !((bal[j] >> lane) & 1ULL);
                const u64 m = __ballot(same);
                if (same) {
                    const u32 o = at + 2u * (u32)__popcll(m & below);
                    members[o] = (u32)(k[j] & lowmask);
                    members[o + 1] = (u32)(pred & lowmask);
                }
                at += 2u * (u32)__popcll(m);
            }
        }
    }
}

// Did any block of the previous round split?  The previous round's
// rank_packed_kernel listed every node of every block of two or more nodes
// (members, some twice; singletons cannot split).  Each listed node names
// itself its block's representative (any member will do: the racing stores
// leave one of them), then compares its signature with the
// representative's; any difference sets *split.  No difference means the
// round's (signature, colour) pairs are as many as the old colours: the
// block count repeats and the refinement has reached its fixpoint.
__global__ void confirm_rep_kernel(const u32* __restrict__ members, u64 m,
                                   const u32* __restrict__ col, u32* __restrict__ rep) {
    const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        const u32 v = members[t];
        rep[col[v]] = v;
    }
}
__global__ void confirm_check_kernel(const u32* __restrict__ members, u64 m,
                                     const u32* __restrict__ col, const u64* __restrict__ sig,
                                     const u32* __restrict__ rep, u32* __restrict__ split) {
    const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        const u32 v = members[t];
        if (sig[v] != sig[rep[col[v]]]) atomicOr(split, 1u);
    }
}

// One old colour (the first round from a uniform start): the dense rank of
// (sig, colour) is the rank of sig among the distinct signatures, so no sort
// of all nodes is needed.  Each node finds or claims a slot of an
// open-addressed table of T = 2n slots (owner = node + 1; keys compared in
// full through sig[owner - 1]; linear probing, load at most 1/2); the
// distinct signatures are listed and sorted, each finds its slot again and
// stores its position there; a node's colour is its slot's position.
__device__ __forceinline__ u64 sig_slot(u64 s, u64 T) { return __umul64hi(rbsc_randomize(s), T); }

// Rows of one workgroup that share a signature first meet in an LDS table
// (INS_LDS slots for INS_BLK rows, so at most half full); only the first of
// them probes the global table, and the others take its slot.  Rows sharing
// a first-round signature (millions here) then reach a global slot's line
// once per workgroup instead of once per row.
constexpr int INS_BLK = 256;
constexpr u32 INS_LDS = 512;

__global__ __launch_bounds__(INS_BLK) void sig_insert_kernel(
    const u64* __restrict__ sig, u64 n, u32* __restrict__ owner, u64 T,
    u32* __restrict__ slot_of, u32* __restrict__ err) {
    __shared__ u64 l_sig[INS_BLK];
    __shared__ u32 l_lead[INS_LDS];  // row (in the workgroup) + 1 owning the slot; 0 = empty
    __shared__ u32 l_slot[INS_BLK];
    const int t = threadIdx.x;
    const u64 u = (u64)blockIdx.x * INS_BLK + t;
    const bool valid = u < n;
    const u64 s = valid ? sig[u] : 0;
    l_sig[t] = s;
    for (u32 i = t; i < INS_LDS; i += INS_BLK) l_lead[i] = 0;
    __syncthreads();
    const u64 r = rbsc_randomize(s);
    int lead = t;
    if (valid) {
        u32 h = (u32)r & (INS_LDS - 1);
        for (u32 p = 0; p < INS_LDS; ++p) {
            const u32 o = atomicCAS(&l_lead[h], 0u, (u32)(t + 1));
            if (o == 0) break;
            if (l_sig[o - 1] == s) {
                lead = (int)(o - 1);
                break;
            }
            h = (h + 1) & (INS_LDS - 1);
        }
    }
    if (valid && lead == t) {
        u64 h = __umul64hi(r, T);
        u32 found = 0;
        for (u64 probes = 0; probes < T; ++probes) {
            u32 o = __hip_atomic_load(&owner[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (o == 0) {
                const u32 prev = atomicCAS(&owner[h], 0u, (u32)(u + 1));
                if (prev == 0) {
                    found = 1;
                    break;
                }
                o = prev;
            }
            if (sig[o - 1] == s) {
                found = 1;
                break;
            }
            h = h + 1 == T ? 0 : h + 1;
        }
        if (!found) atomicOr(err, KERR_PROBE);  // not reached while the table is at most half full
        l_slot[t] = (u32)h;
    }
    __syncthreads();
    if (valid) slot_of[u] = l_slot[lead];
}

// Signatures of the occupied slots, appended in any order.  A workgroup
// covers LIST_ITEMS * 256 slots and reserves its range of the output with
// one atomic (the occupied slots are few and scattered, and one atomic per
// wave that met any serialised on the counter's address); inside it,
// positions come from ballots, so stores stay packed.
constexpr int LIST_ITEMS = 64;
constexpr u64 LIST_TILE = 256 * LIST_ITEMS;

__global__ __launch_bounds__(256) void sig_list_kernel(const u32* __restrict__ owner, u64 T,
                                                       const u64* __restrict__ sig,
                                                       u32* __restrict__ count,
                                                       u64* __restrict__ keys) {
    __shared__ u32 s_wave[4];
    __shared__ u32 s_base;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const u64 base = (u64)blockIdx.x * LIST_TILE + (u64)wave * (64 * LIST_ITEMS);
    u32 o[LIST_ITEMS];
    u32 wn = 0;
#pragma unroll
    for (int j = 0; j < LIST_ITEMS; ++j) {
        const u64 h = base + 64 * j + lane;
        o[j] = h < T ? owner[h] : 0u;
    }
#pragma unroll
    for (int j = 0; j < LIST_ITEMS; ++j) wn += (u32)__popcll(__ballot(o[j] != 0));
    if (lane == 0) s_wave[wave] = wn;
    __syncthreads();
    if (t == 0) {
        const u32 tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        s_base = tot ? atomicAdd(count, tot) : 0u;
    }
    __syncthreads();
    u32 off = s_base;
    for (int w = 0; w < wave; ++w) off += s_wave[w];
    const u64 below = lane ? (~0ULL >> (64 - lane)) : 0ULL;
#pragma unroll
    for (int j = 0; j < LIST_ITEMS; ++j) {
        const u64 bal = __ballot(o[j] != 0);
        if (o[j]) keys[off + (u32)__popcll(bal & below)] = sig[o[j] - 1];
        off += (u32)__popcll(bal);
    }
}

// sorted[i] (a listed signature) finds its slot and stores i there.
__global__ void sig_number_kernel(const u64* __restrict__ sorted, u64 D,
                                  const u32* __restrict__ owner, u64 T, const u64* __restrict__ sig,
                                  u32* __restrict__ num, u32* __restrict__ err) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const u64 s = sorted[i];
    u64 h = sig_slot(s, T);
    for (u64 probes = 0; probes < T; ++probes) {
        const u32 o = owner[h];
        if (o && sig[o - 1] == s) {
            num[h] = (u32)i;
            return;
        }
        h = h + 1 == T ? 0 : h + 1;
    }
    atomicOr(err, KERR_PROBE);  // every listed signature has a slot
}

__global__ void sig_colour_kernel(const u32* __restrict__ slot_of, const u32* __restrict__ num,
                                  u64 n, u64 rank_offset, u32* __restrict__ col_out) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n) col_out[u] = (u32)(rank_offset + num[slot_of[u]]);
}

// keys[i] = colour of node i, P[i] = (colour << 32) | i (raw u32 colours).
__global__ void colour_keys_kernel(const u32* __restrict__ col, u64 n, u64* __restrict__ keys,
                                   u64* __restrict__ P) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        keys[i] = col[i];
        P[i] = ((u64)col[i] << 32) | i;
    }
}

// Raw colours by internal node from the packed order (nodes beyond n keep
// the colour they were filled with).
__global__ void colors_kernel(const u64* __restrict__ P, u32* __restrict__ col, u64 n) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) col[(u32)P[i]] = (u32)(P[i] >> 32);
}
__global__ void fill_u32(u32* p, u64 n, u32 v) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void fill_u32_kernel(u32* p, u64 n, u32 v) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
// Colours by original id.
__global__ void map_colors_kernel(const u32* __restrict__ col_int, const u32* __restrict__ pi,
                                  u64 N, u32* __restrict__ out) {
    const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < N) out[v] = col_int[pi[v]];
}

// First-occurrence canonical labels on the device (the relabel of
// rbsc.cu:431-445): first[c] = least original id of colour c; the first
// occurrences, flagged in original id order and scanned, number the blocks;
// a node's label is the number of its colour's first occurrence.
// first[] comes from a stable sort of (colour, id) pairs by colour: each
// colour's run starts with its least id (one random atomic per node, the
// earlier form, ran at about 3 G nodes/s on graphs of mostly singleton
// blocks).  (Colours above N are not expected; one sets *bad, and the
// caller then relabels on the host.)
__global__ void canon_pairs_kernel(const u32* __restrict__ col, u64 N, u64* __restrict__ pairs,
                                   u32* __restrict__ bad) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= N) return;
    const u32 c = col[u];
    if (c > N) atomicOr(bad, 1u);
    pairs[u] = ((u64)c << 32) | u;
}
__global__ void canon_first_kernel(const u64* __restrict__ sorted, u64 N, u32* __restrict__ first) {
    const u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const u64 k = sorted[p];
    if (p == 0 || (sorted[p - 1] >> 32) != (k >> 32)) first[k >> 32] = (u32)k;
}
__global__ void canon_flag_kernel(const u32* __restrict__ col, const u32* __restrict__ first, u64 N,
                                  u64* __restrict__ flag) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= N) return;
    const u32 c = col[u];
    flag[u] = c <= N && first[c] == (u32)u ? 1 : 0;
}
__global__ void canon_label_kernel(const u32* __restrict__ col, const u32* __restrict__ first,
                                   const u64* __restrict__ rank, u64 N, u32* __restrict__ out) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= N) return;
    const u32 c = col[u];
    out[u] = c <= N ? (u32)rank[first[c]] : 0u;
}


// Block sizes of a colouring with K dense colours over N nodes, in
// descending order, then zeros (N entries): size[c] nodes per colour, then
// bys[i] = blocks of size N - i, exclusively scanned, so entry j < K of the
// descending order is N - i for the last i whose scan is <= j.
__global__ void colour_size_kernel(const u32* __restrict__ col, u64 N, u32 K,
                                   u32* __restrict__ size) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < N && col[u] < K) atomicAdd(&size[col[u]], 1u);
}
__global__ void size_count_kernel(const u32* __restrict__ size, u32 K, u64 N,
                                  u64* __restrict__ bys) {
    const u64 c = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < K && size[c] >= 1 && size[c] <= N) atomicAdd((unsigned long long*)&bys[N - size[c]], 1ull);
}
__global__ void sizes_desc_kernel(const u64* __restrict__ bys, u64 N, u32 K, u32* __restrict__ out) {
    const u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    u32 v = 0;
    if (j < K) {
        u64 lo = 0, hi = N - 1;
        while (lo < hi) {
            const u64 mid = (lo + hi + 1) >> 1;
            if (bys[mid] <= j) lo = mid; else hi = mid - 1;
        }
        v = (u32)(N - lo);
    }
    out[j] = v;
}

}  // namespace

// ---------------------------------------------------------------- host side

u32 canonicalize_host(u32 N, const u32* colors, u32* out) {
    std::vector<u32> map;
    u32 next = 0;
    u32 maxc = 0;
    for (u32 i = 0; i < N; ++i) maxc = std::max(maxc, colors[i]);
    map.assign((size_t)maxc + 1, UINT32_MAX);
    for (u32 i = 0; i < N; ++i) {
        u32& m = map[colors[i]];
        if (m == UINT32_MAX) m = next++;
        out[i] = m;
    }
    return next;
}

void prof_begin(rbsc_ctx* ctx, int phase) {
    if (!ctx->prof) return;
    if (ctx->prof_used + 2 > ctx->prof_ev.size()) {
        for (int i = 0; i < 64; ++i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            ctx->prof_ev.push_back(e);
        }
    }
    (void)hipEventRecord(ctx->prof_ev[ctx->prof_used], ctx->stream);
    ctx->prof_phase.push_back(phase);
    ctx->prof_used += 1;
}

void prof_end(rbsc_ctx* ctx) {
    if (!ctx->prof || ctx->prof_used % 2 == 0) return;
    (void)hipEventRecord(ctx->prof_ev[ctx->prof_used], ctx->stream);
    ctx->prof_used += 1;
}

int prof_collect(rbsc_ctx* ctx) {
    if (!ctx->prof) return RBSC_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i + 1 < ctx->prof_used; i += 2) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->prof_ev[i], ctx->prof_ev[i + 1]));
        const int ph = ctx->prof_phase[i / 2];
        ctx->prof_ms[ph] += ms;
        ctx->prof_n[ph] += 1;
    }
    ctx->prof_used = 0;
    ctx->prof_phase.clear();
    return RBSC_OK;
}

int check_kernel_errors(rbsc_ctx* ctx) {
    u32 e = 0;
    HIP_TRY(hipMemcpy(&e, ctx->d_err, sizeof(u32), hipMemcpyDeviceToHost));
    if (e) {
        HIP_TRY(hipMemset(ctx->d_err, 0, sizeof(u32)));
        rbsc_set_error("kernel error flags 0x%x (%s)", e,
                       (e & KERR_SPIN) ? "look-back wait expired" : "signature table full");
        return RBSC_ERR_KERNEL;
    }
    return RBSC_OK;
}

void ws_release(rbsc_ctx* ctx) {
    Workspace& w = ctx->ws;
    hipStream_t s = ctx->stream;
    for (int i = 0; i < 2; ++i) {
        ctx_free(w.keys[i], s);
        ctx_free(w.P[i], s);
    }
    for (void* p : {(void*)w.sig, (void*)w.hc, (void*)w.hc2, (void*)w.rank_status, w.cold,
                    (void*)w.tie_runs, (void*)w.tie_runs_m, (void*)w.tie_runs_l,
                    (void*)w.tie_pre, (void*)w.tie_tmp, (void*)w.tie_mark, (void*)w.tie_buf,
                    (void*)w.carry})
        ctx_free(p, s);
    w = Workspace();
    ctx->sort.extra_status = nullptr;
    ctx->sort.extra_words = 0;
}

int ws_reserve(rbsc_ctx* ctx, u64 n) {
    Workspace& w = ctx->ws;
    if (n <= w.cap_n && w.keys[0]) return RBSC_OK;
    // grows by half at least, so a context serving graphs of growing sizes
    // reallocates a few times only
    const u64 cap = std::max<u64>(std::max<u64>(n, 4096), w.cap_n + w.cap_n / 2);
    ws_release(ctx);
    hipStream_t s = ctx->stream;
    for (int i = 0; i < 2; ++i) {
        HIP_TRY(ctx_alloc(&w.keys[i], sizeof(u64) * cap, s));
        HIP_TRY(ctx_alloc(&w.P[i], sizeof(u64) * cap, s));
    }
    HIP_TRY(ctx_alloc(&w.sig, sizeof(u64) * (cap + 1), s));  // + spare slot for padding
    HIP_TRY(ctx_alloc(&w.hc, sizeof(u64) * cap, s));
    HIP_TRY(ctx_alloc(&w.hc2, sizeof(u64) * cap, s));
    w.tie_cap = cap / 16 + 1024;
    HIP_TRY(ctx_alloc(&w.tie_runs, sizeof(u64) * 2 * w.tie_cap, s));
    HIP_TRY(ctx_alloc(&w.tie_runs_m, sizeof(u64) * 2 * TIE_MID_CAP, s));
    HIP_TRY(ctx_alloc(&w.tie_runs_l, sizeof(u64) * 2 * TIE_LONG_CAP, s));
    HIP_TRY(ctx_alloc(&w.tie_pre, sizeof(u64) * (TIE_LONG_CAP + 1), s));
    HIP_TRY(ctx_alloc(&w.tie_tmp, sizeof(u64) * scan_tmp_words(TIE_LONG_CAP), s));
    HIP_TRY(ctx_alloc(&w.tie_mark, sizeof(u32) * cap, s));
    HIP_TRY(hipMemsetAsync(w.tie_mark, 0, sizeof(u32) * cap, s));
    w.tie_epoch = 0;
    w.rank_tiles = (cap + RANK_TILE - 1) / RANK_TILE + 1;
    HIP_TRY(ctx_alloc(&w.rank_status, sizeof(u64) * w.rank_tiles, s));
    HIP_TRY(hipMemsetAsync(w.rank_status, 0, sizeof(u64) * w.rank_tiles, s));
    w.cap_n = cap;
    ctx->sort.extra_status = w.rank_status;
    ctx->sort.extra_words = w.rank_tiles;
    int rc = sort_ws_init(&ctx->sort, cap, ctx->d_err, s);
    if (rc) return rc;
    return RBSC_OK;
}

void dgraph_release(rbsc_dgraph* g, hipStream_t s) {
    if (!g) return;
    ctx_free(g->row_ptr, s);
    ctx_free(g->pi, s);
    ctx_free(g->dst, s);
    ctx_free(g->wt, s);
    ctx_free(g->init, s);
    ctx_free(g->end_bits, s);
    ctx_free(g->seg_row, s);
    ctx_free(g->d_part_tile, s);
    if (g->dst_host) (void)hipHostFree(g->dst_host);
    if (g->wt_host) (void)hipHostFree(g->wt_host);
    ctx_free(g->tile_row, s);
    ctx_free(g->carry_row, s);
    ctx_free(g->dup_rep, s);
    ctx_free(g->dang_w, s);
    ctx_free(g->row_w, s);
    if (g->bins) bins_release(g->bins, s);
    *g = rbsc_dgraph();
}

// Host-side compression of a graph into rows (validates ids and grouping).
static int compress_host(const rbsc_graph* gr, std::vector<u64>& row_ptr,
                         std::vector<u32>& row_node) {
    const u32 N = gr->n_nodes;
    const u64 E = gr->n_edges;
    row_ptr.clear();
    row_node.clear();
    if (gr->row_offsets) {
        const u64* ro = gr->row_offsets;
        if (ro[0] != 0 || ro[N] != E) {
            rbsc_set_error("row_offsets must start at 0 and end at n_edges");
            return RBSC_ERR_INPUT;
        }
        for (u32 u = 0; u < N; ++u) {
            if (ro[u + 1] < ro[u]) {
                rbsc_set_error("row_offsets not monotone at node %u", u);
                return RBSC_ERR_INPUT;
            }
            if (ro[u + 1] > ro[u]) {
                row_ptr.push_back(ro[u]);
                row_node.push_back(u);
            }
        }
    } else {
        if (!gr->src && E) {
            rbsc_set_error("graph needs src or row_offsets");
            return RBSC_ERR_INVALID;
        }
        std::vector<unsigned char> seen(N, 0);
        for (u64 e = 0; e < E; ++e) {
            const u32 s = gr->src[e];
            if (s >= N) {
                rbsc_set_error("line %llu: source %u >= N=%u", (unsigned long long)e, s, N);
                return RBSC_ERR_INPUT;
            }
            if (e == 0 || gr->src[e - 1] != s) {
                if (seen[s]) {
                    rbsc_set_error("line %llu: lines of source %u are not grouped",
                                   (unsigned long long)e, s);
                    return RBSC_ERR_INPUT;
                }
                seen[s] = 1;
                row_ptr.push_back(e);
                row_node.push_back(s);
            }
        }
    }
    row_ptr.push_back(E);
    for (u64 e = 0; e < E; ++e) {
        if (gr->dst[e] >= N) {
            rbsc_set_error("line %llu: destination %u >= N=%u", (unsigned long long)e, gr->dst[e], N);
            return RBSC_ERR_INPUT;
        }
    }
    return RBSC_OK;
}

// ---------------------------------------------------------------- refine
static void launch_sig(rbsc_ctx* ctx, rbsc_dgraph* g, const u32* dst, const u64* wt,
                       u64 t_begin, u64 t_end, u64 line_base, u32 gather_limit, u64 dang_h,
                       u64 salt, bool c32, bool xcd = false) {
    if (t_end <= t_begin) return;
    Workspace& w = ctx->ws;
    unsigned grid = grid_for((t_end - t_begin) * 64);
    if (xcd) {  // 8 partitions, each on its own XCD
        u64 most = 0;
        for (u32 p = 0; p < 8; ++p) most = std::max(most, g->part_tile[p + 1] - g->part_tile[p]);
        grid = (unsigned)(8 * ((most + 3) / 4));
    }
    const unsigned char* eb = reinterpret_cast<const unsigned char*>(g->end_bits);
    const u64 E = g->n_parts > 1 ? g->E_pad : g->E;
#define SIG_LAUNCH(W, PT, C)                                                                   \
    sig_kernel<W, PT, C><<<grid, BLK, 0, ctx->stream>>>(dst, W ? wt : nullptr, E, eb, w.hc,    \
                                                        gather_limit, dang_h, salt, w.sig,     \
                                                        g->seg_row, g->tile_row, w.carry,      \
                                                        t_begin, t_end, line_base, g->d_part_tile)
#define SIG_C(W, PT)                   \
    if (c32) SIG_LAUNCH(W, PT, true);  \
    else SIG_LAUNCH(W, PT, false)
    if (xcd) {
        if (wt) { SIG_C(true, 2); }
        else { SIG_C(false, 2); }
    } else if (g->n_parts > 1) {
        if (wt) { SIG_C(true, 1); }
        else { SIG_C(false, 1); }
    } else {
        if (wt) { SIG_C(true, 0); }
        else { SIG_C(false, 0); }
    }
#undef SIG_C
#undef SIG_LAUNCH
}

static int stream_reserve(rbsc_ctx* ctx, u64 lines, bool weighted) {
    if (ctx->sbuf_lines >= lines && (!weighted || ctx->swbuf[0])) return RBSC_OK;
    for (int b = 0; b < 2; ++b) {
        ctx_free(ctx->sbuf[b], ctx->stream);
        ctx_free(ctx->swbuf[b], ctx->stream);
        ctx->sbuf[b] = nullptr;
        ctx->swbuf[b] = nullptr;
    }
    ctx->sbuf_lines = 0;
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(ctx_alloc(&ctx->sbuf[b], sizeof(u32) * lines + 32, ctx->stream));
        if (weighted) HIP_TRY(ctx_alloc(&ctx->swbuf[b], sizeof(u64) * lines, ctx->stream));
    }
    ctx->sbuf_lines = lines;
    return RBSC_OK;
}

int refine_core(rbsc_ctx* ctx, rbsc_dgraph* g, const RefineParams& prm, RefineOut* out) {
    hipStream_t s = ctx->stream;
    const u64 N = g->N;
    *out = RefineOut();
    if (N == 0) return RBSC_OK;
    int rc = ws_reserve(ctx, N);
    if (rc) return rc;
    Workspace& w = ctx->ws;
    if (g->n_tiles > w.carry_cap || !w.carry) {
        ctx_free(w.carry, s);
        w.carry = nullptr;
        w.carry_cap = 0;
        HIP_TRY(ctx_alloc(&w.carry, sizeof(u64) * std::max<u64>(g->n_tiles, 1), s));
        w.carry_cap = std::max<u64>(g->n_tiles, 1);
    }
    // Without initial colours, nodes without lines are always colour 0: their
    // pair (0, 0) is the least of every round (others have colour >= 1 from
    // c0 = N+1 on), so they are left out of the node side, ranks of the rest
    // are shifted by one, and their hashed colour is the constant
    // randomize(0 + r).  With initial colours every node takes part.
    const bool excluded = g->init == nullptr;
    if (prm.d_sig_fill && excluded) {
        rbsc_set_error("signature defaults need initial colours");
        return RBSC_ERR_INVALID;
    }
    // Without initial colours, a row that repeats a stored row has the same
    // signature as it every round and starts with the same colour, so it is
    // always in its stored row's block: repeated rows [R, U) are left out of
    // the node side too (their pairs equal their stored rows' pairs, so the
    // dense ranks are unchanged) and take their stored rows' colours.  With
    // initial colours every node takes part.
    const u64 n_rank = excluded ? g->R : N;
    const u64 rank_offset = excluded && N > g->U ? 1 : 0;
    const u32 gather_limit = excluded ? g->U : (u32)N;
    const bool streaming = g->dst == nullptr && g->E > 0;
    // The binned layout serves resident rounds.  Its hashed form keeps u64
    // hashed colours per node; its raw form gathers u32 raw colours (and is
    // left out when hashed colours are asked for).
    const bool bins_fit = g->bins && !streaming && g->n_parts == 1 &&
                          g->bins->GL == gather_limit && !(prm.flags & RBSC_FLAG_NOBINS);
    const bool binned = bins_fit && (g->bins->hashed || !(prm.flags & RBSC_FLAG_HASHED64));
    const int c32 = binned ? (g->bins->hashed ? 0 : 1) : ((prm.flags & RBSC_FLAG_HASHED64) ? 0 : 1);
    if (binned && bins_cold_bytes(g->bins) > w.cold_bytes) {
        // the rounds' cold-slot colours belong to this refinement, not to the
        // graph, so concurrent refinements of one graph do not share them
        ctx_free(w.cold, s);
        w.cold = nullptr;
        w.cold_bytes = 0;
        HIP_TRY(ctx_alloc(&w.cold, bins_cold_bytes(g->bins), s));
        HIP_TRY(hipMemsetAsync(w.cold, 0, bins_cold_bytes(g->bins), s));
        w.cold_bytes = bins_cold_bytes(g->bins);
    }
    const bool follow = excluded && g->U > g->R;
    const bool node_order = excluded && c32 && !(prm.flags & RBSC_FLAG_COLOUR_ORDER);

    // Round-0 state: P sorted by c0, hc = randomize(c0 + 0).
    prof_begin(ctx, 4);
    fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar + 3, 1, 0);
    int which = 0;
    int pc = 0;
    if (excluded) {
        if (n_rank)
            init_uniform_kernel<<<grid_for(n_rank), BLK, 0, s>>>(w.P[0], w.hc, n_rank, N + 1, c32);
    } else {
        init_given_kernel<<<grid_for(N), BLK, 0, s>>>(g->init, w.keys[0], w.P[0], w.hc, N, c32);
        u64* kk[2] = {w.keys[0], w.keys[1]};
        u64* pp[2] = {w.P[0], w.P[1]};
        rc = radix_sort_u64(&ctx->sort, kk, pp, N, 0, 32, &which, s);
        if (rc) return rc;
        pc = which;
    }
    // Signatures of nodes without lines are 0 (rbsc.cu:104-116) and no line
    // kernel writes them; without initial colours those nodes are not ranked
    // and every round writes the stored rows' signatures before reading them.
    if (prm.d_sig_fill)
        permute_fill_kernel<<<grid_for(N), BLK, 0, s>>>(prm.d_sig_fill, g->pi, N, w.sig);
    else if (!excluded)
        fill_u64<<<grid_for(N), BLK, 0, s>>>(w.sig, N, 0);
    prof_end(ctx);

    // Streaming set-up: slices of whole wave tiles, double-buffered.
    u64 Bs = 0, n_slices = 0;
    const bool weighted = streaming ? g->wt_host != nullptr : g->wt != nullptr;
    if (streaming) {
        if (!g->dst_host) {
            rbsc_set_error("graph has no line array");
            return RBSC_ERR_INVALID;
        }
        const u64 want = prm.slice_lines ? prm.slice_lines : g->E;
        Bs = std::max<u64>(SIG_TW, (want / SIG_TW) * SIG_TW);
        n_slices = (g->E + Bs - 1) / Bs;
        if (n_slices <= 2) Bs = std::min<u64>(Bs, ((g->E + SIG_TW - 1) / SIG_TW) * SIG_TW);
        rc = stream_reserve(ctx, Bs, weighted);
        if (rc) return rc;
        for (int b = 0; b < 2; ++b) HIP_TRY(hipEventRecord(ctx->ev_done[b], s));
    }
    bool prefetched[2] = {false, false};
    bool streamed_once = false;  // a round has streamed the slices into the buffers
    auto issue_copy = [&](u64 sl) -> int {
        const int b = (int)(sl & 1);
        const u64 lo = sl * Bs, hi = std::min<u64>(g->E, lo + Bs);
        HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, ctx->ev_done[b], 0));
        HIP_TRY(hipMemcpyAsync(ctx->sbuf[b], g->dst_host + lo, sizeof(u32) * (hi - lo),
                               hipMemcpyHostToDevice, ctx->copy_stream));
        if (weighted)
            HIP_TRY(hipMemcpyAsync(ctx->swbuf[b], g->wt_host + lo, sizeof(u64) * (hi - lo),
                                   hipMemcpyHostToDevice, ctx->copy_stream));
        HIP_TRY(hipEventRecord(ctx->ev_copy[b], ctx->copy_stream));
        out->pcie_bytes += (double)(hi - lo) * (weighted ? 12.0 : 4.0);
        return RBSC_OK;
    };

    HIP_TRY(hipEventRecord(ctx->ev0, s));
    u32 K = prm.k0, Knew = prm.k0;
    u32 it = 0;
    const u64 n_tiles = g->n_tiles;
    // Repeated rows' colours are kept only where lines still gather them (the
    // binned layout points its lines at the stored rows instead).
    const bool copy_rep = follow && !(binned && g->bins->remap);
    auto rep_colors = [&](u64* hcbuf) {
        if (copy_rep)
            dup_colors_kernel<<<grid_for(g->U - g->R), BLK, 0, s>>>(g->dup_rep, g->R,
                                                                    g->U - g->R, hcbuf, c32 ? 0 : 1);
    };
    rep_colors(w.hc);
    // The last round (the one whose block count repeats) is confirmed from
    // the previous round's blocks of two or more nodes instead of being
    // sorted and ranked, when only the partition is read afterwards: its
    // colours would name the same blocks as the previous round's (see
    // confirm_check_kernel).  Callers that read raw colours or stop after a
    // given round get the ranked round.
    const bool confirm_ok = node_order && !prm.max_rounds && !prm.exact_colours &&
                            !(prm.flags & RBSC_FLAG_FULLROUNDS);
    bool have_members = false;  // members lists the previous round's multi-node blocks
    u64 members_n = 0;
    u32* members = reinterpret_cast<u32*>(w.keys[1]);  // 2 cap_n words; free between rounds
    do {
        K = Knew;
        const u64 dang_h = rbsc_randomize(0 + (u64)it);
        // First round from a uniform start: signatures in closed form from the
        // per-row weights (first_round_sig_kernel), on every layout; the
        // colour of the nodes with lines is N+1, hashed as the round-0 state
        // holds it (raw u32 or u64).
        const bool closed_first =
            excluded && it == 0 && g->dang_w && !(prm.flags & RBSC_FLAG_GATHER_FIRST);
        if (binned && !closed_first) {
            prof_begin(ctx, 6);
            bins_scatter(ctx, g, w.hc, w.cold);
            prof_end(ctx);
        }
        if (closed_first) {  // timed with the round-0 set-up: it reads no line
            prof_begin(ctx, 4);
            const u64 hot_h = rbsc_randomize(c32 ? (u64)(u32)(N + 1) : N + 1);
            if (g->R)
                first_round_sig_kernel<<<grid_for(g->R), BLK, 0, s>>>(g->row_ptr, g->row_w, g->dang_w,
                                                                     g->R, hot_h, dang_h, w.sig);
            prof_end(ctx);
        } else {
        prof_begin(ctx, 0);
        if (binned) {
            bins_sig(ctx, g, w.hc, w.cold, (u64)it, dang_h, w.sig);
        } else if (n_tiles && !streaming && g->n_parts == 8 && (prm.flags & RBSC_FLAG_XCD)) {
            fill_u64<<<grid_for(g->U), BLK, 0, s>>>(w.sig, g->U, 0);
            launch_sig(ctx, g, g->dst, g->wt, 0, n_tiles, 0, gather_limit, dang_h, it, c32, true);
        } else if (n_tiles && !streaming && g->n_parts > 1) {
            fill_u64<<<grid_for(g->U), BLK, 0, s>>>(w.sig, g->U, 0);
            for (u32 p = 0; p < g->n_parts; ++p)
                launch_sig(ctx, g, g->dst, g->wt, g->part_tile[p], g->part_tile[p + 1], 0,
                           gather_limit, dang_h, it, c32);
        } else if (n_tiles && !streaming) {
            launch_sig(ctx, g, g->dst, g->wt, 0, n_tiles, 0, gather_limit, dang_h, it, c32);
        } else if (n_tiles) {
            // both slices stay in the buffers once a round has streamed them
            const bool resident = n_slices <= 2 && streamed_once;
            for (u64 sl = 0; sl < n_slices; ++sl) {
                const int b = (int)(sl & 1);
                if (!resident && sl < 2 && !prefetched[b]) {  // later slices were issued
                    rc = issue_copy(sl);                         // two iterations earlier
                    if (rc) return rc;
                }
                if (!resident) HIP_TRY(hipStreamWaitEvent(s, ctx->ev_copy[b], 0));
                const u64 lo = sl * Bs, hi = std::min<u64>(g->E, lo + Bs);
                launch_sig(ctx, g, ctx->sbuf[b], weighted ? ctx->swbuf[b] : nullptr, lo / SIG_TW,
                           (hi + SIG_TW - 1) / SIG_TW, lo, gather_limit, dang_h, it, c32);
                HIP_TRY(hipEventRecord(ctx->ev_done[b], s));
                if (!resident && sl + 2 < n_slices) {
                    rc = issue_copy(sl + 2);
                    if (rc) return rc;
                }
            }
            prefetched[0] = prefetched[1] = false;
            streamed_once = true;
            if (n_slices > 2) {  // next round's first slices overlap this round's node phase
                for (u64 sl = 0; sl < 2; ++sl) {
                    rc = issue_copy(sl);
                    if (rc) return rc;
                    prefetched[sl] = true;
                }
            }
        }
        prof_end(ctx);
        }  // the line pass
        prof_begin(ctx, 5);
        if (n_tiles && !binned && !closed_first)
            sig_fixup_kernel<<<grid_for(n_tiles), BLK, 0, s>>>(g->carry_row, w.carry,
                                                               g->n_parts > 1 ? g->seg_row : nullptr,
                                                               w.sig, n_tiles);
        if (g->U > g->R && !follow)  // ranked repeats need their signatures
            dup_sig_kernel<<<grid_for(g->U - g->R), BLK, 0, s>>>(g->dup_rep, g->R, g->U - g->R,
                                                                 w.sig);
        prof_end(ctx);
        if (confirm_ok && have_members && members_n <= n_rank / 4) {
            // no block split: the block count repeats, fixpoint (K stays)
            prof_begin(ctx, 3);
            u32* rep = reinterpret_cast<u32*>(w.P[0]);  // one word per old colour (< K)
            const u32* hc32 = reinterpret_cast<const u32*>(w.hc);
            fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar + 14, 1, 0);
            if (members_n) {
                confirm_rep_kernel<<<grid_for(members_n), BLK, 0, s>>>(members, members_n, hc32, rep);
                confirm_check_kernel<<<grid_for(members_n), BLK, 0, s>>>(members, members_n, hc32,
                                                                         w.sig, rep,
                                                                         ctx->d_scalar + 14);
            }
            HIP_TRY(hipMemcpyAsync(ctx->h_scalar + 14, ctx->d_scalar + 14, sizeof(u32),
                                   hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            prof_end(ctx);
            if (ctx->h_scalar[14] == 0) {
                ++ctx->confirmed_rounds;
                ++it;
                break;
            }
        }
        have_members = false;
        // Node side: keys in colour order, stable sort by signature, dense rank.
        // sorted_pc is the P buffer holding the sorted order (input of rank).
        int sorted_pc = pc;
        // Orders every run of equal top bits by the low bits (see
        // tie_find_kernel): small and mid runs on the device; long runs are
        // left for fix_long after the round's host sync (d_scalar[10] counts
        // them).  Sets d_scalar[1] when a run list overflows.
        const u32 lo = (prm.flags & RBSC_FLAG_SPLIT16) ? 48 : 32;
        // packed node-order sort this round (see sort_and_rank_nodes); K is
        // the previous round's block count
        const bool packed_round =
            node_order && !(prm.flags & (RBSC_FLAG_FULLSORT | RBSC_FLAG_UNPACKED)) &&
            ((prm.flags & RBSC_FLAG_PACKED) || (u64)K * 2 >= (u64)n_rank) &&
            (n_rank >> lo) == 0;
        auto fix_ties = [&](u64* keys, u64* P, const u64* psig) {
            const u64 cap = (prm.flags & RBSC_FLAG_TIESMALL) ? 2 : w.tie_cap;
            if (++w.tie_epoch == 0) {  // wrapped: clear the stamps
                fill_u32_kernel<<<grid_for(w.cap_n), BLK, 0, s>>>(w.tie_mark, w.cap_n, 0);
                w.tie_epoch = 1;
            }
            fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar + 8, 3, 0);
            const unsigned tg = (unsigned)std::min<u64>((n_rank + TF_TILE - 1) / TF_TILE, TIE_GRID);
            tie_find_kernel<<<tg, TIE_BLK, 0, s>>>(keys, psig, n_rank, lo, w.tie_mark, w.tie_epoch,
                                                   w.tie_runs, cap, w.tie_runs_m, w.tie_runs_l,
                                                   ctx->d_scalar + 8, ctx->d_scalar + 1);
            tie_small_kernel<<<1024, 256, 0, s>>>(keys, P, psig, w.tie_runs, ctx->d_scalar + 8, cap,
                                                  lo);
            tie_mid_kernel<<<256, 256, 0, s>>>(keys, P, psig, w.tie_runs_m, ctx->d_scalar + 9, lo);
        };
        // The nl long runs listed this round (host-driven): members gathered
        // into a compact array, sorted by low bits and then by run (both
        // stable), scattered back.
        auto fix_long = [&](u64* keys, u64* P, u32 nl, const u64* psig) -> int {
            tie_len_kernel<<<grid_for(TIE_LONG_CAP), BLK, 0, s>>>(w.tie_runs_l, ctx->d_scalar + 10,
                                                                  TIE_LONG_CAP, w.tie_pre);
            int r3 = scan_excl_u64(w.tie_pre, w.tie_pre, TIE_LONG_CAP, w.tie_pre + TIE_LONG_CAP,
                                   w.tie_tmp, s);
            if (r3) return r3;
            u64 total = 0;
            HIP_TRY(hipMemcpyAsync(&total, w.tie_pre + TIE_LONG_CAP, sizeof(u64),
                                   hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (total > w.tie_buf_cap) {
                ctx_free(w.tie_buf, s);
                w.tie_buf = nullptr;
                w.tie_buf_cap = 0;
                const u64 want = std::max<u64>(total, 4096);
                HIP_TRY(ctx_alloc(&w.tie_buf, sizeof(u64) * 8 * want, s));
                w.tie_buf_cap = want;
            }
            const u64 C = w.tie_buf_cap;
            u64* cl[2] = {w.tie_buf, w.tie_buf + C};
            u64* cv[2] = {w.tie_buf + 2 * C, w.tie_buf + 3 * C};
            u64 *crun = w.tie_buf + 4 * C, *cfull = w.tie_buf + 5 * C, *cP = w.tie_buf + 6 * C,
                *corig = w.tie_buf + 7 * C;
            tie_gather_kernel<<<grid_for(total), BLK, 0, s>>>(keys, P, psig, w.tie_runs_l, w.tie_pre, nl,
                                                             total, lo, cl[0], cv[0], crun, cfull,
                                                             cP, corig);
            int wa = 0;
            r3 = radix_sort_u64(&ctx->sort, cl, cv, total, 0, (int)lo, &wa, s);  // by low bits
            if (r3) return r3;
            u64* ck[2] = {cl[wa ^ 1], cl[wa]};
            u64* cw[2] = {cv[wa], cv[wa ^ 1]};
            tie_key2_kernel<<<grid_for(total), BLK, 0, s>>>(cv[wa], crun, total, ck[0]);
            int rb = 8;
            while (rb < 32 && ((u64)(nl - 1) >> rb)) rb += 8;
            int wb = 0;
            r3 = radix_sort_u64(&ctx->sort, ck, cw, total, 0, rb, &wb, s);  // then by run
            if (r3) return r3;
            tie_scatter_kernel<<<grid_for(total), BLK, 0, s>>>(cw[wb], cfull, cP, corig, total,
                                                               keys, P);
            return RBSC_OK;
        };
        // In rounds whose previous block count is below half the ranked
        // nodes (the early rounds, where a large class sharing top bits with
        // another signature makes a run too long for the in-kernel sorts),
        // the long runs are read back and fixed before the rank pass, so it
        // runs once; in the other rounds they are rare and fixed after the
        // round's sync, with the rank pass run again.
        const bool long_before_rank = (u64)K * 2 < (u64)n_rank;
        auto fix_long_first = [&](u64* keys, u64* P, const u64* psig) -> int {
            HIP_TRY(hipMemcpyAsync(ctx->h_scalar, ctx->d_scalar, 12 * sizeof(u32),
                                   hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            const u32 nl = ctx->h_scalar[10];
            if (ctx->h_scalar[1] || !nl) return RBSC_OK;  // overflow: the redo handles it
            ++ctx->long_run_rounds;
            const int r3 = fix_long(keys, P, nl, psig);
            if (r3) return r3;
            fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar + 10, 1, 0);
            return RBSC_OK;
        };
        // Rank input of this round (sorted keys and positions) and the
        // positions' output (null in node order); kept for a second rank pass
        // after long runs are fixed.
        u64 *rk_keys = nullptr, *rk_P = nullptr, *rk_Pout = nullptr;
        bool rk_packed = false;  // rk_keys are packed keys (node order)
        auto launch_rank = [&]() {
            const u64 tiles = (n_rank + RANK_TILE - 1) / RANK_TILE;
            const u32 ep = lb_next_epoch(&ctx->sort, s);
            prof_begin(ctx, 3);
            if (rk_packed) {
                if (confirm_ok) fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar + 15, 1, 0);
                rank_packed_kernel<<<(unsigned)tiles, 256, 0, s>>>(
                    rk_keys, w.sig, reinterpret_cast<const u32*>(w.hc),
                    reinterpret_cast<u32*>(w.hc2), n_rank, lo, rank_offset, w.rank_status,
                    ctx->sort.ticket, ctx->sort.ticket_total, ep, ctx->d_scalar, ctx->d_scalar + 3,
                    ctx->d_err, confirm_ok ? members : nullptr, ctx->d_scalar + 15);
            } else
                rank_kernel<<<(unsigned)tiles, 256, 0, s>>>(
                    rk_keys, rk_P, rk_Pout, w.hc2, n_rank, rank_offset, (u64)it + 1, c32,
                    w.rank_status, ctx->sort.ticket, ctx->sort.ticket_total, ep, ctx->d_scalar,
                    ctx->d_scalar + 3, ctx->d_err);
            rep_colors(w.hc2);
            prof_end(ctx);
            ctx->sort.ticket_total += tiles;
        };
        auto sort_and_rank = [&](bool full, bool gather) -> int {
            const int bit0 = full ? 0 : (int)lo;
            // the fused histogram covers at most SORT_HD digits
            const bool fused = gather && (64 - bit0) / 8 <= SORT_HD;
            if (gather && !fused) {
                prof_begin(ctx, 1);
                gather_keys_kernel<<<grid_for(n_rank), BLK, 0, s>>>(w.P[pc], w.sig, w.keys[0],
                                                                    n_rank);
                prof_end(ctx);
            } else if (fused) {
                prof_begin(ctx, 1);
                const unsigned gb = (unsigned)std::min<u64>((n_rank + 4095) / 4096, GK_BLOCKS);
                gather_keys_hist_kernel<<<gb, 256, 0, s>>>(w.P[pc], w.sig, w.keys[0], n_rank, bit0,
                                                           (64 - bit0) / 8, ctx->sort.hist);
                prof_end(ctx);
            }
            u64* kk[2] = {w.keys[0], w.keys[1]};
            u64* pp[2] = {w.P[pc], w.P[pc ^ 1]};
            prof_begin(ctx, 2);
            int r2 = radix_sort_u64(&ctx->sort, kk, pp, n_rank, bit0, 64, &which, s, fused);
            if (r2) return r2;
            if (!full) fix_ties(kk[which], pp[which], nullptr);
            else fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar + 8, 3, 0);
            if (!full && long_before_rank) {
                r2 = fix_long_first(kk[which], pp[which], nullptr);
                if (r2) return r2;
            }
            prof_end(ctx);
            sorted_pc = (pp[which] == w.P[0]) ? 0 : 1;
            rk_keys = kk[which];
            rk_P = pp[which];
            rk_Pout = pp[which ^ 1];
            rk_packed = false;
            launch_rank();
            return RBSC_OK;
        };
        // Node order (no initial colours, raw colours): the sort reads the
        // signatures and colours by node id, so no gather; inside equal
        // signatures the order is by node, which is (signature, colour) order
        // as long as equal signatures have equal colours.  From a uniform
        // start they do: equal signatures mean equal neighbour-colour
        // multisets in this and, the colours refining, every earlier round,
        // hence equal colours -- unless two different multisets hash to the
        // same 64-bit sum.  The rank pass flags any pair out of colour order
        // and the round's node side is then redone from positions in colour
        // order.
        //
        // Packed (rounds whose previous block count is at least half the
        // ranked nodes, so few neighbours share top bits): the sort moves 8
        // bytes per node, the signature's top bits with the node below them,
        // and the tie fix-up and rank read the low bits and colours by node
        // where top bits are equal.
        auto sort_and_rank_nodes = [&](bool full) -> int {
            u64* kk[2] = {w.keys[0], w.keys[1]};
            u64* pp[2] = {w.P[0], w.P[1]};
            prof_begin(ctx, 2);
            if (!full && packed_round) {
                int r2 = radix_sort_nodes_packed(&ctx->sort, w.sig, kk, n_rank, (int)lo, 64, s);
                if (r2) return r2;
                fix_ties(kk[0], nullptr, w.sig);
                prof_end(ctx);
                rk_keys = kk[0];
                rk_P = nullptr;
                rk_Pout = nullptr;
                rk_packed = true;
                launch_rank();
                return RBSC_OK;
            }
            int r2 = radix_sort_nodes(&ctx->sort, w.sig, reinterpret_cast<const u32*>(w.hc), kk,
                                      pp, n_rank, full ? 0 : (int)lo, 64, s);
            if (r2) return r2;
            if (!full) fix_ties(kk[0], pp[0], nullptr);
            else fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar + 8, 3, 0);
            if (!full && long_before_rank) {
                r2 = fix_long_first(kk[0], pp[0], nullptr);
                if (r2) return r2;
            }
            prof_end(ctx);
            rk_keys = kk[0];
            rk_P = pp[0];
            rk_Pout = nullptr;
            rk_packed = false;
            launch_rank();
            return RBSC_OK;
        };
        const bool full_sort = (prm.flags & RBSC_FLAG_FULLSORT) != 0;
        // Node order with one old colour (see sig_insert_kernel): ranks from
        // the distinct signatures through a hash table.
        auto rank_one_class = [&]() -> int {
            const u64 T = 2 * n_rank;
            u32* owner = reinterpret_cast<u32*>(w.P[0]);    // T slots
            u32* num = reinterpret_cast<u32*>(w.P[1]);      // T slots
            u32* slot_of = reinterpret_cast<u32*>(w.hc);    // the old colours, all equal
            prof_begin(ctx, 2);
            fill_u32_kernel<<<grid_for(T), BLK, 0, s>>>(owner, T, 0);
            fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar + 13, 1, 0);
            sig_insert_kernel<<<(unsigned)((n_rank + INS_BLK - 1) / INS_BLK), INS_BLK, 0, s>>>(
                w.sig, n_rank, owner, T, slot_of, ctx->d_err);
            sig_list_kernel<<<(unsigned)((T + LIST_TILE - 1) / LIST_TILE), 256, 0, s>>>(
                owner, T, w.sig, ctx->d_scalar + 13, w.keys[0]);
            HIP_TRY(hipMemcpyAsync(ctx->h_scalar + 13, ctx->d_scalar + 13, sizeof(u32),
                                   hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            const u64 D = ctx->h_scalar[13];
            u64* kk[2] = {w.keys[0], w.keys[1]};
            int wh = 0;
            int r2 = radix_sort_u64(&ctx->sort, kk, nullptr, D, 0, 64, &wh, s);
            if (r2) return r2;
            sig_number_kernel<<<grid_for(D), BLK, 0, s>>>(kk[wh], D, owner, T, w.sig, num,
                                                          ctx->d_err);
            prof_end(ctx);
            prof_begin(ctx, 3);
            sig_colour_kernel<<<grid_for(n_rank), BLK, 0, s>>>(slot_of, num, n_rank, rank_offset,
                                                               reinterpret_cast<u32*>(w.hc2));
            rep_colors(w.hc2);
            fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar + 1, 11, 0);  // no overflow, disorder, runs
            fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar, 1, (u32)(rank_offset + D));
            prof_end(ctx);
            return RBSC_OK;
        };
        const bool one_class = node_order && K == 1 && !full_sort &&
                               !(prm.flags & RBSC_FLAG_SORTRANK) &&
                               n_rank < (1ull << 30);
        if (n_rank) {
            rc = one_class ? rank_one_class()
               : node_order ? sort_and_rank_nodes(full_sort) : sort_and_rank(full_sort, true);
            if (rc) return rc;
        }
        HIP_TRY(hipGetLastError());
        auto read_scalars = [&]() -> int {
            HIP_TRY(hipMemcpyAsync(ctx->h_scalar, ctx->d_scalar, 16 * sizeof(u32),
                                   hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(ctx->h_scalar + 4, ctx->d_err, sizeof(u32),
                                   hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            return RBSC_OK;
        };
        if ((rc = read_scalars())) return rc;
        if (ctx->h_scalar[4]) return check_kernel_errors(ctx);  // K of a failed round is void
        bool redone = false;
        if (n_rank && !ctx->h_scalar[1] && ctx->h_scalar[10]) {
            // Runs too long for the in-kernel sorts: sort them in a compact
            // array and rank again (the rank pass is a pure function of the
            // sorted arrays).
            ++ctx->long_run_rounds;
            rc = fix_long(rk_keys, rk_P, ctx->h_scalar[10], rk_packed ? w.sig : nullptr);
            if (rc) return rc;
            fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar + 3, 1, 0);
            launch_rank();
            if ((rc = read_scalars())) return rc;
            if (ctx->h_scalar[4]) return check_kernel_errors(ctx);
        }
        if (n_rank && node_order && (ctx->h_scalar[1] || ctx->h_scalar[3])) {
            // Run list overflow or pairs out of colour order: redo the node
            // side from positions sorted by the (unchanged) old colours, with
            // the stable full 64-bit sort by signature.
            ++ctx->full_sort_redos;
            redone = true;
            fill_u32<<<1, 32, 0, s>>>(ctx->d_scalar + 1, 3, 0);
            colour_keys_kernel<<<grid_for(n_rank), BLK, 0, s>>>(
                reinterpret_cast<const u32*>(w.hc), n_rank, w.keys[0], w.P[0]);
            u64* kk[2] = {w.keys[0], w.keys[1]};
            u64* pp[2] = {w.P[0], w.P[1]};
            int wc = 0;
            rc = radix_sort_u64(&ctx->sort, kk, pp, n_rank, 0, 32, &wc, s);
            if (rc) return rc;
            pc = wc;
            rc = sort_and_rank(true, true);
            if (rc) return rc;
            if ((rc = read_scalars())) return rc;
            if (ctx->h_scalar[4]) return check_kernel_errors(ctx);
            if (ctx->h_scalar[1] || ctx->h_scalar[3]) {
                rbsc_set_error("node side redo left pairs out of order");
                return RBSC_ERR_KERNEL;
            }
        } else if (n_rank && ctx->h_scalar[1]) {
            // A run list overflowed: sort this round's (signature, P) again
            // on all 64 bits.  Inside equal signatures the arrays still hold
            // colour order, so the full stable sort gives (signature, colour)
            // order.
            ++ctx->full_sort_redos;
            redone = true;
            fill_u32<<<1, 1, 0, s>>>(ctx->d_scalar + 1, 1, 0);
            pc = sorted_pc;
            rc = sort_and_rank(true, false);
            if (rc) return rc;
            if ((rc = read_scalars())) return rc;
            if (ctx->h_scalar[4]) return check_kernel_errors(ctx);
        }
        // the member list stands if this round's ranks came from the packed
        // rank pass (the redo paths below it sort through keys[1])
        have_members = confirm_ok && rk_packed && !redone && n_rank;
        members_n = have_members ? ctx->h_scalar[15] : 0;
        if (n_rank) pc = sorted_pc ^ 1;
        std::swap(w.hc, w.hc2);  // next round reads the colours ranked here
        Knew = n_rank ? ctx->h_scalar[0] : (u32)rank_offset;
        ++it;
        // K grows every round that is not the last and cannot pass N.
        if (it > N + 2) {
            rbsc_set_error("refinement passed %llu rounds without a fixpoint",
                           (unsigned long long)it);
            return RBSC_ERR_KERNEL;
        }
        if (prm.max_rounds && it >= prm.max_rounds) break;
    } while (K != Knew);
    HIP_TRY(hipEventRecord(ctx->ev1, s));
    if (streaming) HIP_TRY(hipStreamSynchronize(ctx->copy_stream));
    rc = check_kernel_errors(ctx);
    if (rc) return rc;
    rc = prof_collect(ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    w.pc = pc;
    w.n_rank = n_rank;
    w.follow = follow;
    w.node_order = node_order;
    out->K = Knew;
    out->rounds = it;
    out->device_ms = ms;
    return RBSC_OK;
}

void device_colors(rbsc_ctx* ctx, rbsc_dgraph* g, u32* d_int) {
    const u64 N = g->N;
    Workspace& w = ctx->ws;
    hipStream_t s = ctx->stream;
    if (w.n_rank < N) fill_u32_kernel<<<grid_for(N), BLK, 0, s>>>(d_int, N, 0);
    if (w.n_rank && w.node_order)
        (void)hipMemcpyAsync(d_int, w.hc, sizeof(u32) * w.n_rank, hipMemcpyDeviceToDevice, s);
    else if (w.n_rank)
        colors_kernel<<<grid_for(w.n_rank), BLK, 0, s>>>(w.P[w.pc], d_int, w.n_rank);
    if (w.follow)
        dup_colors_kernel<<<grid_for(g->U - g->R), BLK, 0, s>>>(g->dup_rep, g->R, g->U - g->R,
                                                                d_int, 0);
}

int download_colors(rbsc_ctx* ctx, rbsc_dgraph* g, u32* host) {
    const u64 N = g->N;
    if (N == 0) return RBSC_OK;
    Workspace& w = ctx->ws;
    hipStream_t s = ctx->stream;
    u32* d_int = reinterpret_cast<u32*>(w.keys[0]);
    u32* d_out = reinterpret_cast<u32*>(w.keys[1]);
    device_colors(ctx, g, d_int);
    map_colors_kernel<<<grid_for(N), BLK, 0, s>>>(d_int, g->pi, N, d_out);
    HIP_TRY(hipMemcpyAsync(host, d_out, sizeof(u32) * N, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return RBSC_OK;
}

// Canonical labels (and optionally raw colours) by original id into host
// memory, computed on the device; reuses workspace arrays the refinement no
// longer needs (call it last).  Colours are below N + 1.
int device_partition(rbsc_ctx* ctx, rbsc_dgraph* g, u32* partition_out, u32* colors_out) {
    const u64 N = g->N;
    if (N == 0) return RBSC_OK;
    Workspace& w = ctx->ws;
    hipStream_t s = ctx->stream;
    auto on_host = [&]() -> int {
        std::vector<u32> col;
        u32* hcol = colors_out;
        if (!hcol) {
            col.resize(N);
            hcol = col.data();
        }
        int rc = download_colors(ctx, g, hcol);
        if (rc) return rc;
        if (partition_out) canonicalize_host((u32)N, hcol, partition_out);
        return RBSC_OK;
    };
    if (w.cap_n < N || scan_tmp_words(N) > w.cap_n) return on_host();  // workspace too small
    u32* d_int = reinterpret_cast<u32*>(w.keys[0]);
    u32* d_col = reinterpret_cast<u32*>(w.keys[1]);  // colours by original id
    device_colors(ctx, g, d_int);
    map_colors_kernel<<<grid_for(N), BLK, 0, s>>>(d_int, g->pi, N, d_col);
    if (colors_out)
        HIP_TRY(hipMemcpyAsync(colors_out, d_col, sizeof(u32) * N, hipMemcpyDeviceToHost, s));
    if (partition_out) {
        u32* first = reinterpret_cast<u32*>(w.keys[0]);  // N + 1 entries (d_int is done)
        // labels straight into the caller's page-locked buffer when it is
        // one from rbsc_host_alloc (no staging copy), else through hc2
        u32* direct = static_cast<u32*>(host_device_ptr(partition_out, sizeof(u32) * N));
        u32* d_lab = direct ? direct : reinterpret_cast<u32*>(w.hc2);
        fill_u32_kernel<<<1, 32, 0, s>>>(ctx->d_scalar + 12, 1, 0);
        u64* pairs[2] = {w.P[0], w.P[1]};
        canon_pairs_kernel<<<grid_for(N), BLK, 0, s>>>(d_col, N, pairs[0], ctx->d_scalar + 12);
        int cbits = 8;  // colours are at most N
        while (cbits < 32 && (N >> cbits)) cbits += 8;
        int wp = 0;
        int rc = radix_sort_u64(&ctx->sort, pairs, nullptr, N, 32, 32 + cbits, &wp, s);
        if (rc) return rc;
        canon_first_kernel<<<grid_for(N), BLK, 0, s>>>(pairs[wp], N, first);
        u64* flag = pairs[wp];  // the sorted pairs are done
        canon_flag_kernel<<<grid_for(N), BLK, 0, s>>>(d_col, first, N, flag);
        rc = scan_excl_u64(flag, flag, N, nullptr, w.sig, s);
        if (rc) return rc;
        canon_label_kernel<<<grid_for(N), BLK, 0, s>>>(d_col, first, flag, N, d_lab);
        if (!direct)
            HIP_TRY(hipMemcpyAsync(partition_out, d_lab, sizeof(u32) * N, hipMemcpyDeviceToHost, s));
    }
    u32 bad = 0;
    if (partition_out)
        HIP_TRY(hipMemcpyAsync(&bad, ctx->d_scalar + 12, sizeof(u32), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return bad ? on_host() : RBSC_OK;
}

int device_block_sizes(rbsc_ctx* ctx, rbsc_dgraph* g, u32 K, u32* sizes_out) {
    const u64 N = g->N;
    if (N == 0 || !sizes_out) return RBSC_OK;
    Workspace& w = ctx->ws;
    hipStream_t s = ctx->stream;
    if (w.cap_n < N || scan_tmp_words(N) > w.cap_n || K > N) {
        rbsc_set_error("block sizes: workspace smaller than the graph");
        return RBSC_ERR_INVALID;
    }
    // Scratch only in arrays that hold no colours (device_colors reads hc
    // or P), so the canonical partition can follow.
    u32* d_int = reinterpret_cast<u32*>(w.keys[0]);
    u32* size = reinterpret_cast<u32*>(w.keys[1]);   // K <= N entries
    u32* d_out = size + N;                            // N entries (keys[1] holds 2 cap)
    u64* bys = w.sig;                                 // N entries
    device_colors(ctx, g, d_int);
    fill_u32_kernel<<<grid_for(K), BLK, 0, s>>>(size, K, 0);
    fill_u64<<<grid_for(N), BLK, 0, s>>>(bys, N, 0);
    colour_size_kernel<<<grid_for(N), BLK, 0, s>>>(d_int, N, K, size);
    size_count_kernel<<<grid_for(K), BLK, 0, s>>>(size, K, N, bys);
    int rc = scan_excl_u64(bys, bys, N, nullptr, w.hc2, s);
    if (rc) return rc;
    sizes_desc_kernel<<<grid_for(N), BLK, 0, s>>>(bys, N, K, d_out);
    HIP_TRY(hipMemcpyAsync(sizes_out, d_out, sizeof(u32) * N, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return RBSC_OK;
}

static int refine_resident(rbsc_ctx* ctx, rbsc_dgraph* g, const rbsc_opts* opts,
                           rbsc_result* res, uint32_t* partition_out, uint32_t* colors_out) {
    const u64 N = g->N;
    auto t0 = std::chrono::steady_clock::now();
    RefineParams prm;
    prm.max_rounds = opts ? opts->max_rounds : 0;
    prm.exact_colours = colors_out != nullptr;
    prm.flags = opts ? opts->flags : 0;
    if (opts && opts->max_batch_edges && opts->max_batch_edges < g->E_in)
        prm.slice_lines = opts->max_batch_edges;
    RefineOut o;
    int rc = refine_core(ctx, g, prm, &o);
    if (rc) return rc;
    int32_t verified = 0;
    if (opts && opts->verify && N) {
        u32* d_int = reinterpret_cast<u32*>(ctx->ws.keys[0]);
        device_colors(ctx, g, d_int);
        u64 bad = 0;
        rc = verify_partition(ctx, g, d_int, o.K, &bad);
        if (rc) return rc;
        verified = bad ? -1 : 1;
    }
    // last: reuses workspace arrays
    if (N && (partition_out || colors_out)) {
        rc = device_partition(ctx, g, partition_out, colors_out);
        if (rc) return rc;
    }
    auto t1 = std::chrono::steady_clock::now();
    rbsc_result r{};
    r.passes = 1;
    r.edges_per_round = g->E_in;
    r.nodes = N ? o.K : 0;
    r.done = 1;
    r.rounds = o.rounds;
    r.device_ms = o.device_ms;
    r.time_s = std::chrono::duration<double>(t1 - t0).count();
    r.verified = verified;
    if (res) *res = r;
    return RBSC_OK;
}

// ---------------------------------------------------------------- C ABI
extern "C" int rbsc_ctx_create(int device, uint64_t hbm_budget, rbsc_ctx** out) {
    if (!out) {
        rbsc_set_error("rbsc_ctx_create: out is null");
        return RBSC_ERR_INVALID;
    }
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        rbsc_set_error("no HIP device available");
        return RBSC_ERR_NODEVICE;
    }
    if (device < 0 || device >= ndev) {
        rbsc_set_error("device %d out of range (%d devices)", device, ndev);
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(device));
    {
        const int rc = batch_device_scratch_init(device);
        if (rc) return rc;
    }
    {
        // the device's memory pool (ctx_alloc) keeps freed memory for reuse
        hipMemPool_t pool = nullptr;
        HIP_TRY(hipDeviceGetDefaultMemPool(&pool, device));
        uint64_t keep = ~0ull;
        HIP_TRY(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    }
    rbsc_ctx* c = new rbsc_ctx();
    c->device = device;
    c->budget = hbm_budget;
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&c->d_err, sizeof(u32)));
    HIP_TRY(hipMemset(c->d_err, 0, sizeof(u32)));
    HIP_TRY(hipMalloc(&c->d_scalar, sizeof(u32) * 16));
    HIP_TRY(hipMemset(c->d_scalar, 0, sizeof(u32) * 16));
    HIP_TRY(hipHostMalloc(&c->h_scalar, sizeof(u32) * 16, hipHostMallocDefault));
    HIP_TRY(hipEventCreate(&c->ev0));
    HIP_TRY(hipEventCreate(&c->ev1));
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(hipEventCreateWithFlags(&c->ev_copy[b], hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&c->ev_done[b], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    for (int k = 0; k < 4; ++k)
        HIP_TRY(hipEventCreateWithFlags(&c->ev_chunk[k], hipEventDisableTiming));
    *out = c;
    return RBSC_OK;
}

extern "C" int rbsc_ctx_destroy(rbsc_ctx* c) {
    if (!c) return RBSC_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->copy_stream);
    ws_release(c);
    sort_ws_free(&c->sort, c->stream);
    for (int b = 0; b < 2; ++b) {
        ctx_free(c->sbuf[b], c->stream);
        ctx_free(c->swbuf[b], c->stream);
        c->sbuf[b] = nullptr;
        c->swbuf[b] = nullptr;
    }
    (void)hipStreamSynchronize(c->stream);
    {
        // hand the pool's unused memory back (memory other contexts hold stays)
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, c->device) == hipSuccess) (void)hipMemPoolTrimTo(pool, 0);
    }
    if (c->d_err) (void)hipFree(c->d_err);
    if (c->d_scalar) (void)hipFree(c->d_scalar);
    if (c->h_scalar) (void)hipHostFree(c->h_scalar);
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    for (int b = 0; b < 2; ++b) {
        if (c->ev_copy[b]) (void)hipEventDestroy(c->ev_copy[b]);
        if (c->ev_done[b]) (void)hipEventDestroy(c->ev_done[b]);
    }
    for (hipEvent_t e : c->ev_chunk)
        if (e) (void)hipEventDestroy(e);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    delete c;
    return RBSC_OK;
}

// ------------------------------------------------------ upload: rows on the device
// The lines of a host graph are checked and compressed into rows on the
// device after the copy (the host loops over every line cost seconds at
// 2^31 lines): rows are the nodes with lines, in input order; ids below N,
// offsets monotone, each source's lines contiguous.  A flagged error is
// described by the host check, which finds its first line.
namespace {
constexpr u32 UPE_BAD = 1u;
// CSR input: flag[u] = 1 for nodes with lines; bad offsets set the flag.
__global__ void csr_flag_kernel(const u64* __restrict__ ro, u64 N, u64 E, u64* __restrict__ flag,
                                u32* __restrict__ err) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u == 0 && (ro[0] != 0 || ro[N] != E)) atomicOr(err, UPE_BAD);
    if (u >= N) return;
    const u64 a = ro[u], b = ro[u + 1];
    if (b < a || b > E) atomicOr(err, UPE_BAD);
    flag[u] = b > a ? 1 : 0;
}
__global__ void csr_rows_kernel(const u64* __restrict__ ro, const u64* __restrict__ pos, u64 N,
                                u64* __restrict__ row_ptr, u32* __restrict__ row_node) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= N || ro[u + 1] <= ro[u]) return;
    row_ptr[pos[u]] = ro[u];
    row_node[pos[u]] = (u32)u;
}
// COO input, tiles of LT_LINES lines: group starts per tile (and source ids
// checked), then each start's row slot from the tile's offset and its rank.
__device__ __forceinline__ u32 coo_starts(const u32* __restrict__ src, u64 E, u64 b0,
                                          u32 f[LT_PER]) {
    u32 c = 0;
#pragma unroll
    for (int j = 0; j < LT_PER; ++j) {
        const u64 e = b0 + j;
        f[j] = (e < E && (e == 0 || src[e] != src[e - 1])) ? 1u : 0u;
        c += f[j];
    }
    return c;
}
__global__ __launch_bounds__(LT_THREADS) void coo_count_kernel(const u32* __restrict__ src, u64 E,
                                                               u64 N, u64* __restrict__ tcount,
                                                               u32* __restrict__ err) {
    __shared__ u32 tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    const u64 b0 = (u64)blockIdx.x * LT_LINES + (u64)threadIdx.x * LT_PER;
    u32 f[LT_PER];
    const u32 c = coo_starts(src, E, b0, f);
    bool bad = false;
#pragma unroll
    for (int j = 0; j < LT_PER; ++j)
        if (b0 + j < E && src[b0 + j] >= N) bad = true;
    if (bad) atomicOr(err, UPE_BAD);
    if (c) atomicAdd(&tot, c);
    __syncthreads();
    if (threadIdx.x == 0) tcount[blockIdx.x] = tot;
}
__global__ __launch_bounds__(LT_THREADS) void coo_rows_kernel(const u32* __restrict__ src, u64 E,
                                                              u64 N, const u64* __restrict__ toff,
                                                              u32* __restrict__ seen,
                                                              u64* __restrict__ row_ptr,
                                                              u32* __restrict__ row_node,
                                                              u32* __restrict__ err) {
    __shared__ u32 wsum[LT_THREADS / 64];
    const u64 b0 = (u64)blockIdx.x * LT_LINES + (u64)threadIdx.x * LT_PER;
    u32 f[LT_PER];
    const u32 c = coo_starts(src, E, b0, f);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    u32 x = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const u32 y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    u32 off = 0;
#pragma unroll
    for (int w = 0; w < LT_THREADS / 64; ++w)
        if (w < wave) off += wsum[w];
    u64 p = toff[blockIdx.x] + off + x - c;
#pragma unroll
    for (int j = 0; j < LT_PER; ++j) {
        if (!f[j]) continue;
        const u64 e = b0 + j;
        const u32 sv = src[e];
        row_ptr[p] = e;
        row_node[p] = sv;
        ++p;
        // a source starting two groups: its lines are not contiguous
        if (sv < N && atomicAdd(&seen[sv], 1u) != 0) atomicOr(err, UPE_BAD);
    }
}
__global__ void dst_check_kernel(const u32* __restrict__ dst, u64 E, u64 N, u32* __restrict__ err) {
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E && dst[e] >= N) atomicOr(err, UPE_BAD);
}
__global__ void set_u64_kernel(u64* p, u64 v) { *p = v; }
}  // namespace

// Rows of the uploaded lines into raw (row_ptr, row_node, U); on a bad
// input, RBSC_ERR_INPUT with the host check's message.
static int rows_on_device(rbsc_ctx* ctx, const rbsc_graph* gr, RawGraph* raw) {
    hipStream_t s = ctx->stream;
    const u64 N = gr->n_nodes, E = gr->n_edges;
    u64 *d_ro = nullptr, *flag = nullptr, *tmp = nullptr, *d_tot = nullptr;
    u32 *d_src = nullptr, *seen = nullptr, *d_err = nullptr;
    u64 h_tot = 0;
    u32 h_err = 0;
    int rc = RBSC_OK;
#define U_TRY(expr)                                                                  \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            rbsc_set_error("%s failed: %s", #expr, hipGetErrorString(_e));           \
            rc = _e == hipErrorOutOfMemory ? RBSC_ERR_NOMEM : RBSC_ERR_HIP;          \
            goto done;                                                               \
        }                                                                            \
    } while (0)
    U_TRY(ctx_alloc(&d_err, sizeof(u32), s));
    U_TRY(ctx_alloc(&d_tot, sizeof(u64), s));
    U_TRY(hipMemsetAsync(d_err, 0, sizeof(u32), s));
    if (gr->row_offsets) {
        U_TRY(ctx_alloc(&d_ro, sizeof(u64) * (N + 1), s));
        U_TRY(ctx_alloc(&flag, sizeof(u64) * std::max<u64>(N, 1), s));
        U_TRY(ctx_alloc(&tmp, sizeof(u64) * scan_tmp_words(N + 1), s));
        U_TRY(hipMemcpyAsync(d_ro, gr->row_offsets, sizeof(u64) * (N + 1), hipMemcpyHostToDevice, s));
        csr_flag_kernel<<<grid_for(std::max<u64>(N, 1)), BLK, 0, s>>>(d_ro, N, E, flag, d_err);
        if ((rc = scan_excl_u64(flag, flag, N, d_tot, tmp, s))) goto done;
        U_TRY(hipMemcpyAsync(&h_tot, d_tot, sizeof(u64), hipMemcpyDeviceToHost, s));
        U_TRY(hipMemcpyAsync(&h_err, d_err, sizeof(u32), hipMemcpyDeviceToHost, s));
        U_TRY(hipStreamSynchronize(s));
        if (h_err) goto done;
        raw->U = h_tot;
        U_TRY(ctx_alloc(&raw->row_ptr, sizeof(u64) * (raw->U + 1), s));
        U_TRY(ctx_alloc(&raw->row_node, sizeof(u32) * std::max<u64>(raw->U, 1), s));
        if (N) csr_rows_kernel<<<grid_for(N), BLK, 0, s>>>(d_ro, flag, N, raw->row_ptr, raw->row_node);
    } else {
        const u64 nt = lt_tiles(E);
        U_TRY(ctx_alloc(&d_src, sizeof(u32) * std::max<u64>(E, 1), s));
        U_TRY(ctx_alloc(&flag, sizeof(u64) * (nt + 1), s));
        U_TRY(ctx_alloc(&tmp, sizeof(u64) * scan_tmp_words(nt + 1), s));
        U_TRY(ctx_alloc(&seen, sizeof(u32) * std::max<u64>(N, 1), s));
        if (E) U_TRY(hipMemcpyAsync(d_src, gr->src, sizeof(u32) * E, hipMemcpyHostToDevice, s));
        U_TRY(hipMemsetAsync(seen, 0, sizeof(u32) * std::max<u64>(N, 1), s));
        U_TRY(hipMemsetAsync(flag + nt, 0, sizeof(u64), s));
        if (nt) coo_count_kernel<<<(unsigned)nt, LT_THREADS, 0, s>>>(d_src, E, N, flag, d_err);
        if ((rc = scan_excl_u64(flag, flag, nt + 1, nullptr, tmp, s))) goto done;
        U_TRY(hipMemcpyAsync(&h_tot, flag + nt, sizeof(u64), hipMemcpyDeviceToHost, s));
        U_TRY(hipMemcpyAsync(&h_err, d_err, sizeof(u32), hipMemcpyDeviceToHost, s));
        U_TRY(hipStreamSynchronize(s));
        if (h_err) goto done;
        raw->U = h_tot;
        U_TRY(ctx_alloc(&raw->row_ptr, sizeof(u64) * (raw->U + 1), s));
        U_TRY(ctx_alloc(&raw->row_node, sizeof(u32) * std::max<u64>(raw->U, 1), s));
        if (nt)
            coo_rows_kernel<<<(unsigned)nt, LT_THREADS, 0, s>>>(d_src, E, N, flag, seen,
                                                                raw->row_ptr, raw->row_node, d_err);
    }
    set_u64_kernel<<<1, 1, 0, s>>>(raw->row_ptr + raw->U, E);
    // heads of lines copied in chunks are checked by the build, chunk by chunk
    if (E && !raw->n_chunks) dst_check_kernel<<<grid_for(E), BLK, 0, s>>>(raw->dst, E, N, d_err);
    U_TRY(hipMemcpyAsync(&h_err, d_err, sizeof(u32), hipMemcpyDeviceToHost, s));
    U_TRY(hipGetLastError());
    U_TRY(hipStreamSynchronize(s));
done:
    for (void* p : {(void*)d_ro, (void*)flag, (void*)tmp, (void*)d_tot, (void*)d_src, (void*)seen,
                    (void*)d_err})
        ctx_free(p, s);
    if (!rc && h_err) {
        // the host check names the first bad line
        std::vector<u64> row_ptr;
        std::vector<u32> row_node;
        rc = compress_host(gr, row_ptr, row_node);
        if (!rc) {
            rbsc_set_error("rbsc_graph_upload: malformed lines");
            rc = RBSC_ERR_INPUT;
        }
    }
    return rc;
#undef U_TRY
}

// Line arrays from this size on are copied in chunks (see graph_upload_host).
constexpr u64 CHUNK_MIN_LINES = 1ull << 20;

int graph_upload_host(rbsc_ctx* ctx, const rbsc_graph* gr, bool lines_on_host, rbsc_dgraph** out,
                      bool sample_indeg) {
    if (!ctx || !gr || !out) {
        rbsc_set_error("rbsc_graph_upload: null argument");
        return RBSC_ERR_INVALID;
    }
    *out = nullptr;
    if (gr->n_nodes == 0xFFFFFFFFu) {
        rbsc_set_error("n_nodes must be below 2^32-1 (colour N+1 must fit 32 bits)");
        return RBSC_ERR_INVALID;
    }
    if (gr->n_edges && !gr->dst) {
        rbsc_set_error("rbsc_graph_upload: dst is null");
        return RBSC_ERR_INVALID;
    }
    if (gr->n_edges && !gr->row_offsets && !gr->src) {
        rbsc_set_error("graph needs src or row_offsets");
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = RBSC_OK;
    const u64 N = gr->n_nodes, E = gr->n_edges;
    const u64 line_bytes = lines_on_host ? 0 : 4ull * E + (gr->weight ? 8ull * E : 0);
    // rows are at most the nodes
    const u64 need = 8 * (N + 1) + 8 * N + line_bytes + E / 8 + 48 * N;
    if (ctx->budget && need > ctx->budget) {
        rbsc_set_error("graph needs %llu bytes, budget is %llu", (unsigned long long)need,
                       (unsigned long long)ctx->budget);
        return RBSC_ERR_BUDGET;
    }
    hipStream_t s = ctx->stream;
    RawGraph raw;
    raw.N = N;
    raw.E = E;
    raw.sample_indeg = sample_indeg;
    auto fail = [&](hipError_t e, const char* what) {
        rbsc_set_error("%s failed: %s", what, hipGetErrorString(e));
        if (raw.n_chunks) (void)hipStreamSynchronize(ctx->copy_stream);  // chunks in flight
        raw_release(&raw, s);
        return e == hipErrorOutOfMemory ? RBSC_ERR_NOMEM : RBSC_ERR_HIP;
    };
    hipError_t e;
    if ((e = ctx_alloc(&raw.dst, sizeof(u32) * std::max<u64>(E, 1), s)) != hipSuccess) return fail(e, "hipMalloc");
    if (gr->weight && (e = ctx_alloc(&raw.wt, sizeof(u64) * std::max<u64>(E, 1), s)) != hipSuccess)
        return fail(e, "hipMalloc");
    if (gr->init_color && (e = ctx_alloc(&raw.init, sizeof(u32) * std::max<u64>(N, 1), s)) != hipSuccess)
        return fail(e, "hipMalloc");
    // Large line arrays cross the link in four chunks on the second stream;
    // the build's line kernels take each chunk as it lands (chunk ends on
    // whole line tiles of the row-fingerprint kernel).
    const bool chunked = E >= CHUNK_MIN_LINES && ctx->copy_stream;
    if (chunked) {
        if ((e = ctx_alloc(&raw.dst_err, sizeof(u32), s)) != hipSuccess) return fail(e, "hipMalloc");
        fill_u32<<<1, 32, 0, s>>>(raw.dst_err, 1, 0);
        if ((e = hipEventRecord(ctx->ev_fork, s)) ||
            (e = hipStreamWaitEvent(ctx->copy_stream, ctx->ev_fork, 0)))
            return fail(e, "hipEventRecord");
        raw.n_chunks = 4;
        for (int c = 0; c < 4; ++c) {
            const u64 lo = c ? raw.chunk_end[c - 1] : 0;
            u64 hi = c == 3 ? E : std::min<u64>(E, (E * (c + 1) / 4 + LT_LINES - 1) / LT_LINES * LT_LINES);
            if (hi < lo) hi = lo;
            raw.chunk_end[c] = hi;
            raw.chunk_ev[c] = ctx->ev_chunk[c];
            if ((hi > lo && (e = hipMemcpyAsync(raw.dst + lo, gr->dst + lo, 4 * (hi - lo),
                                                hipMemcpyHostToDevice, ctx->copy_stream))) ||
                (hi > lo && gr->weight &&
                 (e = hipMemcpyAsync(raw.wt + lo, gr->weight + lo, 8 * (hi - lo),
                                     hipMemcpyHostToDevice, ctx->copy_stream))) ||
                (e = hipEventRecord(raw.chunk_ev[c], ctx->copy_stream)))
                return fail(e, "hipMemcpyAsync");
        }
    } else if ((E && (e = hipMemcpyAsync(raw.dst, gr->dst, 4 * E, hipMemcpyHostToDevice, s))) ||
               (E && gr->weight &&
                (e = hipMemcpyAsync(raw.wt, gr->weight, 8 * E, hipMemcpyHostToDevice, s)))) {
        return fail(e, "hipMemcpyAsync");
    }
    if (N && gr->init_color && (e = hipMemcpyAsync(raw.init, gr->init_color, 4 * N, hipMemcpyHostToDevice, s)))
        return fail(e, "hipMemcpyAsync");
    // the lines' memory goes back on this stream only after every chunk landed
    auto settle = [&]() {
        if (chunked) (void)hipStreamWaitEvent(s, raw.chunk_ev[3], 0);
    };
    rc = rows_on_device(ctx, gr, &raw);
    if (rc) {
        settle();
        raw_release(&raw, s);
        return rc;
    }
    rbsc_dgraph* g = new rbsc_dgraph();
    rc = dgraph_build(ctx, &raw, g);
    settle();
    if (rc == RBSC_ERR_INPUT && chunked) {
        // the host check names the first bad line
        std::vector<u64> row_ptr;
        std::vector<u32> row_node;
        if (compress_host(gr, row_ptr, row_node) == RBSC_OK)
            rbsc_set_error("rbsc_graph_upload: malformed lines");
    }
    raw_release(&raw, s);
    if (rc) {
        dgraph_release(g, s);
        delete g;
        return rc;
    }
    if (lines_on_host && E) {
        rc = rbsc_graph_lines_to_host(ctx, g);
        if (rc) {
            dgraph_release(g, s);
            delete g;
            return rc;
        }
    }
    *out = g;
    return RBSC_OK;
}

extern "C" int rbsc_graph_upload(rbsc_ctx* ctx, const rbsc_graph* gr, rbsc_dgraph** out) {
    return graph_upload_host(ctx, gr, false, out);
}

static void drop_bins(rbsc_dgraph* g, hipStream_t s) {
    if (!g->bins) return;
    g->bytes -= g->bins->bytes;
    bins_release(g->bins, s);
    g->bins = nullptr;
}

extern "C" int rbsc_graph_bin(rbsc_ctx* ctx, rbsc_dgraph* g, const rbsc_bin_opts* opts) {
    if (!ctx || !g) {
        rbsc_set_error("rbsc_graph_bin: null argument");
        return RBSC_ERR_INVALID;
    }
    rbsc_bin_opts o = opts ? *opts : rbsc_bin_opts{0, 0, 0, 0, 0, 0, 0};
    if (o.hot_heads == 0) o.hot_heads = 1u << 22;
    if (o.dense_heads == 0) o.dense_heads = 1u << 21;
    if (o.rows_per_bin == 0) o.rows_per_bin = 1u << 13;
    if (o.chunk_slots == 0) o.chunk_slots = 1ull << 19;
    if (o.scatter_lines == 0) o.scatter_lines = (o.flags & RBSC_BIN_SLOT_ORDER) ? 1u << 11 : 1u << 14;
    if (o.rows_per_bin & (o.rows_per_bin - 1)) {
        rbsc_set_error("rbsc_graph_bin: rows_per_bin must be a power of two");
        return RBSC_ERR_INVALID;
    }
    u32 log_s = 0;
    while ((1u << log_s) < o.rows_per_bin) ++log_s;
    HIP_TRY(hipSetDevice(ctx->device));
    if (g->R == 0) {
        drop_bins(g, ctx->stream);
        return RBSC_OK;
    }
    return dgraph_bin(ctx, g, o.hot_heads, o.dense_heads, log_s, o.chunk_slots, o.scatter_lines,
                      (o.flags & RBSC_BIN_SLOT_ORDER) == 0, (o.flags & RBSC_BIN_HASHED) != 0);
}

extern "C" int rbsc_graph_unbin(rbsc_ctx* ctx, rbsc_dgraph* g) {
    if (!g) return RBSC_OK;
    if (ctx) HIP_TRY(hipSetDevice(ctx->device));
    drop_bins(g, ctx ? ctx->stream : nullptr);
    return RBSC_OK;
}

extern "C" int rbsc_graph_bin_info(const rbsc_dgraph* g, uint64_t* info) {
    if (!g || !info) {
        rbsc_set_error("rbsc_graph_bin_info: null argument");
        return RBSC_ERR_INVALID;
    }
    const rbsc_bins* b = g->bins;
    const uint64_t v[10] = {b ? 1u : 0u, b ? b->n_bins : 0, b ? b->n_hot : 0, b ? b->n_cold : 0,
                            b ? b->n_chunks : 0, b ? b->n_zero : 0, b ? b->bytes : 0,
                            b ? ((u64)b->H | ((u64)b->log_s << 32) | ((u64)b->hashed << 40)) : 0,
                            b ? b->D : 0, b ? b->n_warm : 0};
    for (int i = 0; i < 10; ++i) info[i] = v[i];
    return RBSC_OK;
}

extern "C" int rbsc_graph_relayout(rbsc_ctx* ctx, rbsc_dgraph* g, uint32_t n_parts) {
    if (!ctx || !g) {
        rbsc_set_error("rbsc_graph_relayout: null argument");
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    drop_bins(g, ctx->stream);
    return dgraph_relayout(ctx, g, n_parts, 0);
}

extern "C" int rbsc_graph_split_hot(rbsc_ctx* ctx, rbsc_dgraph* g, uint32_t hot_heads) {
    if (!ctx || !g || hot_heads == 0) {
        rbsc_set_error("rbsc_graph_split_hot: null argument or empty hot range");
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    drop_bins(g, ctx->stream);
    return dgraph_relayout(ctx, g, 2, std::min<u32>(hot_heads, g->U ? g->U : 1));
}

extern "C" int rbsc_graph_lines_to_host(rbsc_ctx* ctx, rbsc_dgraph* g) {
    if (!ctx || !g) {
        rbsc_set_error("rbsc_graph_lines_to_host: null argument");
        return RBSC_ERR_INVALID;
    }
    if (g->n_parts > 1) {
        rbsc_set_error("rbsc_graph_lines_to_host: graph has a partitioned line layout");
        return RBSC_ERR_INVALID;
    }
    if (!g->dst) return RBSC_OK;  // already on the host (or no lines)
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    drop_bins(g, s);  // streamed rounds read the plain line array
    const u64 E = g->E;
    HIP_TRY(hipHostMalloc(&g->dst_host, sizeof(u32) * std::max<u64>(E, 1), hipHostMallocDefault));
    HIP_TRY(hipMemcpy(g->dst_host, g->dst, sizeof(u32) * E, hipMemcpyDeviceToHost));
    if (g->wt) {
        HIP_TRY(hipHostMalloc(&g->wt_host, sizeof(u64) * std::max<u64>(E, 1), hipHostMallocDefault));
        HIP_TRY(hipMemcpy(g->wt_host, g->wt, sizeof(u64) * E, hipMemcpyDeviceToHost));
        ctx_free(g->wt, s);
        g->wt = nullptr;
        g->bytes -= sizeof(u64) * E;
    }
    ctx_free(g->dst, s);
    g->dst = nullptr;
    g->bytes -= sizeof(u32) * E;
    return RBSC_OK;
}

extern "C" int rbsc_graph_free(rbsc_ctx* ctx, rbsc_dgraph* g) {
    if (!g) return RBSC_OK;
    if (ctx) (void)hipSetDevice(ctx->device);
    // without a context the free is ordered on the null stream (the
    // context's streams do not block on it; the graph is idle by contract)
    dgraph_release(g, ctx ? ctx->stream : nullptr);
    delete g;
    return RBSC_OK;
}

extern "C" uint64_t rbsc_graph_bytes(const rbsc_dgraph* g) { return g ? g->bytes : 0; }

extern "C" int rbsc_graph_dims(const rbsc_dgraph* g, uint32_t* n_nodes, uint64_t* n_edges) {
    if (!g) {
        rbsc_set_error("rbsc_graph_dims: null graph");
        return RBSC_ERR_INVALID;
    }
    if (n_nodes) *n_nodes = g->N;
    if (n_edges) *n_edges = g->E_in;
    return RBSC_OK;
}

extern "C" uint32_t rbsc_graph_rows(const rbsc_dgraph* g) { return g ? g->R : 0; }
extern "C" uint64_t rbsc_graph_stored_lines(const rbsc_dgraph* g) { return g ? g->E : 0; }

extern "C" int rbsc_ctx_set_dedup(rbsc_ctx* ctx, int on) {
    if (!ctx) {
        rbsc_set_error("rbsc_ctx_set_dedup: null context");
        return RBSC_ERR_INVALID;
    }
    ctx->dedup = on != 0;
    return RBSC_OK;
}

extern "C" int rbsc_graph_download(rbsc_ctx* ctx, const rbsc_dgraph* g, uint64_t* row_offsets,
                                   uint32_t* dst) {
    if (!ctx || !g) {
        rbsc_set_error("rbsc_graph_download: null argument");
        return RBSC_ERR_INVALID;
    }
    if (g->n_parts > 1) {
        rbsc_set_error("rbsc_graph_download: graph has a partitioned line layout");
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    const u64 N = g->N, U = g->U, R = g->R, E = g->E;
    std::vector<u64> rp(R + 1);
    std::vector<u32> pi(N), inv(N), d(E), dup(U - R);
    HIP_TRY(hipMemcpy(rp.data(), g->row_ptr, sizeof(u64) * (R + 1), hipMemcpyDeviceToHost));
    if (U > R)
        HIP_TRY(hipMemcpy(dup.data(), g->dup_rep, sizeof(u32) * (U - R), hipMemcpyDeviceToHost));
    if (N) HIP_TRY(hipMemcpy(pi.data(), g->pi, sizeof(u32) * N, hipMemcpyDeviceToHost));
    if (E) {
        if (g->dst) HIP_TRY(hipMemcpy(d.data(), g->dst, sizeof(u32) * E, hipMemcpyDeviceToHost));
        else memcpy(d.data(), g->dst_host, sizeof(u32) * E);
    }
    for (u64 v = 0; v < N; ++v) inv[pi[v]] = (u32)v;
    // Rows come back in original node order (the input is assumed sorted by
    // source, as the generators produce it): offsets serially, then the
    // lines by node ranges on several threads.
    auto stored = [&](u64 v) -> u64 {
        u64 j = pi[v];
        return j < U && j >= R ? dup[j - R] : j;  // a repeated row: the stored row's lines
    };
    std::vector<u64> ro(N + 1, 0);
    for (u64 v = 0; v < N; ++v) {
        const u64 j = stored(v);
        ro[v + 1] = ro[v] + (pi[v] < U ? rp[j + 1] - rp[j] : 0);
    }
    if (row_offsets) memcpy(row_offsets, ro.data(), sizeof(u64) * (N + 1));
    if (dst && N) {
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                const u64 v0 = N * t / nt, v1 = N * (t + 1) / nt;
                for (u64 v = v0; v < v1; ++v) {
                    if (pi[v] >= U) continue;
                    const u64 j = stored(v);
                    u64 o = ro[v];
                    for (u64 p = rp[j]; p < rp[j + 1]; ++p, ++o) dst[o] = inv[d[p]];
                }
            });
        for (auto& x : th) x.join();
    }
    return RBSC_OK;
}

extern "C" int rbsc_graph_refine(rbsc_ctx* ctx, rbsc_dgraph* g, const rbsc_opts* opts,
                                 rbsc_result* result, uint32_t* partition_out,
                                 uint32_t* colors_out) {
    if (!ctx || !g) {
        rbsc_set_error("rbsc_graph_refine: null argument");
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    return refine_resident(ctx, g, opts, result, partition_out, colors_out);
}

extern "C" int rbsc_refine(rbsc_ctx* ctx, const rbsc_graph* graph, const rbsc_opts* opts,
                           rbsc_result* result, uint32_t* partition_out) {
    if (!ctx || !graph) {
        rbsc_set_error("rbsc_refine: null argument");
        return RBSC_ERR_INVALID;
    }
    const u64 B = opts ? opts->max_batch_edges : 0;
    const bool batched = B && B < graph->n_edges;
    const u32 mode = opts ? opts->mode : RBSC_MODE_STREAM;
    if (mode != RBSC_MODE_STREAM && mode != RBSC_MODE_QUOTIENT &&
        mode != RBSC_MODE_QUOTIENT_AGGRESSIVE) {
        rbsc_set_error("rbsc_refine: unknown mode %u", mode);
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    if (batched && mode != RBSC_MODE_STREAM)
        return quotient_refine(ctx, graph, opts, result, partition_out);
    auto t0 = std::chrono::steady_clock::now();
    rbsc_dgraph* g = nullptr;
    // refined once: sampled in-degrees are enough for the plain layout's order
    int rc = graph_upload_host(ctx, graph, batched, &g, true);
    if (rc) return rc;
    rbsc_result r{};
    rc = refine_resident(ctx, g, opts, &r, partition_out, nullptr);
    rbsc_graph_free(ctx, g);
    if (rc) return rc;
    auto t1 = std::chrono::steady_clock::now();
    r.time_s = std::chrono::duration<double>(t1 - t0).count();
    if (result) *result = r;
    return RBSC_OK;
}

extern "C" int rbsc_ctx_profile(rbsc_ctx* ctx, int enable) {
    if (!ctx) {
        rbsc_set_error("rbsc_ctx_profile: null context");
        return RBSC_ERR_INVALID;
    }
    ctx->prof = enable != 0;
    ctx->prof_used = 0;
    ctx->prof_phase.clear();
    return RBSC_OK;
}

extern "C" int rbsc_ctx_stats(rbsc_ctx* ctx, uint64_t* out, int reset) {
    if (!ctx || !out) {
        rbsc_set_error("rbsc_ctx_stats: null argument");
        return RBSC_ERR_INVALID;
    }
    out[0] = ctx->full_sort_redos;
    out[1] = ctx->long_run_rounds;
    out[2] = ctx->confirmed_rounds;
    if (reset) ctx->full_sort_redos = ctx->long_run_rounds = ctx->confirmed_rounds = 0;
    return RBSC_OK;
}

extern "C" int rbsc_ctx_profile_read(rbsc_ctx* ctx, double* ms, uint64_t* launches, int reset) {
    if (!ctx) {
        rbsc_set_error("rbsc_ctx_profile_read: null context");
        return RBSC_ERR_INVALID;
    }
    for (int i = 0; i < rbsc_ctx::NPHASE; ++i) {
        if (ms) ms[i] = ctx->prof_ms[i];
        if (launches) launches[i] = ctx->prof_n[i];
        if (reset) {
            ctx->prof_ms[i] = 0;
            ctx->prof_n[i] = 0;
        }
    }
    return RBSC_OK;
}
