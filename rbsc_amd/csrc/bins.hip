// bins.hip -- binned signature layout of a resident graph.
//
// The plain signature kernel gathers one colour per line at a random head.
// On a large graph every such gather is its own L2 request, and gathers of
// the many rarely-used ("cold") heads miss to the fabric, moving a 128-byte
// line for 4 useful bytes.  This layout computes the same signatures
//   sig(u) = sum over u's lines (u, v, w) of randomize(c(v) + r) * w  (mod 2^64)
// (rbsc.cu:80-87 set_values + rbsc.cu:366-377 ReduceByKey) in two passes:
//
//   cold_scatter_kernel  walks the cold lines (H <= head < gather limit) in
//                        head order: colours are read in order, and each
//                        line's colour is written into its slot.  The slots
//                        of one bin are in head order too, so every bin's
//                        slot run fills front to back; bins are split into 8
//                        groups, one per XCD, so a slot's cache line is
//                        written from one L2 only.
//   bin_sig_kernel       one workgroup per chunk of a bin of 2^log_s rows:
//                        row sums in LDS, hot slots (head < H) gathered from
//                        the L2-resident hot colours in head order (a wave's
//                        64 gathers fall into few cache lines), cold slots
//                        read as written, lines into heads that are never
//                        gathered folded in per row as dang_h * dsum[row].
//
// Only the order of the additions differs from the plain kernel, and they are
// exact in Z/2^64: signatures are identical.
#include "internal.h"
#include "scan.h"
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
constexpr int SCATTER_V = 8;                          // lines per thread
constexpr u64 SCATTER_WG = (u64)BLK * SCATTER_V;      // lines per workgroup
constexpr int SIG_THREADS = 512;
constexpr int SIG_VB = 8;                             // slots per thread per step

// ---------------------------------------------------------------- build
// Per stored row: lines into hot heads, into cold heads, and the total
// weight of lines into heads that are never gathered.  Entry R is the zero
// that makes the exclusive scans end with the totals.
template <bool HAS_WT>
__global__ void bin_count_kernel(const u64* __restrict__ row_ptr, const u32* __restrict__ dst,
                                 const u64* __restrict__ wt, u64 R, u32 H, u32 GL,
                                 u64* __restrict__ n_hot, u64* __restrict__ n_cold,
                                 u64* __restrict__ dsum) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u > R) return;
    if (u == R) {
        n_hot[R] = 0;
        n_cold[R] = 0;
        return;
    }
    u64 h = 0, c = 0, d = 0;
    for (u64 e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
        const u32 v = dst[e];
        if (v < H) ++h;
        else if (v < GL) ++c;
        else d += HAS_WT ? wt[e] : 1;
    }
    n_hot[u] = h;
    n_cold[u] = c;
    dsum[u] = d;
}

// Line records of one class in row-major order: key (bin, head), value
// (line << 16 | row in bin).
template <bool HOT>
__global__ void bin_fill_kernel(const u64* __restrict__ row_ptr, const u32* __restrict__ dst,
                                u64 R, u32 H, u32 GL, u32 log_s, int hb,
                                const u64* __restrict__ off, u64* __restrict__ keys,
                                u64* __restrict__ vals) {
    const u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= R) return;
    u64 p = off[u];
    const u64 bin = u >> log_s;
    const u64 rib = u & ((1ull << log_s) - 1);
    for (u64 e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
        const u32 v = dst[e];
        const bool take = HOT ? v < H : (v >= H && v < GL);
        if (take) {
            keys[p] = (bin << hb) | v;
            vals[p] = (e << 16) | rib;
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
// (group, head) -> slot records of the scatter pass.
template <bool HAS_WT>
__global__ void place_cold_kernel(const u64* __restrict__ keys, const u64* __restrict__ vals,
                                  u64 n, int gb, const u64* __restrict__ bin_slot,
                                  const u64* __restrict__ bin_hot,
                                  const u64* __restrict__ cold_start,
                                  const u32* __restrict__ grp_of_bin, const u64* __restrict__ wt,
                                  u32* __restrict__ slot_key,
                                  unsigned short* __restrict__ slot_row,
                                  u64* __restrict__ slot_wt, u64* __restrict__ k2,
                                  u64* __restrict__ v2) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 k = keys[i], v = vals[i];
    const u64 b = k >> gb;
    const u64 head = k & ((1ull << gb) - 1);
    const u64 s = bin_slot[b] + bin_hot[b] + (i - cold_start[b]);
    slot_key[s] = 0;
    slot_row[s] = (unsigned short)(v & 0xFFFF);
    if (HAS_WT) slot_wt[s] = wt[v >> 16];
    k2[i] = ((u64)grp_of_bin[b] << gb) | head;
    v2[i] = s;
}

__global__ void csc_fill_kernel(const u64* __restrict__ k2, const u64* __restrict__ v2, u64 n,
                                int gb, u32* __restrict__ csc_head, u32* __restrict__ csc_slot) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    csc_head[i] = (u32)(k2[i] & ((1ull << gb) - 1));
    csc_slot[i] = (u32)v2[i];
}

// ---------------------------------------------------------------- rounds
// Writes the current colour of every cold line's head into the line's slot.
// Workgroup b serves group b mod 8 (workgroups are dealt to the XCDs round
// robin), lines [grp_line[g] + (b / 8) * SCATTER_WG, ...) in head order.
__global__ __launch_bounds__(BLK) void cold_scatter_kernel(const u32* __restrict__ csc_head,
                                                           const u32* __restrict__ csc_slot,
                                                           const u64* __restrict__ grp_line,
                                                           const u32* __restrict__ col,
                                                           u32* __restrict__ slot_key) {
    const u32 g = blockIdx.x & (BIN_GROUPS - 1);
    const u64 lo = grp_line[g] + (u64)(blockIdx.x / BIN_GROUPS) * SCATTER_WG;
    const u64 end = grp_line[g + 1];
    if (lo >= end) return;
    const u64 hi = lo + SCATTER_WG < end ? lo + SCATTER_WG : end;
    u32 hd[SCATTER_V], sl[SCATTER_V];
#pragma unroll
    for (int k = 0; k < SCATTER_V; ++k) {
        const u64 j = lo + (u64)k * BLK + threadIdx.x;
        hd[k] = j < hi ? csc_head[j] : 0u;
        sl[k] = j < hi ? csc_slot[j] : 0u;
    }
#pragma unroll
    for (int k = 0; k < SCATTER_V; ++k) {
        const u64 j = lo + (u64)k * BLK + threadIdx.x;
        if (j < hi) slot_key[sl[k]] = col[hd[k]];
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

template <int LOG_S, bool HAS_WT>
__global__ __launch_bounds__(SIG_THREADS) void bin_sig_kernel(
    const BinChunk* __restrict__ chunks, const u64* __restrict__ bin_slot,
    const u64* __restrict__ bin_hot, const u32* __restrict__ slot_key,
    const unsigned short* __restrict__ slot_row, const u64* __restrict__ slot_wt,
    const u64* __restrict__ dsum, const u32* __restrict__ col, u64 salt, u64 dang_h, u64 R,
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
    for (u64 base = ch.lo; base < ch.hi; base += (u64)SIG_THREADS * SIG_VB) {
        u32 key[SIG_VB], row[SIG_VB];
        u64 wv[SIG_VB];
#pragma unroll
        for (int k = 0; k < SIG_VB; ++k) {
            const u64 s = base + (u64)k * SIG_THREADS + threadIdx.x;
            const bool ok = s < ch.hi;
            key[k] = ok ? slot_key[s] : 0u;
            row[k] = ok ? (u32)slot_row[s] : 0u;
            if (HAS_WT) wv[k] = ok ? slot_wt[s] : 0ull;
        }
#pragma unroll
        for (int k = 0; k < SIG_VB; ++k) {
            const u64 s = base + (u64)k * SIG_THREADS + threadIdx.x;
            if (s < cold0 && s < ch.hi) key[k] = col[key[k]];
        }
#pragma unroll
        for (int k = 0; k < SIG_VB; ++k) {
            const u64 s = base + (u64)k * SIG_THREADS + threadIdx.x;
            if (s < ch.hi) {
                u64 h = rbsc_randomize((u64)key[k] + salt);
                if (HAS_WT) h *= wv[k];
                atomicAdd((unsigned long long*)&sums[row[k]], (unsigned long long)h);
            }
        }
    }
    __syncthreads();
    const bool only = (ch.flags & CHUNK_ONLY) != 0;
    for (u32 i = threadIdx.x; i < nrows; i += SIG_THREADS) {
        if (only) sig[row0 + i] = sums[i];
        else atomicAdd((unsigned long long*)&sig[row0 + i], (unsigned long long)sums[i]);
    }
}

template <int LOG_S>
void launch_bin_sig(hipStream_t s, const rbsc_bins* b, const u32* col, u64 salt, u64 dang_h,
                    u64 R, u64* sig) {
    if (b->slot_wt)
        bin_sig_kernel<LOG_S, true><<<(unsigned)b->n_chunks, SIG_THREADS, 0, s>>>(
            b->chunks, b->bin_slot, b->bin_hot, b->slot_key, b->slot_row, b->slot_wt, b->dsum, col,
            salt, dang_h, R, sig);
    else
        bin_sig_kernel<LOG_S, false><<<(unsigned)b->n_chunks, SIG_THREADS, 0, s>>>(
            b->chunks, b->bin_slot, b->bin_hot, b->slot_key, b->slot_row, nullptr, b->dsum, col,
            salt, dang_h, R, sig);
}
}  // namespace

void bins_scatter(rbsc_ctx* ctx, const rbsc_dgraph* g, const u32* col) {
    const rbsc_bins* b = g->bins;
    if (!b->n_cold) return;
    u64 most = 0;
    for (int k = 0; k < BIN_GROUPS; ++k) most = std::max(most, b->grp_line[k + 1] - b->grp_line[k]);
    const u64 per = (most + SCATTER_WG - 1) / SCATTER_WG;
    cold_scatter_kernel<<<(unsigned)(per * BIN_GROUPS), BLK, 0, ctx->stream>>>(
        b->csc_head, b->csc_slot, b->d_grp_line, col, b->slot_key);
}

void bins_sig(rbsc_ctx* ctx, const rbsc_dgraph* g, const u32* col, u64 salt, u64 dang_h,
              u64* sig) {
    const rbsc_bins* b = g->bins;
    hipStream_t s = ctx->stream;
    if (b->n_zero)
        bin_zero_kernel<<<grid_for(b->n_zero << b->log_s), BLK, 0, s>>>(b->zero_bins, b->n_zero,
                                                                     b->log_s, g->R, sig);
    switch (b->log_s) {
        case 8: launch_bin_sig<8>(s, b, col, salt, dang_h, g->R, sig); break;
        case 9: launch_bin_sig<9>(s, b, col, salt, dang_h, g->R, sig); break;
        case 10: launch_bin_sig<10>(s, b, col, salt, dang_h, g->R, sig); break;
        case 11: launch_bin_sig<11>(s, b, col, salt, dang_h, g->R, sig); break;
        case 12: launch_bin_sig<12>(s, b, col, salt, dang_h, g->R, sig); break;
        case 13: launch_bin_sig<13>(s, b, col, salt, dang_h, g->R, sig); break;
        default: launch_bin_sig<14>(s, b, col, salt, dang_h, g->R, sig); break;
    }
}

void bins_release(rbsc_bins* b) {
    if (!b) return;
    for (void* p : {(void*)b->bin_slot, (void*)b->bin_hot, (void*)b->slot_key, (void*)b->slot_row,
                    (void*)b->slot_wt, (void*)b->dsum, (void*)b->csc_head, (void*)b->csc_slot,
                    (void*)b->d_grp_line, (void*)b->chunks, (void*)b->zero_bins})
        if (p) (void)hipFree(p);
    delete b;
}

int dgraph_bin(rbsc_ctx* ctx, rbsc_dgraph* g, u32 H, u32 log_s, u64 chunk_slots) {
    if (!g->dst || g->n_parts != 1) {
        rbsc_set_error("bin: needs the plain layout with lines resident on the device");
        return RBSC_ERR_INVALID;
    }
    if (log_s < 8 || log_s > 14) {
        rbsc_set_error("bin: rows per bin must be a power of two from 256 to 16384");
        return RBSC_ERR_INVALID;
    }
    hipStream_t s = ctx->stream;
    const u64 R = g->R;
    // Heads at or beyond GL are never gathered: nodes without lines, when no
    // initial colours are given (their colour is 0 every round).
    const u32 GL = g->init ? g->N : g->U;
    H = std::min(H, GL);
    const u64 n_bins = (R + (1ull << log_s) - 1) >> log_s;
    if (chunk_slots == 0) chunk_slots = 1ull << 19;
    rbsc_bins* b = new rbsc_bins();
    b->log_s = log_s;
    b->H = H;
    b->GL = GL;
    b->n_bins = n_bins;
    u64 *n_hot = nullptr, *n_cold = nullptr, *tmp = nullptr, *tot = nullptr, *hot_start = nullptr,
        *cold_start = nullptr, *bin_n = nullptr, *keys[2] = {nullptr, nullptr},
        *vals[2] = {nullptr, nullptr};
    u32* grp_of_bin = nullptr;
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
    B_TRY(hipMalloc(&n_hot, sizeof(u64) * (R + 1)));
    B_TRY(hipMalloc(&n_cold, sizeof(u64) * (R + 1)));
    B_TRY(hipMalloc(&b->dsum, sizeof(u64) * std::max<u64>(R, 1)));
    B_TRY(hipMalloc(&tmp, sizeof(u64) * scan_tmp_words(std::max<u64>(R, n_bins) + 1)));
    B_TRY(hipMalloc(&tot, sizeof(u64) * 2));
    if (g->wt)
        bin_count_kernel<true><<<grid_for(R + 1), BLK, 0, s>>>(g->row_ptr, g->dst, g->wt, R, H, GL,
                                                             n_hot, n_cold, b->dsum);
    else
        bin_count_kernel<false><<<grid_for(R + 1), BLK, 0, s>>>(g->row_ptr, g->dst, nullptr, R, H,
                                                              GL, n_hot, n_cold, b->dsum);
    if ((rc = scan_excl_u64(n_hot, n_hot, R + 1, tot, tmp, s))) goto done;
    if ((rc = scan_excl_u64(n_cold, n_cold, R + 1, tot + 1, tmp, s))) goto done;
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
    B_TRY(hipMalloc(&hot_start, sizeof(u64) * (n_bins + 1)));
    B_TRY(hipMalloc(&cold_start, sizeof(u64) * (n_bins + 1)));
    B_TRY(hipMalloc(&bin_n, sizeof(u64) * (n_bins + 1)));
    B_TRY(hipMalloc(&b->bin_hot, sizeof(u64) * (n_bins + 1)));
    B_TRY(hipMalloc(&b->bin_slot, sizeof(u64) * (n_bins + 1)));
    bin_sizes_kernel<<<grid_for(n_bins + 1), BLK, 0, s>>>(n_hot, n_cold, R, log_s, n_bins, hot_start,
                                                          cold_start, bin_n, b->bin_hot);
    h_bin_n.resize(n_bins + 1);
    h_bin_hot.resize(n_bins + 1);
    B_TRY(hipMemcpyAsync(h_bin_n.data(), bin_n, sizeof(u64) * (n_bins + 1), hipMemcpyDeviceToHost, s));
    B_TRY(hipMemcpyAsync(h_bin_hot.data(), b->bin_hot, sizeof(u64) * n_bins, hipMemcpyDeviceToHost, s));
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
        std::fill(b->grp_line, b->grp_line + BIN_GROUPS + 1, 0);
        for (u64 k = 0; k < n_bins; ++k) b->grp_line[h_grp[k] + 1] += h_bin_n[k] - h_bin_hot[k];
        for (int k = 0; k < BIN_GROUPS; ++k) b->grp_line[k + 1] += b->grp_line[k];
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
    B_TRY(hipMalloc(&b->chunks, sizeof(BinChunk) * std::max<u64>(b->n_chunks, 1)));
    B_TRY(hipMemcpyAsync(b->chunks, h_chunks.data(), sizeof(BinChunk) * b->n_chunks,
                         hipMemcpyHostToDevice, s));
    if (b->n_zero) {
        B_TRY(hipMalloc(&b->zero_bins, sizeof(u32) * b->n_zero));
        B_TRY(hipMemcpyAsync(b->zero_bins, h_zero.data(), sizeof(u32) * b->n_zero,
                             hipMemcpyHostToDevice, s));
    }
    B_TRY(hipMalloc(&b->d_grp_line, sizeof(u64) * (BIN_GROUPS + 1)));
    B_TRY(hipMemcpyAsync(b->d_grp_line, b->grp_line, sizeof(u64) * (BIN_GROUPS + 1),
                         hipMemcpyHostToDevice, s));
    // slot arrays
    B_TRY(hipMalloc(&b->slot_key, sizeof(u32) * std::max<u64>(b->n_slots, 1)));
    B_TRY(hipMalloc(&b->slot_row, sizeof(unsigned short) * std::max<u64>(b->n_slots, 1)));
    if (g->wt) B_TRY(hipMalloc(&b->slot_wt, sizeof(u64) * std::max<u64>(b->n_slots, 1)));
    if ((rc = sort_ws_init(&ctx->sort, std::max(b->n_hot, b->n_cold), ctx->d_err, s))) goto done;
    // hot slots: records sorted by (bin, head)
    if (b->n_hot) {
        for (int i = 0; i < 2; ++i) {
            B_TRY(hipMalloc(&keys[i], sizeof(u64) * b->n_hot));
            B_TRY(hipMalloc(&vals[i], sizeof(u64) * b->n_hot));
        }
        bin_fill_kernel<true><<<grid_for(R), BLK, 0, s>>>(g->row_ptr, g->dst, R, H, GL, log_s, hb,
                                                          n_hot, keys[0], vals[0]);
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
        B_TRY(hipStreamSynchronize(s));
        for (int i = 0; i < 2; ++i) {
            (void)hipFree(keys[i]);
            (void)hipFree(vals[i]);
            keys[i] = vals[i] = nullptr;
        }
    }
    // cold slots and the scatter records
    if (b->n_cold) {
        for (int i = 0; i < 2; ++i) {
            B_TRY(hipMalloc(&keys[i], sizeof(u64) * b->n_cold));
            B_TRY(hipMalloc(&vals[i], sizeof(u64) * b->n_cold));
        }
        B_TRY(hipMalloc(&grp_of_bin, sizeof(u32) * (n_bins + 1)));
        B_TRY(hipMemcpyAsync(grp_of_bin, h_grp.data(), sizeof(u32) * (n_bins + 1),
                             hipMemcpyHostToDevice, s));
        bin_fill_kernel<false><<<grid_for(R), BLK, 0, s>>>(g->row_ptr, g->dst, R, H, GL, log_s, gb,
                                                           n_cold, keys[0], vals[0]);
        int w = 0;
        int bits = (gb + bb + 7) / 8 * 8;
        if ((rc = radix_sort_u64(&ctx->sort, keys, vals, b->n_cold, 0, bits, &w, s))) goto done;
        if (g->wt)
            place_cold_kernel<true><<<grid_for(b->n_cold), BLK, 0, s>>>(
                keys[w], vals[w], b->n_cold, gb, b->bin_slot, b->bin_hot, cold_start, grp_of_bin,
                g->wt, b->slot_key, b->slot_row, b->slot_wt, keys[w ^ 1], vals[w ^ 1]);
        else
            place_cold_kernel<false><<<grid_for(b->n_cold), BLK, 0, s>>>(
                keys[w], vals[w], b->n_cold, gb, b->bin_slot, b->bin_hot, cold_start, grp_of_bin,
                nullptr, b->slot_key, b->slot_row, nullptr, keys[w ^ 1], vals[w ^ 1]);
        u64* k2[2] = {keys[w ^ 1], keys[w]};
        u64* v2[2] = {vals[w ^ 1], vals[w]};
        int w2 = 0;
        bits = (gb + 3 + 7) / 8 * 8;
        if ((rc = radix_sort_u64(&ctx->sort, k2, v2, b->n_cold, 0, bits, &w2, s))) goto done;
        B_TRY(hipMalloc(&b->csc_head, sizeof(u32) * b->n_cold));
        B_TRY(hipMalloc(&b->csc_slot, sizeof(u32) * b->n_cold));
        csc_fill_kernel<<<grid_for(b->n_cold), BLK, 0, s>>>(k2[w2], v2[w2], b->n_cold, gb,
                                                           b->csc_head, b->csc_slot);
    }
    B_TRY(hipGetLastError());
    B_TRY(hipStreamSynchronize(s));
    b->bytes = 16 * (n_bins + 1) + 6 * b->n_slots + (g->wt ? 8 * b->n_slots : 0) + 8 * R +
               8 * b->n_cold + sizeof(BinChunk) * b->n_chunks + 4 * b->n_zero;
    if (g->bins) {
        g->bytes -= g->bins->bytes;
        bins_release(g->bins);
    }
    g->bins = b;
    g->bytes += b->bytes;
    b = nullptr;
done:
    for (u64* p : {n_hot, n_cold, tmp, tot, hot_start, cold_start, bin_n, keys[0], keys[1],
                   vals[0], vals[1]})
        if (p) (void)hipFree(p);
    if (grp_of_bin) (void)hipFree(grp_of_bin);
    if (b) bins_release(b);
    return rc;
#undef B_TRY
}
