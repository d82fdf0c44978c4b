// verify.hip -- exact stability check of a partition, without hashing.
//
// A partition is the coarsest stable one reached by the round function
// exactly when every node has the same multiset of (block of head, summed
// weight) over its lines as every other node of its block (the lumping
// condition, part_ref/MDP_minimize.cc:55-65) and no block mixes initial
// colours.  The hashed rounds can only merge nodes that this condition
// separates (a 64-bit signature collision), never split nodes it joins, so
// a partition that passes is the exact result; one that fails names a
// collision.
//
// Lines are sorted by (row, colour of head); each node's sorted run list is
// compared with that of one representative of its block (the least internal
// id of the block).  Unweighted lines compare position by position (equal
// sorted colour sequences <=> equal multisets); weighted lines are first
// summed over runs of equal (row, colour).
#include "internal.h"
#include "scan.h"
#include <algorithm>
#include <vector>

namespace {

constexpr int VB = 256;
inline unsigned vgrid(u64 n) { return (unsigned)std::max<u64>(1, (n + VB - 1) / VB); }

// Row of line e: the row holding its tile's first line plus the rows that
// end inside the tile before e (row-end bitmap, one bit per line).
__device__ inline u64 line_row(const u32* __restrict__ end_bits, const u32* __restrict__ tile_row,
                               u64 e) {
    const u64 t = e / SIG_TW;
    u64 r = tile_row[t];
    const u64 w0 = (t * SIG_TW) >> 5, we = e >> 5;
    for (u64 w = w0; w < we; ++w) r += (u64)__popc(end_bits[w]);
    r += (u64)__popc(end_bits[we] & ((1u << (e & 31)) - 1u));
    return r;
}

__global__ void vkeys_kernel(u64 E, const u32* __restrict__ dst, const u32* __restrict__ end_bits,
                             const u32* __restrict__ tile_row, const u32* __restrict__ col,
                             u32 cbits, u64* __restrict__ keys) {
    const u64 e = (u64)blockIdx.x * VB + threadIdx.x;
    if (e < E) keys[e] = (line_row(end_bits, tile_row, e) << cbits) | (u64)col[dst[e]];
}

__global__ void vfill_u32(u32* p, u64 n, u32 v) {
    const u64 i = (u64)blockIdx.x * VB + threadIdx.x;
    if (i < n) p[i] = v;
}

// rep[c] = least node of colour c.  The recorded value is read first: the
// nodes of a large class (all nodes without lines share one colour) mostly
// find a smaller id there and skip the atomic instead of queuing on one
// address.
__global__ void vrep_kernel(const u32* __restrict__ col, u64 N, u32* __restrict__ rep) {
    const u64 u = (u64)blockIdx.x * VB + threadIdx.x;
    if (u >= N) return;
    u32* r = &rep[col[u]];
    if ((u32)u < __hip_atomic_load(r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMin(r, (u32)u);
}

// Nodes with and without lines never share a block; nor do initial colours.
__global__ void vnodes_kernel(const u32* __restrict__ col, const u32* __restrict__ rep,
                              const u32* __restrict__ init, u64 N, u64 U,
                              unsigned long long* bad) {
    const u64 u = (u64)blockIdx.x * VB + threadIdx.x;
    if (u >= N) return;
    const u64 r = rep[col[u]];
    bool b = (u < U) != (r < U);
    if (init && init[u] != init[r]) b = true;
    if (b) atomicAdd(bad, 1ULL);
}

// Unweighted: the sorted colour sequence of each row equals its
// representative's, position by position.
__global__ void vlines_kernel(const u64* __restrict__ sk, u64 E, const u64* __restrict__ row_ptr,
                              const u32* __restrict__ col, const u32* __restrict__ rep, u64 U,
                              u32 cbits, unsigned long long* bad) {
    const u64 e = (u64)blockIdx.x * VB + threadIdx.x;
    if (e >= E) return;
    const u64 cmask = (1ULL << cbits) - 1;
    const u64 u = sk[e] >> cbits;
    const u64 j = e - row_ptr[u];
    const u64 r = rep[col[u]];
    if (r >= U) {
        if (j == 0) atomicAdd(bad, 1ULL);
        return;
    }
    const u64 du = row_ptr[u + 1] - row_ptr[u], dr = row_ptr[r + 1] - row_ptr[r];
    if (du != dr) {
        if (j == 0) atomicAdd(bad, 1ULL);
        return;
    }
    if ((sk[e] & cmask) != (sk[row_ptr[r] + j] & cmask)) atomicAdd(bad, 1ULL);
}

__global__ void vheads_kernel(const u64* __restrict__ sk, u64 E, u64* __restrict__ f) {
    const u64 e = (u64)blockIdx.x * VB + threadIdx.x;
    if (e < E) f[e] = (e == 0 || sk[e] != sk[e - 1]) ? 1ULL : 0ULL;
}

// Heads record their position; W is the exclusive prefix of sorted weights.
__global__ void vrun_start_kernel(const u64* __restrict__ sk, const u64* __restrict__ hidx, u64 E,
                                  u64* __restrict__ run_start) {
    const u64 e = (u64)blockIdx.x * VB + threadIdx.x;
    if (e < E && (e == 0 || sk[e] != sk[e - 1])) run_start[hidx[e]] = e;
}

__global__ void vrun_w_kernel(const u64* __restrict__ run_start, const u64* __restrict__ W,
                              const u64* __restrict__ d_wtotal, const u64* __restrict__ d_nruns,
                              u64 E, u64* __restrict__ run_w) {
    const u64 k = (u64)blockIdx.x * VB + threadIdx.x;
    const u64 nr = *d_nruns;
    if (k >= nr) return;
    const u64 a = run_start[k];
    const u64 wa = W[a];
    const u64 wb = k + 1 < nr ? W[run_start[k + 1]] : *d_wtotal;
    (void)E;
    run_w[k] = wb - wa;
}

// First run of each row (a row's first line always starts a run).
__global__ void vfirst_kernel(const u64* __restrict__ row_ptr, const u64* __restrict__ hidx, u64 U,
                              const u64* __restrict__ d_nruns, u64* __restrict__ first) {
    const u64 u = (u64)blockIdx.x * VB + threadIdx.x;
    if (u < U) first[u] = hidx[row_ptr[u]];
    if (u == U) first[U] = *d_nruns;
}

__global__ void vruns_kernel(const u64* __restrict__ sk, const u64* __restrict__ run_start,
                             const u64* __restrict__ run_w, const u64* __restrict__ first,
                             const u64* __restrict__ d_nruns, const u32* __restrict__ col,
                             const u32* __restrict__ rep, u64 U, u32 cbits,
                             unsigned long long* bad) {
    const u64 k = (u64)blockIdx.x * VB + threadIdx.x;
    if (k >= *d_nruns) return;
    const u64 cmask = (1ULL << cbits) - 1;
    const u64 key = sk[run_start[k]];
    const u64 u = key >> cbits;
    const u64 j = k - first[u];
    const u64 r = rep[col[u]];
    if (r >= U) {
        if (j == 0) atomicAdd(bad, 1ULL);
        return;
    }
    const u64 cu = first[u + 1] - first[u], cr = first[r + 1] - first[r];
    if (cu != cr) {
        if (j == 0) atomicAdd(bad, 1ULL);
        return;
    }
    const u64 kr = first[r] + j;
    if ((key & cmask) != (sk[run_start[kr]] & cmask) || run_w[k] != run_w[kr])
        atomicAdd(bad, 1ULL);
}

// Nodes [R, U) repeat a stored row (dup_rep) and have no lines of their
// own: their row must match the row of their block's representative, unless
// it is that very row.  first == nullptr: unweighted (compare sorted colour
// sequences); else compare (colour, summed weight) runs.
__global__ void vdups_kernel(const u64* __restrict__ sk, const u64* __restrict__ row_ptr,
                             const u32* __restrict__ dup_rep, u64 R, u64 U,
                             const u32* __restrict__ col, const u32* __restrict__ rep, u32 cbits,
                             const u64* __restrict__ first, const u64* __restrict__ run_start,
                             const u64* __restrict__ run_w, unsigned long long* bad) {
    const u64 i = (u64)blockIdx.x * VB + threadIdx.x;
    if (i >= U - R) return;
    const u64 d = R + i;
    const u64 r = rep[col[d]];
    if (r >= U) return;  // counted by vnodes_kernel
    const u64 a = dup_rep[i], b = r < R ? r : dup_rep[r - R];
    if (a == b) return;
    const u64 cmask = (1ULL << cbits) - 1;
    if (!first) {
        const u64 n = row_ptr[a + 1] - row_ptr[a];
        if (n != row_ptr[b + 1] - row_ptr[b]) {
            atomicAdd(bad, 1ULL);
            return;
        }
        for (u64 j = 0; j < n; ++j)
            if ((sk[row_ptr[a] + j] & cmask) != (sk[row_ptr[b] + j] & cmask)) {
                atomicAdd(bad, 1ULL);
                return;
            }
    } else {
        const u64 n = first[a + 1] - first[a];
        if (n != first[b + 1] - first[b]) {
            atomicAdd(bad, 1ULL);
            return;
        }
        for (u64 j = 0; j < n; ++j) {
            const u64 ka = first[a] + j, kb = first[b] + j;
            if ((sk[run_start[ka]] & cmask) != (sk[run_start[kb]] & cmask) || run_w[ka] != run_w[kb]) {
                atomicAdd(bad, 1ULL);
                return;
            }
        }
    }
}

__global__ void vpermute_kernel(const u32* __restrict__ host_col, const u32* __restrict__ pi,
                                u64 N, u32* __restrict__ col) {
    const u64 v = (u64)blockIdx.x * VB + threadIdx.x;
    if (v < N) col[pi[v]] = host_col[v];
}

u32 bits_for(u64 x) {  // bits to hold values 0..x (at least 1)
    u32 b = 1;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

}  // namespace

// Counts violations of exact stability of the colouring d_col (internal
// ids, values < K) on g; *bad_out == 0 iff stable.
int verify_partition(rbsc_ctx* ctx, rbsc_dgraph* g, const u32* d_col, u32 K, u64* bad_out) {
    *bad_out = 0;
    const u64 N = g->N, E = g->E, U = g->U, R = g->R;  // R stored rows, U nodes with lines
    if (N == 0) return RBSC_OK;
    if (g->n_parts > 1) {
        rbsc_set_error("exact check needs the default line layout (graph was relaid out)");
        return RBSC_ERR_INVALID;
    }
    hipStream_t s = ctx->stream;
    const bool weighted = g->wt != nullptr || g->wt_host != nullptr;
    const u32 cbits = bits_for(K ? K - 1 : 0);
    const u32 rbits = bits_for(R ? R - 1 : 0);
    if (rbits + cbits > 64) {
        rbsc_set_error("exact check: %u row bits + %u colour bits exceed 64", rbits, cbits);
        return RBSC_ERR_INVALID;
    }
    const int bit_end = (int)(((rbits + cbits) + 7) / 8 * 8);
    std::vector<void*> owned;
    int rc = RBSC_OK;
    auto alloc = [&](void** p, size_t bytes) -> int {
        const hipError_t e = ctx_alloc(p, std::max<size_t>(bytes, 8), s);
        if (e != hipSuccess) {
            rbsc_set_error("exact check: hipMalloc of %zu bytes failed: %s", bytes,
                           hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? RBSC_ERR_NOMEM : RBSC_ERR_HIP;
        }
        owned.push_back(*p);
        return RBSC_OK;
    };
    auto done = [&](int code) {
        (void)hipStreamSynchronize(s);
        for (void* p : owned) ctx_free(p, s);
        return code;
    };
    u32* rep = nullptr;
    unsigned long long* d_bad = nullptr;
    u64* d_tot = nullptr;  // [0] runs, [1] weight total
    if ((rc = alloc((void**)&rep, sizeof(u32) * std::max<u64>(K, 1)))) return done(rc);
    if ((rc = alloc((void**)&d_bad, sizeof(unsigned long long)))) return done(rc);
    if ((rc = alloc((void**)&d_tot, 2 * sizeof(u64)))) return done(rc);
    vfill_u32<<<1, VB, 0, s>>>(reinterpret_cast<u32*>(d_bad), 2, 0);
    vfill_u32<<<vgrid(K), VB, 0, s>>>(rep, K, 0xFFFFFFFFu);
    vrep_kernel<<<vgrid(N), VB, 0, s>>>(d_col, N, rep);
    vnodes_kernel<<<vgrid(N), VB, 0, s>>>(d_col, rep, g->init, N, U, d_bad);
    if (E) {
        // lines on the device (streamed graphs: copied in for the check)
        const u32* dst = g->dst;
        const u64* wt = g->wt;
        if (!dst) {
            u32* t = nullptr;
            if ((rc = alloc((void**)&t, sizeof(u32) * E))) return done(rc);
            if (hipMemcpyAsync(t, g->dst_host, sizeof(u32) * E, hipMemcpyHostToDevice, s) != hipSuccess)
                return done(RBSC_ERR_HIP);
            dst = t;
            if (g->wt_host) {
                u64* tw = nullptr;
                if ((rc = alloc((void**)&tw, sizeof(u64) * E))) return done(rc);
                if (hipMemcpyAsync(tw, g->wt_host, sizeof(u64) * E, hipMemcpyHostToDevice, s) != hipSuccess)
                    return done(RBSC_ERR_HIP);
                wt = tw;
            }
        }
        u64 *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr;
        if ((rc = alloc((void**)&k0, sizeof(u64) * E))) return done(rc);
        if ((rc = alloc((void**)&k1, sizeof(u64) * E))) return done(rc);
        if (weighted) {
            if ((rc = alloc((void**)&v0, sizeof(u64) * E))) return done(rc);
            if ((rc = alloc((void**)&v1, sizeof(u64) * E))) return done(rc);
            if (hipMemcpyAsync(v0, wt, sizeof(u64) * E, hipMemcpyDeviceToDevice, s) != hipSuccess)
                return done(RBSC_ERR_HIP);
        }
        vkeys_kernel<<<vgrid(E), VB, 0, s>>>(E, dst, g->end_bits, g->tile_row, d_col, cbits, k0);
        if ((rc = sort_ws_init(&ctx->sort, E, ctx->d_err, s))) return done(rc);
        u64* kk[2] = {k0, k1};
        u64* vv[2] = {v0, v1};
        int which = 0;
        if ((rc = radix_sort_u64(&ctx->sort, kk, weighted ? vv : nullptr, E, 0, bit_end, &which, s)))
            return done(rc);
        const u64* sk = kk[which];
        if (!weighted) {
            vlines_kernel<<<vgrid(E), VB, 0, s>>>(sk, E, g->row_ptr, d_col, rep, U, cbits, d_bad);
            if (U > R)
                vdups_kernel<<<vgrid(U - R), VB, 0, s>>>(sk, g->row_ptr, g->dup_rep, R, U, d_col, rep,
                                                         cbits, nullptr, nullptr, nullptr, d_bad);
        } else {
            u64* hidx = kk[which ^ 1];  // free after the sort
            u64* W = vv[which];         // sorted weights -> exclusive prefix, in place
            u64* run_start = vv[which ^ 1];
            u64 *run_w = nullptr, *first = nullptr, *tmp = nullptr;
            if ((rc = alloc((void**)&run_w, sizeof(u64) * E))) return done(rc);
            if ((rc = alloc((void**)&first, sizeof(u64) * (R + 1)))) return done(rc);
            if ((rc = alloc((void**)&tmp, sizeof(u64) * scan_tmp_words(E)))) return done(rc);
            vheads_kernel<<<vgrid(E), VB, 0, s>>>(sk, E, hidx);
            if ((rc = scan_excl_u64(hidx, hidx, E, d_tot, tmp, s))) return done(rc);
            if ((rc = scan_excl_u64(W, W, E, d_tot + 1, tmp, s))) return done(rc);
            vrun_start_kernel<<<vgrid(E), VB, 0, s>>>(sk, hidx, E, run_start);
            vrun_w_kernel<<<vgrid(E), VB, 0, s>>>(run_start, W, d_tot + 1, d_tot, E, run_w);
            vfirst_kernel<<<vgrid(R + 1), VB, 0, s>>>(g->row_ptr, hidx, R, d_tot, first);
            vruns_kernel<<<vgrid(E), VB, 0, s>>>(sk, run_start, run_w, first, d_tot, d_col, rep, U,
                                                 cbits, d_bad);
            if (U > R)
                vdups_kernel<<<vgrid(U - R), VB, 0, s>>>(sk, g->row_ptr, g->dup_rep, R, U, d_col, rep,
                                                         cbits, first, run_start, run_w, d_bad);
        }
    }
    if (hipGetLastError() != hipSuccess) return done(RBSC_ERR_HIP);
    u64 bad = 0;
    if (hipMemcpyAsync(&bad, d_bad, sizeof(u64), hipMemcpyDeviceToHost, s) != hipSuccess)
        return done(RBSC_ERR_HIP);
    rc = done(RBSC_OK);
    if (!rc) rc = check_kernel_errors(ctx);
    *bad_out = bad;
    return rc;
}

extern "C" int rbsc_graph_check_partition(rbsc_ctx* ctx, rbsc_dgraph* g, const uint32_t* colors,
                                          uint64_t* bad_out) {
    if (!ctx || !g || !bad_out || (g->N && !colors)) {
        rbsc_set_error("rbsc_graph_check_partition: null argument");
        return RBSC_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    const u64 N = g->N;
    *bad_out = 0;
    if (N == 0) return RBSC_OK;
    hipStream_t s = ctx->stream;
    std::vector<u32> dense(N);
    const u32 K = canonicalize_host((u32)N, colors, dense.data());
    u32 *h = nullptr, *col = nullptr;
    HIP_TRY(ctx_alloc(&h, sizeof(u32) * N, s));
    if (ctx_alloc(&col, sizeof(u32) * N, s) != hipSuccess) {
        ctx_free(h, s);
        rbsc_set_error("rbsc_graph_check_partition: out of device memory");
        return RBSC_ERR_NOMEM;
    }
    int rc = RBSC_OK;
    if (hipMemcpyAsync(h, dense.data(), sizeof(u32) * N, hipMemcpyHostToDevice, ctx->stream) !=
        hipSuccess) {
        rbsc_set_error("rbsc_graph_check_partition: upload failed");
        rc = RBSC_ERR_HIP;
    }
    if (!rc) {
        vpermute_kernel<<<vgrid(N), VB, 0, ctx->stream>>>(h, g->pi, N, col);
        rc = verify_partition(ctx, g, col, K, bad_out);
    }
    (void)hipStreamSynchronize(ctx->stream);
    ctx_free(h, s);
    ctx_free(col, s);
    return rc;
}
