// gen_device.hip -- R-MAT built directly in HBM (same graph as rbsc_gen_rmat),
// so bench-scale graphs (2^31 lines) need no host round trip.
//   draws -> radix sort of (src << 32 | dst) -> per-node row starts ->
//   clone rows take their source's degree -> prefix sum -> compressed rows ->
//   line copy -> signature-kernel tiles.
#include "gen.h"
#include "internal.h"
#include "scan.h"

namespace {
constexpr int BLK = 256;
inline unsigned grid_for(u64 n) { return (unsigned)((n + BLK - 1) / BLK); }

__global__ void rmat_keys_kernel(RmatSpec s, u64* __restrict__ keys, u64 m) {
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < m) keys[e] = rmat_draw(s, e);
}

// rp[x] = first index i with src(keys[i]) >= x, for x in [0, N].
__global__ void row_start_kernel(const u64* __restrict__ keys, u64 m, u64 N, u64* __restrict__ rp) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > m) return;
    const long long a = i == 0 ? -1 : (long long)(keys[i - 1] >> 32);
    const long long b = i == m ? (long long)N : (long long)(keys[i] >> 32);
    for (long long x = a + 1; x <= b; ++x) rp[x] = i;
}

__global__ void clone_deg_kernel(RmatSpec s, const u64* __restrict__ rp, u64 N,
                                 u64* __restrict__ deg, u64* __restrict__ flag) {
    const u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= N) return;
    const u64 y = rmat_source(s, x);
    const u64 d = rp[y + 1] - rp[y];
    deg[x] = d;
    flag[x] = d > 0;
}

__global__ void build_rows_kernel(const u64* __restrict__ ro, const u64* __restrict__ idx,
                                  const u64* __restrict__ deg, u64 N, u64* __restrict__ row_ptr,
                                  u32* __restrict__ row_node) {
    const u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= N || deg[x] == 0) return;
    row_ptr[idx[x]] = ro[x];
    row_node[idx[x]] = (u32)x;
}

__global__ void copy_lines_kernel(RmatSpec s, const u64* __restrict__ rp,
                                  const u64* __restrict__ ro, const u64* __restrict__ keys,
                                  u64 N, u32* __restrict__ dst) {
    const u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= N) return;
    const u64 y = rmat_source(s, x);
    u64 o = ro[x];
    for (u64 i = rp[y]; i < rp[y + 1]; ++i, ++o) dst[o] = (u32)keys[i];
}

__global__ void set_u64(u64* p, u64 v) { *p = v; }
}  // namespace

extern "C" int rbsc_gen_rmat_device(rbsc_ctx* ctx, uint32_t scale, uint64_t n_draws, double a,
                                    double b, double c, double clone_frac, uint64_t seed,
                                    rbsc_dgraph** out) {
    if (!ctx || !out || scale == 0 || scale > 31 || a < 0 || b < 0 || c < 0 ||
        a + b + c > 1.0 || clone_frac < 0 || clone_frac >= 1.0 || n_draws >= (1ULL << 40)) {
        rbsc_set_error("rbsc_gen_rmat_device: invalid arguments");
        return RBSC_ERR_INVALID;
    }
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const RmatSpec spec = rmat_spec(scale, n_draws, a, b, c, clone_frac, seed);
    const u64 N = spec.mask + 1;
    const u64 M = n_draws;
    u64* keys[2] = {nullptr, nullptr};
    u64 *rp = nullptr, *deg = nullptr, *flag = nullptr, *ro = nullptr, *tmp = nullptr,
        *d_tot = nullptr;
    int rc = RBSC_OK;
    rbsc_dgraph* g = nullptr;
    RawGraph raw;
    u64 h_tot[2] = {0, 0};
    auto cleanup = [&]() {
        for (int i = 0; i < 2; ++i)
            ctx_free(keys[i], s);
        ctx_free(rp, s);
        ctx_free(deg, s);
        ctx_free(flag, s);
        ctx_free(ro, s);
        ctx_free(tmp, s);
        ctx_free(d_tot, s);
    };
#define GEN_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            rbsc_set_error("%s failed: %s", #expr, hipGetErrorString(_e));              \
            cleanup();                                                                  \
            raw_release(&raw, s);                                                       \
            return _e == hipErrorOutOfMemory ? RBSC_ERR_NOMEM : RBSC_ERR_HIP;           \
        }                                                                               \
    } while (0)
    GEN_TRY(ctx_alloc(&keys[0], sizeof(u64) * std::max<u64>(M, 1), s));
    GEN_TRY(ctx_alloc(&keys[1], sizeof(u64) * std::max<u64>(M, 1), s));
    GEN_TRY(ctx_alloc(&rp, sizeof(u64) * (N + 1), s));
    GEN_TRY(ctx_alloc(&deg, sizeof(u64) * N, s));
    GEN_TRY(ctx_alloc(&flag, sizeof(u64) * N, s));
    GEN_TRY(ctx_alloc(&ro, sizeof(u64) * (N + 1), s));
    GEN_TRY(ctx_alloc(&tmp, sizeof(u64) * scan_tmp_words(N + 1), s));
    GEN_TRY(ctx_alloc(&d_tot, sizeof(u64) * 2, s));
    if (M) rmat_keys_kernel<<<grid_for(M), BLK, 0, s>>>(spec, keys[0], M);
    rc = sort_ws_init(&ctx->sort, std::max<u64>(M, N), ctx->d_err, s);
    if (rc) { cleanup(); return rc; }
    int which = 0;
    rc = radix_sort_u64(&ctx->sort, keys, nullptr, M, 0, 64, &which, s);
    if (rc) { cleanup(); return rc; }
    const u64* sk = keys[which];
    row_start_kernel<<<grid_for(M + 1), BLK, 0, s>>>(sk, M, N, rp);
    clone_deg_kernel<<<grid_for(N), BLK, 0, s>>>(spec, rp, N, deg, flag);
    rc = scan_excl_u64(deg, ro, N, d_tot, tmp, s);
    if (rc) { cleanup(); return rc; }
    rc = scan_excl_u64(flag, flag, N, d_tot + 1, tmp, s);
    if (rc) { cleanup(); return rc; }
    GEN_TRY(hipMemcpyAsync(h_tot, d_tot, sizeof(u64) * 2, hipMemcpyDeviceToHost, s));
    GEN_TRY(hipStreamSynchronize(s));
    const u64 E = h_tot[0];
    const u64 U = h_tot[1];
    set_u64<<<1, 1, 0, s>>>(ro + N, E);
    raw.N = N;
    raw.E = E;
    raw.U = U;
    GEN_TRY(ctx_alloc(&raw.row_ptr, sizeof(u64) * (U + 1), s));
    GEN_TRY(ctx_alloc(&raw.row_node, sizeof(u32) * std::max<u64>(U, 1), s));
    GEN_TRY(ctx_alloc(&raw.dst, sizeof(u32) * std::max<u64>(E, 1), s));
    build_rows_kernel<<<grid_for(N), BLK, 0, s>>>(ro, flag, deg, N, raw.row_ptr, raw.row_node);
    set_u64<<<1, 1, 0, s>>>(raw.row_ptr + U, E);
    copy_lines_kernel<<<grid_for(N), BLK, 0, s>>>(spec, rp, ro, sk, N, raw.dst);
    GEN_TRY(hipGetLastError());
    GEN_TRY(hipStreamSynchronize(s));
    cleanup();  // the sorted draws are no longer needed
    for (int i = 0; i < 2; ++i) keys[i] = nullptr;
    rp = deg = flag = ro = tmp = d_tot = nullptr;
    g = new rbsc_dgraph();
    rc = dgraph_build(ctx, &raw, g);
    raw_release(&raw, s);
    if (rc) {
        dgraph_release(g, s);
        delete g;
        return rc;
    }
    *out = g;
    return RBSC_OK;
#undef GEN_TRY
}
