// scan.hip -- device exclusive prefix sum of u64 arrays (reduce-then-scan).
#include "scan.h"
#include "../../include/rbsc.h"

namespace {
constexpr int T = 256;
constexpr int I = 8;
constexpr u64 BLK_ITEMS = T * I;

__device__ __forceinline__ u64 block_scan_excl(u64 v, u64* s_wave, u64* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    u64 x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const u64 y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    u64 off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < T / 64; ++w) {
        if (w < wave) off += s_wave[w];
        tot += s_wave[w];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

__global__ __launch_bounds__(T) void reduce_kernel(const u64* __restrict__ in, u64 n,
                                                   u64* __restrict__ bsum) {
    __shared__ u64 s_wave[T / 64];
    const u64 base = (u64)blockIdx.x * BLK_ITEMS + (u64)threadIdx.x * I;
    u64 v = 0;
#pragma unroll
    for (int j = 0; j < I; ++j)
        if (base + j < n) v += in[base + j];
    u64 tot;
    block_scan_excl(v, s_wave, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Single block: exclusive scan of nb block sums in place; total to bsum[nb].
__global__ __launch_bounds__(T) void top_kernel(u64* bsum, u64 nb, u64* d_total) {
    __shared__ u64 s_wave[T / 64];
    u64 carry = 0;
    for (u64 b0 = 0; b0 < nb; b0 += T) {
        const u64 b = b0 + threadIdx.x;
        const u64 v = b < nb ? bsum[b] : 0;
        u64 tot;
        const u64 ex = block_scan_excl(v, s_wave, &tot);
        if (b < nb) bsum[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        bsum[nb] = carry;
        if (d_total) *d_total = carry;
    }
}

__global__ __launch_bounds__(T) void downsweep_kernel(const u64* in, u64* out, u64 n,
                                                      const u64* __restrict__ bsum) {
    __shared__ u64 s_wave[T / 64];
    const u64 base = (u64)blockIdx.x * BLK_ITEMS + (u64)threadIdx.x * I;
    u64 v[I];
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < I; ++j) {
        v[j] = base + j < n ? in[base + j] : 0;
        s += v[j];
    }
    u64 tot;
    u64 run = bsum[blockIdx.x] + block_scan_excl(s, s_wave, &tot);
#pragma unroll
    for (int j = 0; j < I; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
}
}  // namespace

u64 scan_tmp_words(u64 n) { return (n + BLK_ITEMS - 1) / BLK_ITEMS + 1; }

int scan_excl_u64(const u64* in, u64* out, u64 n, u64* d_total, u64* tmp, hipStream_t s) {
    const u64 nb = (n + BLK_ITEMS - 1) / BLK_ITEMS;
    if (nb) reduce_kernel<<<(unsigned)nb, T, 0, s>>>(in, n, tmp);
    top_kernel<<<1, T, 0, s>>>(tmp, nb, d_total);
    if (nb) downsweep_kernel<<<(unsigned)nb, T, 0, s>>>(in, out, n, tmp);
    HIP_TRY(hipGetLastError());
    return RBSC_OK;
}
