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

// Number of keys per onesweep tile.
constexpr u64 SORT_TILE = 4096;

int sort_ws_init(SortWorkspace* ws, u64 max_n, u32* err_flag, hipStream_t s);
void sort_ws_free(SortWorkspace* ws);

// Sort (keys[0], vals[0]) by key bits [bit_begin, bit_end) (multiple of 8 wide),
// stably, ping-ponging with (keys[1], vals[1]).  vals may be null (keys only).
// Returns in *which the buffer index (0 or 1) that holds the result.
int radix_sort_u64(SortWorkspace* ws, u64* keys[2], u64* vals[2], u64 n, int bit_begin,
                   int bit_end, int* which, hipStream_t s);

// Epoch/ticket bookkeeping shared with the other look-back kernels.
u32 lb_next_epoch(SortWorkspace* ws, hipStream_t s);
