// scan.h -- device exclusive prefix sum of u64 arrays (reduce-then-scan),
// used when graphs are built on the device.
#pragma once
#include "common.h"

// Scratch words needed by scan_excl_u64 for n items.
u64 scan_tmp_words(u64 n);
// out[i] = sum in[0..i) for i < n; *d_total (device, nullable) = sum of all.
// in and out may alias.
int scan_excl_u64(const u64* in, u64* out, u64 n, u64* d_total, u64* tmp, hipStream_t s);
