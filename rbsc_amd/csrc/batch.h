// batch.h -- device layout of a batch of small graphs (batch.hip).
#pragma once
#include "common.h"

// Largest graph (nodes) refined inside one workgroup's LDS.
constexpr u32 BATCH_NMAX = 5120;
// Graphs of at most BATCH_NMID nodes run in a second launch whose LDS layout
// is sized for them (half of BATCH_NMAX), two workgroups per CU; graphs of
// at most BATCH_NLOW nodes in a third, four per CU; graphs of at most
// BATCH_NTINY nodes in a fourth, eight per CU.
constexpr u32 BATCH_NMID = 2560;
constexpr u32 BATCH_NLOW = 1280;
constexpr u32 BATCH_NTINY = 640;

struct BatchDesc {
    u64 node_base;  // first node of the graph in the batch's node numbering
    u64 line_base;  // first line in the small-graph line arrays
    u64 rp_base;    // first local row offset (n + 1 entries)
    u32 n;          // nodes
    u32 pad;
};

struct BatchDevice {
    BatchDesc* desc = nullptr;      // per small graph
    u32* order = nullptr;           // small graphs, largest first
    u32 n_small = 0;
    u32 n_hi = 0;                   // leading graphs of order with more than BATCH_NMID nodes
    u32 n_mid = 0;                  // ... with more than BATCH_NLOW nodes
    u32 n_low = 0;                  // ... with more than BATCH_NTINY nodes
    int cus = 256;                  // compute units of the device
    u32* rp = nullptr;              // local row offsets
    unsigned short* dst = nullptr;  // local heads
    u64* wt = nullptr;              // or null
    u32* init = nullptr;            // per node (batch numbering) or null
    int rb[4] = {8, 8, 8, 8};       // LDS radix digit bits per tier (8 or 4; autotuned)
};

struct rbsc_ctx;
// split_lo: the signatures are sorted by their bits [split_lo, 64) and runs
// colliding there ordered afterwards (32; 48 is a verification hook).
// tiers: bit t launches size tier t (the autotuner times them one by one).
// d_sizes (nullable): per batch node, each graph's block sizes in
// descending order, then zeros.
int batch_launch(rbsc_ctx* ctx, const BatchDevice& b, u32 max_rounds, u32* d_labels, u32* d_colors,
                 u32* d_sizes, u32* d_blocks, u32* d_rounds, u32* d_queue, int threads, int n_wg,
                 int split_lo = 32, u32 tiers = 0xFu);
// Looks up, once per device (that device current), the device address of
// the fixed scratch array of rbsc_batch_refine_device.
int batch_device_scratch_init(int device);
