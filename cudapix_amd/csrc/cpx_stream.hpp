// cpx_stream.hpp -- shared layout of the streaming (continuous-batching)
// launch: a persistent kernel that keeps taking work while the host keeps
// publishing new cohorts of queries into it.
//
// A cohort is one planned batch of admitted queries.  Its entries occupy a
// fresh range of every arena array (ranges are handed out in increasing
// order and 256-byte aligned; none is reused while the launch runs).  The
// host writes its plan into a pinned image of the arena, publishes the
// cohort's record and advances the published ends of two FIFOs: copy units
// (the kernel moves the plan from the pinned image into the arena -- no
// host-side copy may be queued while the persistent kernel holds every CU,
// since the runtime may run small copies as kernels) and main tiles (segment,
// easy-leaf, P2-lookup, ordinary and trivial tiles -- the tiles a batch
// launch runs, rebased to the cohort's entries).  Blocks claim main tiles by
// compare-and-swap below that end.  When a cohort's last main tile ends, its
// carry units become claimable; after them its combine units (one per
// query), which write the results to pinned host memory; the block that ends
// the last one resets the slot and signals the host.
//
// Coherence without cache flushes: a plan is read only at addresses no block
// touched earlier in the launch, after copy units stored it written-through;
// rows one block writes and another reads (carry rows, partial sums) are
// stored written-through too (st_out); counters and heads are device-scope
// atomics; the host's published state is pinned memory the kernel reads with
// system-scope loads; results and reports go to pinned host memory with
// system-scope stores.  Order is kept by waiting for outstanding memory
// operations before a signal and after seeing one.
#pragma once
#include <stdint.h>
#include "cpx_internal.hpp"

#define CPX_STREAM_SLOTS 8

// Unit counts per cohort fit 24 bits, virtual indices 40 bits: a range is
// packed as (start << 24) | count so that one 64-bit load reads it whole.
#define CPX_STREAM_NBITS 24
static inline __host__ __device__ uint64_t stream_pack(uint64_t start, uint64_t n)
{
  return (start << CPX_STREAM_NBITS) | n;
}

// The published record of the cohort in a slot.
struct StreamRec {
  uint64_t seq;          // cohort number (from 1)
  uint64_t copy_range;   // stream_pack(first copy unit, copy units)
  uint64_t main_range;   // stream_pack(first main tile, main tiles)
  uint64_t n_main;       // main tiles
  uint64_t n_fin;        // carry descriptors
  uint64_t n_comb;       // queries (combine units)
  uint64_t q_base, tile_base, fin_base;   // its first query, main tile and carry descriptor
  uint64_t nranges;      // its entries in the plan region: bytes [off, off + len) of each
  uint64_t off[12], len[12];   // (multiples of 4; copy units walk their concatenation)
  uint64_t copy_bytes;   // sum of len
};

// What the host publishes to a streaming launch (pinned host memory, host
// stores in order: a record before the ends that cover it).
struct StreamPub {
  uint64_t copy_end;         // published copy units
  uint64_t main_end;         // published main tiles
  uint64_t cohorts;          // published cohorts
  uint64_t close;            // no more cohorts: exit once they are all done
  const char *image;         // the pinned image of the plan region (device address)
  char *arena;               // the plan region
  StreamRec rec[CPX_STREAM_SLOTS];
};

// Pinned host mailbox: what a streaming launch reports (device stores with
// system scope).
struct StreamMailbox {
  uint64_t done_seq[CPX_STREAM_SLOTS];   // seq of the last cohort completed in the slot
  uint64_t t_first[CPX_STREAM_SLOTS];    // device wall clock: its first main tile began
  uint64_t t_last[CPX_STREAM_SLOTS];     // ... its last main tile ended
  uint64_t main_claimed;     // main tiles claimed so far (backlog hint)
  uint32_t error;            // a copy wait ran out (1), the idle limit hit (2), a unit had no slot (4)
  uint32_t pad1;
};

// Device-side state of one streaming launch (zeroed before the launch, the
// phase heads set to "not ready").
struct StreamCtl {
  uint64_t copy_head;        // next copy unit to claim
  uint64_t main_head;        // next main tile to claim
  uint64_t completed;        // cohorts completed by this launch
  uint64_t error;            // nonzero: every block leaves (mirrored to the mailbox)
  // per slot; reset by the block that completes the slot's cohort
  uint64_t fin_head[CPX_STREAM_SLOTS];   // (seq << 32) | next carry unit; low word ~0: not ready
  uint64_t comb_head[CPX_STREAM_SLOTS];  // (seq << 32) | next combine unit
  uint32_t copied[CPX_STREAM_SLOTS];
  uint32_t main_done[CPX_STREAM_SLOTS];
  uint32_t fin_done[CPX_STREAM_SLOTS];
  uint32_t comb_done[CPX_STREAM_SLOTS];
  uint64_t t_first[CPX_STREAM_SLOTS];
  uint64_t t_last[CPX_STREAM_SLOTS];
};

// Kernel arguments of a streaming launch.
struct StreamArgs {
  BatchDev bd;                       // the arena (every cohort's entries in each array) and
                                     // the shared tables
  const BatchDev *bd_mem;            // the same record in device memory (for out-of-line code)
  StreamCtl *ctl;
  const StreamPub *pub;
  StreamMailbox *mb;
  uint64_t idle_ticks;               // device wall-clock ticks without work before a block leaves
  uint64_t wait_ticks;               // ... a main tile waits for its cohort's copy units at most
  uint32_t nslots;
};

// Bytes of plan one copy unit moves.
#define CPX_STREAM_COPY_BYTES (256u << 10)

// TileDev.group of the ordinary and trivial tiles of a streaming cohort
// (`seg` is the descriptor index); a batch launch runs those terms as
// kernels of their own instead.
#define CPX_TILE_ORD 0xfffffffdu
#define CPX_TILE_TRIV 0xfffffffcu
#define CPX_TILE_NOP 0xfffffffbu     // no work (a cohort whose queries are all table lookups)

// A block that finds no work for this long leaves even if the host never
// closes; a main tile waits at most this long for its cohort's copy units
// (which other blocks claimed before it).
#define CPX_STREAM_IDLE_MS 3000.0
#define CPX_STREAM_WAIT_MS 2000.0

// Arena entries of a cohort start at multiples of this (every element size
// is a multiple of 4 bytes, so ranges start on 256-byte boundaries and no two
// cohorts share a cache line).
#define CPX_STREAM_ALIGN 64

hipError_t launch_stream(const StreamArgs &a, SegShape shape, uint32_t grid, hipStream_t s);
bool stream_shape_supported(SegShape s);
int stream_blocks_per_cu(SegShape shape);
