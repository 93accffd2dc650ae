// Library-internal entry points shared between translation units (not part
// of the C ABI).
#pragma once
#include <stdint.h>

#include <hip/hip_runtime.h>

#include "ispan_hip.h"

namespace ispan {

// Device bytes of the per-vertex solve state a CSR of V vertices and E
// edges needs (labels, bitmaps, queues, WCC and mice arrays, hub items).
int64_t solve_state_bytes(int64_t V, int64_t E);

// A CSR that holds many graphs side by side with global ids (the server's
// arena), with its first-neighbour arrays.
struct ArenaView {
    const int32_t* fwb;
    const int32_t* fwa;
    const int32_t* bwb;
    const int32_t* bwa;
    const int32_t* fw_first;
    const int32_t* bw_first;
    int32_t V;
    int64_t E;
};

// Solves the graphs of the arena with vertex ranges [vb[g], ve[g]) (n of
// them, pairwise disjoint) in one pass; only their vertices take part.
// Graph g's local labels go to out[g] (device, ve[g] - vb[g] entries).  The
// graphs must be valid CSR (not checked here).  ctx's bound graph or batch
// is released.
int solve_arena(ispan_ctx* ctx, const ArenaView& av, int32_t n, const int64_t* vb, const int64_t* ve,
                int32_t* const* out, const ispan_params* P, ispan_stats* st);

// Structural check of a CSR pair already on the device (offsets monotone
// from 0 to E, ids in [0, V)), on stream s.  ISPAN_OK or ISPAN_ERR_CSR.
int check_csr_device(int32_t V, int64_t E, const int32_t* fwb, const int32_t* fwa, const int32_t* bwb,
                     const int32_t* bwa, hipStream_t s);

}  // namespace ispan
