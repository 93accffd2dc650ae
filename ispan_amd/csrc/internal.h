// Library-internal entry points shared between translation units (not part
// of the C ABI).
#pragma once
#include <stdint.h>

#include <hip/hip_runtime.h>

#include "ispan_hip.h"

namespace ispan {

// Device bytes a batch of V vertices and E edges needs beyond its graphs
// (union CSR, first arrays, per-vertex solve state, hub items).
int64_t batch_workspace_bytes(int64_t V, int64_t E);

// Solves graphs [0, n), each given by device pointers to its CSR pair (local
// ids), as one union; graph g's local labels go to d_out + v_off[g].  The
// graphs must be valid CSR (they are not checked here).  ctx's bound graph
// or batch is released.
int batch_solve_tables(ispan_ctx* ctx, int32_t n, const int64_t* v_off, const int64_t* e_off,
                       const int32_t* const* fwb, const int32_t* const* fwa, const int32_t* const* bwb,
                       const int32_t* const* bwa, const ispan_params* P, int32_t* d_out, ispan_stats* st);

// Structural check of a CSR pair already on the device (offsets monotone
// from 0 to E, ids in [0, V)), on stream s.  ISPAN_OK or ISPAN_ERR_CSR.
int check_csr_device(int32_t V, int64_t E, const int32_t* fwb, const int32_t* fwa, const int32_t* bwb,
                     const int32_t* bwa, hipStream_t s);

}  // namespace ispan
