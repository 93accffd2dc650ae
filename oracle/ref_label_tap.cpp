// TEST INFRASTRUCTURE: label tap for the compiled reference (oracle/_ref only).
//
// The reference exposes its per-vertex labels only through the callback
// get_scc_result(scc_id, vert_count) (src/scc_common.h:25-28, called at
// src/scc_core.cpp:1076).  The reference objects are linked unmodified with
// -Wl,--wrap=_Z14get_scc_resultPii, so that call lands here first: the raw
// scc_id[0..V) is written to $ISPAN_REF_LABELS (int32, overwritten by every
// run) and the reference's own summary printer then runs as usual.
#include <stdio.h>
#include <stdlib.h>

extern "C" void __real__Z14get_scc_resultPii(int* scc_id, int vert_count);

extern "C" void __wrap__Z14get_scc_resultPii(int* scc_id, int vert_count) {
    const char* path = getenv("ISPAN_REF_LABELS");
    if (path && *path) {
        FILE* f = fopen(path, "wb");
        if (f) {
            fwrite(scc_id, sizeof(int), (size_t)vert_count, f);
            fclose(f);
        }
    }
    __real__Z14get_scc_resultPii(scc_id, vert_count);
}
