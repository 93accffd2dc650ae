// Drop-in `ispan` executable: same argv and stdout grammar as the reference
// CLI (src/scc_main.cpp:7-57, per-run block src/scc_core.cpp:1062-1068,
// result block src/load_print.cpp:58, averages src/load_print.cpp:69-90),
// solving on the GPU through libispan_hip.so.
//
//   ispan <fw_beg> <fw_adj> <bw_beg> <bw_adj> <thread_count> <alpha> <beta>
//         <gamma> <theta> <run_times>
//
// thread_count is accepted and printed; the GPU path does not use it.  The
// result block counts SCCs by size from the canonical labels (the reference
// counts by the phase that found them and adds one phantom entry).  With
// ISPAN_LABELS_OUT=<file> the canonical labels of the last run are written as
// int32[V].  An optional 11th argument `--histogram` (not in the reference)
// adds, after each result block, the SCC count per size range:
//   SCC size histogram:
//   size 1, <n>
//   size 2-3, <n>
//   size 4-7, <n>      (one line per non-empty power-of-two range)
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <vector>

#include "ispan_hip.h"

static double wtime() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

static int die(const char* what, int rc, ispan_ctx* ctx) {
    fprintf(stderr, "%s failed: %s (%d)%s%s\n", what, ispan_strerror(rc), rc, ctx ? ": " : "",
            ctx ? ispan_ctx_last_error(ctx) : "");
    return 1;
}

int main(int argc, char** argv) {
    printf("args = %d\n", argc);
    const bool histogram = argc == 12 && strcmp(argv[11], "--histogram") == 0;
    if (argc != 11 && !histogram) {
        printf("Usage: ./scc_cpu <fw_beg_file> <fw_csr_file> <bw_beg_file> <bw_csr_file> <thread_count> <alpha> "
               "<beta> <gamma> <theta> <run_times>\n");
        return 255;
    }
    ispan_params P;
    ispan_default_params(&P);
    P.thread_count = atoi(argv[5]);
    P.alpha = (int)atof(argv[6]);  // truncated to int, as src/scc_main.cpp:21-23
    P.beta = (int)atof(argv[7]);
    P.gamma = (int)atof(argv[8]);
    P.theta = atof(argv[9]);
    const int run_times = atoi(argv[10]);
    P.timing = 1;
    printf("Thread = %d, alpha = %d, beta = %d, gamma = %d, theta = %g, run_times = %d\n", P.thread_count, P.alpha,
           P.beta, P.gamma, P.theta, run_times);

    const double t0 = wtime();
    ispan_host_graph g;
    int rc = ispan_graph_load(argv[1], argv[2], argv[3], argv[4], &g);
    if (rc) return die("graph load", rc, nullptr);
    printf("Graph load (success): %d verts, %lld edges %g second(s)\n", g.V, (long long)g.E, wtime() - t0);

    ispan_ctx* ctx = nullptr;
    if ((rc = ispan_ctx_create(&ctx, 0))) return die("context", rc, nullptr);
    if ((rc = ispan_graph_upload(ctx, g.V, g.E, g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj, &P)))
        return die("graph upload", rc, ctx);

    double avg[16] = {0};
    std::vector<int32_t> labels((size_t)g.V + 1);
    for (int r = 1; r <= run_times; ++r) {
        printf("\nRuntime: %d\n", r);
        ispan_stats st;
        if ((rc = ispan_solve(ctx, &P, nullptr, &st))) return die("solve", rc, ctx);
        const double* m = st.ms;
        // per-run block (src/scc_core.cpp:1066)
        printf("\ntime size_1_first, %.3lf\ntime size_1, %.3lf\ntime pivot, %.3lf\nlargest fw, %.3lf\nlargest bw, "
               "%.3lf\nlargest fw/bw, %.3lf\ntrim size_2, %.3lf\ntrim size_3, %.3lf\nwcc time, %.3lf\nmice fw-bw "
               "time, %.3lf\nmice scc time, %.3lf\ntotal time, %.3lf\n",
               m[11], m[14], m[6], m[7], m[8], m[6] + m[7] + m[8], m[5], m[13], m[9], m[10], m[9] + m[10], m[3]);
        for (int i = 0; i < 16; ++i) avg[i] += m[i];
        if ((rc = ispan_labels_download(ctx, labels.data()))) return die("labels", rc, ctx);
        ispan_summary sm;
        if ((rc = ispan_summarize(g.V, labels.data(), &sm))) return die("summary", rc, ctx);
        printf("\nResult:\nlargest, %lld\ntrimmed size_1, %lld\ntrimmed size_2, %lld\ntrimmed size_3, %lld\nothers, "
               "%lld\ntotal, %lld\n",
               (long long)sm.largest, (long long)sm.size_1, (long long)sm.size_2, (long long)sm.size_3,
               (long long)sm.others, (long long)sm.total);
        if (histogram) {
            int64_t bins[32];
            if ((rc = ispan_scc_histogram(ctx, bins, 32, nullptr))) return die("histogram", rc, ctx);
            printf("\nSCC size histogram:\n");
            for (int k = 0; k < 32; ++k) {
                if (!bins[k]) continue;
                if (k == 0) printf("size 1, %lld\n", (long long)bins[k]);
                else printf("size %lld-%lld, %lld\n", 1ll << k, (2ll << k) - 1, (long long)bins[k]);
            }
        }
    }
    const char* out = getenv("ISPAN_LABELS_OUT");
    if (out && *out && run_times > 0) {
        FILE* f = fopen(out, "wb");
        if (!f || fwrite(labels.data(), 4, (size_t)g.V, f) != (size_t)g.V) return die("label write", ISPAN_ERR_IO, ctx);
        fclose(f);
    }
    if (run_times > 0) {
        for (double& a : avg) a /= run_times;
        printf("\nAverage Time Consumption for Running %d Times (ms)\n", run_times);
        printf("Trim, %.3lf\n", avg[0]);
        printf("Elephant SCC, %.3lf\n", avg[1]);
        printf("Mice SCC, %.3lf\n", avg[2]);
        printf("Total time, %.3lf\n", avg[3]);
        printf("\n------Details------\n");
        printf("Trim size_1, %.3lf\n", avg[4]);
        printf("Pivot selection, %.3lf\n", avg[6]);
        printf("FW BFS, %.3lf\n", avg[7]);
        printf("BW BFS, %.3lf\n", avg[8]);
        printf("Trim size_2, %.3lf\n", avg[5]);
        printf("Trim size_3, %.3lf\n", avg[13]);
        printf("Wcc, %.3lf\n", avg[9]);
        printf("Mice fw-bw, %.3lf\n", avg[10]);
    }
    ispan_ctx_destroy(ctx);
    ispan_graph_free(&g);
    return 0;
}
