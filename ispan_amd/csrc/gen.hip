// Device-side producers of the iSpan CSR layout: seeded synthetic edge
// lists (R-MAT, lattice) and an edge-list -> CSR-pair builder (sort, drop
// self-loops, deduplicate, offsets by binary search, transpose).  This is
// the format producer of graph_converter/txt_to_bin_{fw,bw}_int.cpp
// (:31-217), which does not deduplicate; here both directions are built in
// one pass over sorted 64-bit (src,dst) keys.
//
// Randomness is a counter hash of (seed, draw index), so the edge multiset
// depends on nothing but the parameters.
#include <string.h>

#include <cstring>
#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <string>

#include "common.h"

namespace ispan {

namespace {

// SplitMix64 finaliser (Steele, Lea, Flood 2014) as a stateless mixer.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t stream_base(uint64_t seed, uint64_t tag) {
    return mix64(seed * 0x632BE59BD9B4E019ull + tag * 0xD1B54A32D192ED03ull + 1ull);
}

constexpr uint64_t kNoEdge = ~0ull;

// R-MAT draw e: at each of `scale` bit levels (most significant first) a
// 32-bit draw picks the quadrant: r < t1 (0,0), r < t2 (0,1), r < t3 (1,0),
// else (1,1).  Two levels share one 64-bit hash.
__global__ void k_rmat(int32_t V, int scale, int64_t m, uint32_t t1, uint32_t t2, uint32_t t3, uint64_t base,
                       uint64_t* keys) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
        uint32_t s = 0, d = 0;
        uint64_t h = 0;
        for (int k = 0; k < scale; ++k) {
            if ((k & 1) == 0) h = mix64(base ^ (((uint64_t)e << 5) | (uint64_t)(k >> 1)));
            const uint32_t r = (k & 1) ? (uint32_t)(h >> 32) : (uint32_t)h;
            const uint32_t sb = r >= t2;
            const uint32_t db = (r >= t1 && r < t2) || r >= t3;
            s |= sb << (scale - 1 - k);
            d |= db << (scale - 1 - k);
        }
        keys[e] = (s < (uint32_t)V && d < (uint32_t)V && s != d) ? (((uint64_t)s << 32) | d) : kNoEdge;
    }
}

// n x n lattice: draw i < L covers lattice edge i (horizontal edges first,
// row-major), slots 2i and 2i+1; then L/100 uniform random directed edges.
__global__ void k_grid(int32_t n, uint32_t tb, uint64_t b1, uint64_t b2, uint64_t* keys) {
    const int64_t H = (int64_t)n * (n - 1), L = 2 * H, R = L / 100;
    const uint64_t V = (uint64_t)n * (uint64_t)n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < L + R;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i < L) {
            int64_t u, v;
            if (i < H) {
                const int64_t r = i / (n - 1), c = i % (n - 1);
                u = r * n + c;
                v = u + 1;
            } else {
                const int64_t j = i - H;
                const int64_t r = j / n, c = j % n;
                u = r * n + c;
                v = u + n;
            }
            const uint64_t h = mix64(b1 ^ (uint64_t)i);
            const uint64_t uv = ((uint64_t)u << 32) | (uint64_t)v, vu = ((uint64_t)v << 32) | (uint64_t)u;
            if ((uint32_t)h < tb) {
                keys[2 * i] = uv;
                keys[2 * i + 1] = vu;
            } else {
                keys[2 * i] = (h >> 63) == 0 ? uv : vu;
                keys[2 * i + 1] = kNoEdge;
            }
        } else {
            const int64_t k = i - L;
            const uint64_t h = mix64(b2 ^ (uint64_t)k);
            const uint64_t s = ((h & 0xffffffffull) * V) >> 32, d = ((h >> 32) * V) >> 32;
            keys[2 * L + k] = s != d ? ((s << 32) | d) : kNoEdge;
        }
    }
}

__global__ void k_pack(const int32_t* __restrict__ src, const int32_t* __restrict__ dst, int64_t m, int32_t V,
                       uint64_t* keys, int32_t* bad) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = src[i], d = dst[i];
        if (s < 0 || s >= V || d < 0 || d >= V) {
            atomicAdd(bad, 1);
            keys[i] = kNoEdge;
        } else {
            keys[i] = s != d ? (((uint64_t)(uint32_t)s << 32) | (uint32_t)d) : kNoEdge;
        }
    }
}

// adj[i] = low word; beg[v] = first i with key >= v<<32 (binary search).
__global__ void k_csr_from_keys(const uint64_t* __restrict__ keys, int64_t E, int32_t V, int32_t* beg,
                                int32_t* adj) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t i = t; i < E; i += stride) adj[i] = (int32_t)(uint32_t)keys[i];
    for (int64_t v = t; v <= V; v += stride) {
        const uint64_t key = (uint64_t)v << 32;
        int64_t lo = 0, hi = E;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (keys[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        beg[v] = (int32_t)lo;
    }
}

__global__ void k_widen_keys(const uint32_t* __restrict__ k, int64_t m, uint64_t* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (uint64_t)k[i] << 32;
}

// beg[v] = first i with keys[i] >= v<<32 over sorted keys.
__global__ void k_beg_from_keys(const uint64_t* __restrict__ keys, int64_t m, int32_t V, int32_t* beg) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= V; v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = (uint64_t)v << 32;
        int64_t lo = 0, hi = m;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (keys[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        beg[v] = (int32_t)lo;
    }
}

__global__ void k_swap_halves(const uint64_t* __restrict__ in, int64_t E, uint64_t* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = in[i];
        out[i] = (k << 32) | (k >> 32);
    }
}

struct NotSentinel {
    __device__ bool operator()(uint64_t k) const { return k != kNoEdge; }
};

#define GK(call)                                                             \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            err = std::string(#call) + ": " + hipGetErrorString(e_);         \
            rc = e_ == hipErrorOutOfMemory ? -3 : -4;                        \
            goto done;                                                       \
        }                                                                    \
    } while (0)

constexpr int kGrid = 4096;

}  // namespace

// keys[0, m) (kNoEdge = absent) -> CSR pair.  keys is clobbered.
int csr_from_keys(hipStream_t s, int32_t V, int64_t m, uint64_t* keys, int32_t** fwb, int32_t** fwa, int32_t** bwb,
                  int32_t** bwa, int64_t* E_out, std::string& err) {
    int rc = 0;
    uint64_t* sorted = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0, b1 = 0, b2 = 0;
    int64_t* d_count = nullptr;
    int64_t E = 0;
    *fwb = *fwa = *bwb = *bwa = nullptr;
    GK(hipMalloc(&sorted, (size_t)std::max<int64_t>(m, 1) * 8));
    GK(hipMalloc(&d_count, 8));
    GK(rocprim::radix_sort_keys(nullptr, b1, keys, sorted, (size_t)m, 0, 64, s));
    GK(rocprim::unique(nullptr, b2, sorted, keys, d_count, (size_t)m, rocprim::equal_to<uint64_t>(), s));
    tmp_bytes = std::max(b1, b2) + 256;
    GK(hipMalloc(&tmp, tmp_bytes));
    if (m > 0) {
        b1 = tmp_bytes;
        GK(rocprim::radix_sort_keys(tmp, b1, keys, sorted, (size_t)m, 0, 64, s));
        b2 = tmp_bytes;
        GK(rocprim::unique(tmp, b2, sorted, keys, d_count, (size_t)m, rocprim::equal_to<uint64_t>(), s));
        GK(hipMemcpyAsync(&E, d_count, 8, hipMemcpyDeviceToHost, s));
        GK(hipStreamSynchronize(s));
        // the sentinel sorts last: drop it if present
        if (E > 0) {
            uint64_t last = 0;
            GK(hipMemcpy(&last, keys + E - 1, 8, hipMemcpyDeviceToHost));
            if (last == kNoEdge) --E;
        }
    }
    if (E > INT32_MAX) {
        err = "edge count exceeds int32";
        rc = -1;
        goto done;
    }
    GK(hipMalloc(fwb, ((size_t)V + 1) * 4));
    GK(hipMalloc(bwb, ((size_t)V + 1) * 4));
    GK(hipMalloc(fwa, (size_t)std::max<int64_t>(E, 1) * 4));
    GK(hipMalloc(bwa, (size_t)std::max<int64_t>(E, 1) * 4));
    hipLaunchKernelGGL(k_csr_from_keys, dim3(kGrid), dim3(256), 0, s, keys, E, V, *fwb, *fwa);
    GK(hipGetLastError());
    if (E > 0) {
        hipLaunchKernelGGL(k_swap_halves, dim3(kGrid), dim3(256), 0, s, keys, E, sorted);
        GK(hipGetLastError());
        b1 = tmp_bytes;
        GK(rocprim::radix_sort_keys(tmp, b1, sorted, keys, (size_t)E, 0, 64, s));
    }
    hipLaunchKernelGGL(k_csr_from_keys, dim3(kGrid), dim3(256), 0, s, keys, E, V, *bwb, *bwa);
    GK(hipGetLastError());
    GK(hipStreamSynchronize(s));
    *E_out = E;
done:
    if (sorted) (void)hipFree(sorted);
    if (tmp) (void)hipFree(tmp);
    if (d_count) (void)hipFree(d_count);
    if (rc) {
        for (int32_t** p : {fwb, fwa, bwb, bwa}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
    }
    return rc;
}

// Stable CSR of one direction: rows in input order, duplicates and
// self-loops kept (the byte layout of graph_converter/txt_to_bin_fw_int.cpp
// :146-181).  key[i] = row of edge i, val[i] = its other endpoint.
int csr_stable_one(hipStream_t s, int32_t V, int64_t m, const int32_t* key, const int32_t* val, int32_t** beg,
                   int32_t** adj, std::string& err) {
    int rc = 0;
    uint32_t *k2 = nullptr, *kin = nullptr;
    int32_t* v2 = nullptr;
    void* tmp = nullptr;
    size_t bytes = 0;
    uint64_t* keys64 = nullptr;
    *beg = *adj = nullptr;
    unsigned end_bit = 1;
    while (end_bit < 32 && ((int64_t)1 << end_bit) < (int64_t)V) ++end_bit;
    GK(hipMalloc(&kin, (size_t)std::max<int64_t>(m, 1) * 4));
    GK(hipMalloc(&k2, (size_t)std::max<int64_t>(m, 1) * 4));
    GK(hipMalloc(&v2, (size_t)std::max<int64_t>(m, 1) * 4));
    GK(hipMalloc(adj, (size_t)std::max<int64_t>(m, 1) * 4));
    GK(hipMalloc(beg, ((size_t)V + 1) * 4));
    GK(hipMalloc(&keys64, (size_t)std::max<int64_t>(m, 1) * 8));
    if (m > 0) {
        GK(hipMemcpyAsync(kin, key, (size_t)m * 4, hipMemcpyDeviceToDevice, s));
        GK(rocprim::radix_sort_pairs(nullptr, bytes, kin, k2, val, v2, (size_t)m, 0, end_bit, s));
        GK(hipMalloc(&tmp, bytes + 256));
        bytes += 256;
        GK(rocprim::radix_sort_pairs(tmp, bytes, kin, k2, val, v2, (size_t)m, 0, end_bit, s));
        GK(hipMemcpyAsync(*adj, v2, (size_t)m * 4, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_widen_keys, dim3(kGrid), dim3(256), 0, s, k2, (int64_t)m, keys64);
        GK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_beg_from_keys, dim3(kGrid), dim3(256), 0, s, keys64, (int64_t)m, V, *beg);
    GK(hipGetLastError());
    GK(hipStreamSynchronize(s));
done:
    for (void* p : {(void*)kin, (void*)k2, (void*)v2, tmp, (void*)keys64})
        if (p) (void)hipFree(p);
    if (rc) {
        if (*beg) (void)hipFree(*beg);
        if (*adj) (void)hipFree(*adj);
        *beg = *adj = nullptr;
    }
    return rc;
}

int gen_rmat_device(hipStream_t s, int32_t V, double ef, double a, double b, double c, uint64_t seed, int32_t** fwb,
                    int32_t** fwa, int32_t** bwb, int32_t** bwa, int64_t* E, std::string& err) {
    if (V <= 0 || ef < 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return -1;
    int scale = 0;
    while (((int64_t)1 << scale) < (int64_t)V) ++scale;
    const int64_t m = llround(ef * (double)V);
    const double S = 4294967296.0;
    const uint32_t t1 = (uint32_t)(a * S), t2 = (uint32_t)((a + b) * S), t3 = (uint32_t)((a + b + c) * S);
    uint64_t* keys = nullptr;
    int rc = 0;
    GK(hipMalloc(&keys, (size_t)std::max<int64_t>(m, 1) * 8));
    hipLaunchKernelGGL(k_rmat, dim3(kGrid), dim3(256), 0, s, V, scale, m, t1, t2, t3, stream_base(seed, 1), keys);
    GK(hipGetLastError());
    rc = csr_from_keys(s, V, m, keys, fwb, fwa, bwb, bwa, E, err);
done:
    if (keys) (void)hipFree(keys);
    return rc;
}

int gen_grid_device(hipStream_t s, int32_t n, double p_bidir, uint64_t seed, int32_t** fwb, int32_t** fwa,
                    int32_t** bwb, int32_t** bwa, int64_t* E, std::string& err) {
    if (n < 2 || (int64_t)n * n > INT32_MAX || p_bidir < 0 || p_bidir > 1) return -1;
    const int64_t L = 2 * (int64_t)n * (n - 1), R = L / 100, m = 2 * L + R;
    const uint32_t tb = (uint32_t)(p_bidir * 4294967296.0);
    uint64_t* keys = nullptr;
    int rc = 0;
    GK(hipMalloc(&keys, (size_t)m * 8));
    hipLaunchKernelGGL(k_grid, dim3(kGrid), dim3(256), 0, s, n, tb, stream_base(seed, 2), stream_base(seed, 3),
                       keys);
    GK(hipGetLastError());
    rc = csr_from_keys(s, n * n, m, keys, fwb, fwa, bwb, bwa, E, err);
done:
    if (keys) (void)hipFree(keys);
    return rc;
}

int build_csr_device(hipStream_t s, int32_t V, int64_t m, const int32_t* src, const int32_t* dst, int32_t** fwb,
                     int32_t** fwa, int32_t** bwb, int32_t** bwa, int64_t* E, std::string& err) {
    if (V < 0 || m < 0) return -1;
    uint64_t* keys = nullptr;
    int32_t* d_bad = nullptr;
    int32_t bad = 0;
    int rc = 0;
    GK(hipMalloc(&keys, (size_t)std::max<int64_t>(m, 1) * 8));
    GK(hipMalloc(&d_bad, 4));
    GK(hipMemsetAsync(d_bad, 0, 4, s));
    if (m > 0) {
        hipLaunchKernelGGL(k_pack, dim3(kGrid), dim3(256), 0, s, src, dst, m, V, keys, d_bad);
        GK(hipGetLastError());
    }
    GK(hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, s));
    GK(hipStreamSynchronize(s));
    if (bad) {
        err = "edge endpoint outside [0, V)";
        rc = -2;
        goto done;
    }
    rc = csr_from_keys(s, V, m, keys, fwb, fwa, bwb, bwa, E, err);
done:
    if (keys) (void)hipFree(keys);
    if (d_bad) (void)hipFree(d_bad);
    return rc;
}

}  // namespace ispan
