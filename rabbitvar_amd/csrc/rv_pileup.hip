// Pileup kernels for gfx950: one thread per read for the prep and walk passes, one thread per
// dense row for initialisation, one thread per anchor record for the deferred subtraction.
#include <hip/hip_runtime.h>

#include "rv_kernels.h"
#include "rv_pileup_body.h"

namespace rv {

__global__ void __launch_bounds__(256) k_prep(ReadsDev R, Window W, Config cfg, PrepRec* prep, uint32_t* cig,
                                                WalkCounters* ctr, int* max_span) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R.n) prep_one(r, R, W, cfg, prep, cig, ctr, max_span);
}

__global__ void __launch_bounds__(256) k_fill_rows(DenseRow* rows, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) fill_row_one(rows, i);
}

__global__ void __launch_bounds__(256) k_walk(ReadsDev R, Window W, Config cfg, const PrepRec* prep, const uint32_t* cig,
                                                WalkSink sink) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R.n) walk_one(r, R, W, cfg, prep, cig, sink);
}

__global__ void __launch_bounds__(256) k_anchor(const Anchor* A, const unsigned* n_anchors, unsigned cap, DenseRow* rows,
                                                  int dense_lo, int dense_n, Config cfg) {
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned n = *n_anchors;
    if (n > cap) n = cap;
    if (i < n) anchor_one(i, A, rows, dense_lo, dense_n, cfg);
}

// Copies whole 64-row blocks of the dense window into a packed buffer (rows past the window are
// zero-filled).  One workgroup per block; each lane moves 16-byte words.
__global__ void __launch_bounds__(256) k_gather_blocks(const DenseRow* rows, int dense_n, const int32_t* blocks, int nb,
                                                         DenseRow* out) {
    constexpr int WORDS = (int)(sizeof(DenseRow) / 16);
    constexpr int ROWS = 64;
    const int b = blockIdx.x;
    if (b >= nb) return;
    const int64_t first = (int64_t)blocks[b] * ROWS;
    const uint4* src = reinterpret_cast<const uint4*>(rows);
    uint4* dst = reinterpret_cast<uint4*>(out + (int64_t)b * ROWS);
    for (int w = threadIdx.x; w < ROWS * WORDS; w += blockDim.x) {
        const int r = w / WORDS;
        const int64_t g = first + r;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (g >= 0 && g < dense_n) v = src[g * WORDS + (w - r * WORDS)];
        dst[w] = v;
    }
}

// ---------------------------------------------------------------- launch wrappers
static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

hipError_t launch_prep(const ReadsDev& R, const Window& W, const Config& cfg, PrepRec* prep, uint32_t* cig,
                       WalkCounters* ctr, int* max_span, hipStream_t s) {
    if (R.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_prep, dim3(grid_for(R.n, 256)), dim3(256), 0, s, R, W, cfg, prep, cig, ctr, max_span);
    return hipGetLastError();
}
hipError_t launch_fill_rows(DenseRow* rows, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_rows, dim3(grid_for(n, 256)), dim3(256), 0, s, rows, n);
    return hipGetLastError();
}
hipError_t launch_walk(const ReadsDev& R, const Window& W, const Config& cfg, const PrepRec* prep, const uint32_t* cig,
                       const WalkSink& sink, hipStream_t s) {
    if (R.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_walk, dim3(grid_for(R.n, 256)), dim3(256), 0, s, R, W, cfg, prep, cig, sink);
    return hipGetLastError();
}
hipError_t launch_anchor(const Anchor* a, const unsigned* n_anchors, unsigned cap, DenseRow* rows, int dense_lo,
                         int dense_n, const Config& cfg, hipStream_t s) {
    if (cap == 0) return hipSuccess;
    hipLaunchKernelGGL(k_anchor, dim3(grid_for(cap, 256)), dim3(256), 0, s, a, n_anchors, cap, rows, dense_lo, dense_n,
                       cfg);
    return hipGetLastError();
}

hipError_t launch_gather_blocks(const DenseRow* rows, int dense_n, const int32_t* blocks, int nb, DenseRow* out,
                                hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_blocks, dim3((unsigned)nb), dim3(256), 0, s, rows, dense_n, blocks, nb, out);
    return hipGetLastError();
}

}  // namespace rv
