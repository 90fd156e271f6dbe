// Pileup kernels for gfx950: one thread per read for the prep and walk passes, one thread per
// dense row for initialisation, one thread per anchor record for the deferred subtraction.
#include <hip/hip_runtime.h>

#include "rv_kernels.h"
#include "rv_pileup_body.h"

namespace rv {

__global__ void __launch_bounds__(256) k_prep(const WalkArgs* A, PrepRec* prep, uint32_t* cig, WalkCounters* ctr,
                                                int* max_span) {
    // arguments as values, so that the stores below cannot force reloads (see k_walk)
    const ReadsDev R = A->R;
    const Window W = A->W;
    const Config cfg = A->cfg;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R.n) prep_one(r, R, W, cfg, prep, cig, ctr, max_span, A->dup);
}

// -t: marks the reads the record filter drops as duplicates (before prep, which then skips them)
__global__ void __launch_bounds__(256) k_dedup(const WalkArgs* A) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < A->R.n) A->dup[r] = is_duplicate(A->R, A->cfg, r) ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_fill_rows(DenseRow* rows, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) fill_row_one(rows, i);
}

// One workgroup walks nreads consecutive reads (sorted by position), one wavefront per read at a
// time, into an LDS tile that starts TILE_MARGIN bases before the first read; the tile is added to
// the dense rows once at the end.

__global__ void __launch_bounds__(WALK_THREADS) k_walk(const WalkArgs* A, const PrepRec* prep, const uint32_t* cig,
                                                       int nreads) {
    const int waves = (int)(blockDim.x / warpSize);
    // the wavefront index is uniform across its lanes; saying so lets the per-read state live in
    // scalar registers
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / warpSize));
    const int lane = (int)(threadIdx.x % warpSize);
    // the launch arguments as values: read through a pointer, the compiler would reload them from
    // global memory after every store and atomic of the walk (they could alias)
    const ReadsDev R = A->R;
    const Window W = A->W;
    const Config cfg = A->cfg;
    const WalkSink S = A->S;
    const int64_t n = R.n;
    const int64_t r0 = (int64_t)blockIdx.x * nreads;
    const int tile_lo = (r0 < n ? prep[r0].position : 0) - TILE_MARGIN;
    // reads after each wavefront's first are handed out as wavefronts finish (reads differ in cost;
    // every contribution merges the same way whichever wavefront walks it)
    __shared__ int walk_next;
    if (threadIdx.x == 0) walk_next = waves;
    tile_init((int)threadIdx.x, (int)blockDim.x, W, tile_lo);
    __syncthreads();
    for (int k = wave; k < nreads;) {
        const int64_t r = r0 + k;
        if (r < n) walk_one(r, R, W, cfg, prep, cig, S, lane, wave, tile_lo);
        int nk = 0;
        if (lane == 0) nk = atomicAdd(&walk_next, 1);
        k = __builtin_amdgcn_readfirstlane(__shfl(nk, 0));
    }
    __syncthreads();
    tile_flush(tile_lo, S.dense, S.dense_lo, S.dense_n, (int)threadIdx.x, (int)blockDim.x);
}

__global__ void __launch_bounds__(256) k_anchor(const Anchor* A, const unsigned* n_anchors, unsigned cap, DenseRow* rows,
                                                  int dense_lo, int dense_n, Config cfg) {
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned n = *n_anchors;
    if (n > cap) n = cap;
    if (i < n) anchor_one(i, A, rows, dense_lo, dense_n, cfg);
}

// Copies the listed dense rows into a packed buffer (indices outside the window yield zero rows).
// One lane per 16-byte word.
__global__ void __launch_bounds__(256) k_gather_rows(const DenseRow* rows, int dense_n, const int32_t* idx, int n,
                                                       DenseRow* out) {
    constexpr int WORDS = (int)(sizeof(DenseRow) / 16);
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (int64_t)n * WORDS) return;
    const int64_t k = w / WORDS;
    const int j = (int)(w - k * WORDS);
    const int r = idx[k];
    uint4 v = make_uint4(0, 0, 0, 0);
    if (r >= 0 && r < dense_n) v = reinterpret_cast<const uint4*>(rows)[(int64_t)r * WORDS + j];
    reinterpret_cast<uint4*>(out)[w] = v;
}

// noPassingReads over the region's records: counts the records passes_one accepts.  One lane per
// record; one counter update per wavefront.
__global__ void __launch_bounds__(256) k_no_passing(const WalkArgs* A, int start, int end, unsigned* count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool passes = i < A->R.n && passes_one(A->R, i, start, end);
    const unsigned long long m = __ballot(passes);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (unsigned)__popcll(m));
}

// ---------------------------------------------------------------- launch wrappers
static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

hipError_t launch_prep(const WalkArgs* A, int64_t n, PrepRec* prep, uint32_t* cig, WalkCounters* ctr, int* max_span,
                       bool dedup, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (dedup) {
        hipLaunchKernelGGL(k_dedup, dim3(grid_for(n, 256)), dim3(256), 0, s, A);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_prep, dim3(grid_for(n, 256)), dim3(256), 0, s, A, prep, cig, ctr, max_span);
    return hipGetLastError();
}
hipError_t launch_fill_rows(DenseRow* rows, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_rows, dim3(grid_for(n, 256)), dim3(256), 0, s, rows, n);
    return hipGetLastError();
}
hipError_t launch_walk(const WalkArgs* A, int64_t n, const PrepRec* prep, const uint32_t* cig, int nreads,
                       hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (nreads < 1 || nreads > WALK_READS_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_walk, dim3(grid_for(n, nreads)), dim3(WALK_THREADS), 0, s, A, prep, cig, nreads);
    return hipGetLastError();
}
hipError_t launch_no_passing(const WalkArgs* A, int64_t n, int start, int end, unsigned* count, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_no_passing, dim3(grid_for(n, 256)), dim3(256), 0, s, A, start, end, count);
    return hipGetLastError();
}
hipError_t launch_anchor(const Anchor* a, const unsigned* n_anchors, unsigned cap, DenseRow* rows, int dense_lo,
                         int dense_n, const Config& cfg, hipStream_t s) {
    if (cap == 0) return hipSuccess;
    hipLaunchKernelGGL(k_anchor, dim3(grid_for(cap, 256)), dim3(256), 0, s, a, n_anchors, cap, rows, dense_lo, dense_n,
                       cfg);
    return hipGetLastError();
}

hipError_t launch_gather_rows(const DenseRow* rows, int dense_n, const int32_t* idx, int n, DenseRow* out,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    constexpr int WORDS = (int)(sizeof(DenseRow) / 16);
    hipLaunchKernelGGL(k_gather_rows, dim3(grid_for((int64_t)n * WORDS, 256)), dim3(256), 0, s, rows, dense_n, idx, n,
                       out);
    return hipGetLastError();
}

}  // namespace rv
