// Scoring kernels (gfx950): one thread per position.
//   k_score_dense  positions whose alleles all live in the dense table (the vast majority)
//   k_score_cx     positions with descriptors, insertions or realignment changes (CSR records)
// Both run score_position(); the template parameters size the per-thread buffers.
#include <hip/hip_runtime.h>

#include "rv_kernels.h"
#include "rv_score_body.h"

namespace rv {

__global__ void __launch_bounds__(128) k_score_dense(ScoreIn S, const int32_t* pos_list, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int p = pos_list ? pos_list[i] : S.dense_lo + i;
        score_position<128, 4>(S, p, true, p - S.dense_lo);
    }
}

__global__ void __launch_bounds__(64) k_score_cx(ScoreIn S) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S.n_cx) score_position<1536, MAXV>(S, S.cx_pos[i], false, S.dense_n + i);
}

__global__ void __launch_bounds__(256) k_fisher_table(const int* tab, int n, double* out, const double* lgt, int lgt_n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double l, r;
        double two = fisher_two_sided(tab[4 * i], tab[4 * i + 1], tab[4 * i + 2], tab[4 * i + 3], &l, &r,
                                      LGTable{lgt, lgt_n});
        out[3 * i] = two;
        out[3 * i + 1] = l;
        out[3 * i + 2] = r;
    }
}

hipError_t launch_score(const ScoreIn& S, hipStream_t s) {
    if (S.dense_n > 0)
        hipLaunchKernelGGL(k_score_dense, dim3((S.dense_n + 127) / 128), dim3(128), 0, s, S, (const int32_t*)nullptr,
                           S.dense_n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (S.n_cx > 0) hipLaunchKernelGGL(k_score_cx, dim3((S.n_cx + 63) / 64), dim3(64), 0, s, S);
    return hipGetLastError();
}

hipError_t launch_fisher_table(const int* tab, int n, double* out, const double* lgt, int lgt_n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fisher_table, dim3((n + 255) / 256), dim3(256), 0, s, tab, n, out, lgt, lgt_n);
    return hipGetLastError();
}

}  // namespace rv
