// Realigner matching kernels (gfx950): one wavefront per candidate, the candidate's strings staged
// in LDS, the lanes spread over the candidate's start positions.
#include <hip/hip_runtime.h>

#include <climits>

#include "rv_realign_kernels.h"

namespace rv {

constexpr int F35_WAVES = 4;  // wavefronts (pairs) per workgroup

// find35match for one pair per wavefront.
//
// The reference tries every start (i, j), i in [0, l5 - 9], j in [1, l3 - 9] with i the outer loop,
// walks n = 0, 1, ... comparing seq5[i + n] with seq3[l3 - j - n] (a seq5 index past its end counts
// as a mismatch) until a third mismatch or an end, and keeps the first start whose run has more
// than 8 matches, under 10% mismatches, and reaches an end (src/VariationRealigner.cpp:1396-1428).
// Every start on the anti-diagonal D = i + (l3 - j) walks the same cells a = i, i + 1, ..., min(D, l5),
// so one lane scans one diagonal once, keeping the next three mismatches at or after the current
// start, and finds the diagonal's first qualifying start; the answer is the smallest (i, j) over
// the diagonals.
__global__ void __launch_bounds__(64 * F35_WAVES) k_find35(const int32_t* pair, int n, const char* arena,
                                                           F35Result* out) {
    __shared__ char s5[F35_WAVES][F35_MAX_LEN + 1];
    __shared__ char s3[F35_WAVES][F35_MAX_LEN + 1];
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int k = blockIdx.x * F35_WAVES + w;
    if (k >= n) return;  // whole wavefronts leave together: no barrier below spans wavefronts
    const int o5 = pair[4 * k], l5 = pair[4 * k + 1], o3 = pair[4 * k + 2], l3 = pair[4 * k + 3];
    if (l5 > F35_MAX_LEN || l3 > F35_MAX_LEN || l5 < 0 || l3 < 0) {
        if (lane == 0) out[k] = F35Result{-1, -1, -1};
        return;
    }
    for (int t = lane; t < l5; t += 64) s5[w][t] = arena[o5 + t];
    for (int t = lane; t < l3; t += 64) s3[w][t] = arena[o3 + t];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    long long best = LLONG_MAX;  // (i << 32) | j of this lane's best start
    int best_score = 0;
    if (l5 > 8 && l3 > 9) {
        const int d_lo = 9, d_hi = l5 - 9 + l3 - 1;
        for (int D = d_lo + lane; D <= d_hi; D += 64) {
            const int i_lo = max(0, D - l3 + 1), i_hi = min(l5 - 9, D - 9);
            if (i_lo > i_hi) continue;
            const int a_hi = min(D, l5);  // the last cell: seq3 index 0 or seq5's end
            int m0 = 0, m1 = 0, m2 = 0, nq = 0;  // the next mismatches at or after i, ascending
            int scan = i_lo;
            for (int i = i_lo; i <= i_hi; i++) {
                while (nq > 0 && m0 < i) {
                    m0 = m1;
                    m1 = m2;
                    nq--;
                }
                while (nq < 3 && scan <= a_hi) {
                    const bool mm = scan >= l5 || s3[w][D - scan] != s5[w][scan];
                    if (mm) {
                        if (nq == 0)
                            m0 = scan;
                        else if (nq == 1)
                            m1 = scan;
                        else
                            m2 = scan;
                        nq++;
                    }
                    scan++;
                }
                int len, nm;
                if (nq == 3) {
                    len = m2 - i;
                    nm = 3;
                } else {
                    len = a_hi - i + 1;
                    nm = nq;
                }
                const int j = i + l3 - D;
                if (len - nm > 8 && nm / (double)len < 0.1 && (len + j >= l3 || i + len >= l5)) {
                    const long long key = ((long long)i << 32) | (long long)j;
                    if (key < best) {
                        best = key;
                        best_score = len - nm;
                    }
                    break;  // later starts on this diagonal have larger i
                }
            }
        }
    }
    // smallest (i, j) over the wavefront's lanes
    for (int off = 32; off >= 1; off >>= 1) {
        const long long ob = __shfl_xor(best, off, 64);
        const int os = __shfl_xor(best_score, off, 64);
        if (ob < best) {
            best = ob;
            best_score = os;
        }
    }
    if (lane == 0) {
        if (best == LLONG_MAX)
            out[k] = F35Result{0, 0, 0};
        else
            out[k] = F35Result{(int)(best >> 32), (int)(best & 0xffffffffLL), best_score};
    }
}

hipError_t launch_find35(const int32_t* pair, int n, const char* arena, F35Result* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_find35, dim3((n + F35_WAVES - 1) / F35_WAVES), dim3(64 * F35_WAVES), 0, s, pair, n, arena,
                       out);
    return hipGetLastError();
}

}  // namespace rv
