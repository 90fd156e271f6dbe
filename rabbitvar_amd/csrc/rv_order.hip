// nonInsertionVariants keys in creation order: compact the rows whose key exists into
// (first touch, position) pairs, then sort the pairs by first touch (hipCUB radix sort).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "rv_kernels.h"

namespace rv {

// one lane per dense row; one counter update per wavefront
__global__ void __launch_bounds__(256) k_key_list(const DenseRow* rows, int dense_n, int dense_lo,
                                                    unsigned long long* keys, int32_t* pos, unsigned* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool has = i < dense_n && (rows[i].rowflags & 1u);
    const unsigned long long mask = __ballot(has);
    const int lane = threadIdx.x & 63;
    unsigned base = 0;
    if (lane == 0 && mask) base = atomicAdd(count, (unsigned)__popcll(mask));
    base = __shfl(base, 0);
    if (has) {
        const unsigned off = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
        keys[off] = rows[i].first_touch;
        pos[off] = dense_lo + i;
    }
}

hipError_t launch_key_list(const DenseRow* rows, int dense_n, int dense_lo, unsigned long long* keys, int32_t* pos,
                           unsigned* count, hipStream_t s) {
    if (dense_n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_key_list, dim3((unsigned)((dense_n + 255) / 256)), dim3(256), 0, s, rows, dense_n, dense_lo,
                       keys, pos, count);
    return hipGetLastError();
}

// robin_hood slot-order sort key of each key, given its insertion index i: home slot (low hash
// bits), then the info byte's hash bits descending, then insertion index
__global__ void __launch_bounds__(256) k_slot_keys(const int32_t* keys, int n, int cap_log2, int idx_bits,
                                                     unsigned long long* comp, int32_t* val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t k = keys[i];
    const unsigned long long x = (unsigned long long)(long long)k;
    const unsigned long long hi = __umul64hi(x, 0xde5fb9d2630458e9ull);
    const unsigned long long lo = x * 0xde5fb9d2630458e9ull;
    const unsigned long long h = hi + lo;
    const unsigned long long home = h & ((1ull << cap_log2) - 1ull);
    const unsigned long long hb = h >> 59;
    comp[i] = (((home << 5) | (31ull - hb)) << idx_bits) | (unsigned long long)i;
    val[i] = k;
}

hipError_t launch_slot_keys(const int32_t* keys, int n, int cap_log2, int idx_bits, unsigned long long* comp,
                            int32_t* val, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_slot_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, keys, n, cap_log2, idx_bits, comp,
                       val);
    return hipGetLastError();
}

hipError_t sort_key_pairs(void* temp, size_t& temp_bytes, const unsigned long long* kin, unsigned long long* kout,
                          const int32_t* vin, int32_t* vout, int n, int end_bit, hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, kin, kout, vin, vout, n, 0, end_bit, s);
}

// Largest displacement of keys laid out in slot order (sorted by home): a key's slot is its index
// plus the running max of (home - index), so displacement_i = i + max_{j<=i}(home_j - j) - home_i.
__global__ void __launch_bounds__(256) k_disp_in(const int32_t* keys, int n, unsigned long long mask, int32_t* v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long x = (unsigned long long)(long long)keys[i];
    const unsigned long long h = __umul64hi(x, 0xde5fb9d2630458e9ull) + x * 0xde5fb9d2630458e9ull;
    v[i] = (int32_t)(h & mask) - i;
}

// each block reduces 8 keys per lane, then one atomic per block, skipped when it cannot raise *out
__global__ void __launch_bounds__(256) k_disp_max(const int32_t* v, const int32_t* pm, int n, int* out) {
    __shared__ int wmax[4];
    int d = 0;
    for (int j = 0; j < 8; j++) {
        const int i = (blockIdx.x * 8 + j) * 256 + (int)threadIdx.x;
        if (i < n) {
            const int x = pm[i] - v[i];  // (i + pm_i) - home_i = pm_i - (home_i - i)
            d = x > d ? x : d;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int w = __shfl_xor(d, o);
        d = w > d ? w : d;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = wmax[0];
        for (int w = 1; w < 4; w++) m = wmax[w] > m ? wmax[w] : m;
        if (m > *(volatile int*)out) atomicMax(out, m);
    }
}

hipError_t launch_max_displacement(const int32_t* keys, int n, int cap_log2, int32_t* v, int32_t* pm, int* out,
                                   void* temp, size_t& temp_bytes, hipStream_t s) {
    if (temp == nullptr)
        return hipcub::DeviceScan::InclusiveScan(nullptr, temp_bytes, v, pm, hipcub::Max(), n, s);
    if (n <= 0) return hipSuccess;
    const dim3 g((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_disp_in, g, dim3(256), 0, s, keys, n, (1ull << cap_log2) - 1ull, v);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::InclusiveScan(temp, temp_bytes, v, pm, hipcub::Max(), n, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_disp_max, dim3((unsigned)((n + 2047) / 2048)), dim3(256), 0, s, v, pm, n, out);
    return hipGetLastError();
}

}  // namespace rv
