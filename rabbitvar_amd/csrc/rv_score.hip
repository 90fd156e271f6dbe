// Scoring kernels (gfx950): one thread per position.
//   k_score_dense  positions whose alleles all live in the dense table (the vast majority)
//   k_score_cx     positions with descriptors, insertions or realignment changes (CSR records)
// Both run score_position(); the template parameters size the per-thread buffers.
#include <hip/hip_runtime.h>

#include "rv_kernels.h"
#include "rv_score_body.h"

#include <mutex>

namespace rv {

// Picks the dense positions that need scoring (score_needed) into S.pick_list.  Most positions
// end at those checks, and a wavefront of the scoring kernel takes as long as its slowest lane, so
// the scoring kernel then runs full wavefronts of positions that all do the work.  One ballot and
// one counter update per wavefront; the list's order does not matter (every record carries its
// position and rank).
__global__ void __launch_bounds__(256) k_score_pick(ScoreIn S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool need = false;
    const int p = S.dense_lo + i;
    if (i < S.dense_n) need = dense_needed(S, p);
    const unsigned long long m = __ballot(need);
    const int lane = (int)(threadIdx.x & 63u);
    unsigned base = 0;
    if (lane == 0 && m) base = atomicAdd(S.pick_count, (unsigned)__popcll(m));
    base = __shfl(base, 0);
    if (need) S.pick_list[base + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] = p;
}

// k_score_pick's row checks alone (no record list yet), into S.pick_list / S.pick_count
__global__ void __launch_bounds__(256) k_score_prepick(ScoreIn S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool need = false;
    const int p = S.dense_lo + i;
    if (i < S.dense_n) need = dense_row_needed(S, p);
    const unsigned long long m = __ballot(need);
    const int lane = (int)(threadIdx.x & 63u);
    unsigned base = 0;
    if (lane == 0 && m) base = atomicAdd(S.pick_count, (unsigned)__popcll(m));
    base = __shfl(base, 0);
    if (need) S.pick_list[base + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] = p;
}

// k_score_pick from the early list: its positions the record list does not carry
__global__ void __launch_bounds__(256) k_score_filter(ScoreIn S) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned total = *S.pre_count;
    bool need = false;
    int p = 0;
    if (i < total) {
        p = S.pre_list[i];
        need = cx_find(S, p) < 0;
    }
    const unsigned long long m = __ballot(need);
    const int lane = (int)(threadIdx.x & 63u);
    unsigned base = 0;
    if (lane == 0 && m) base = atomicAdd(S.pick_count, (unsigned)__popcll(m));
    base = __shfl(base, 0);
    if (need) S.pick_list[base + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] = p;
}

// n: the number of dense positions (the upper bound of the picked count).
__global__ void __launch_bounds__(128) k_score_dense(ScoreIn S, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = S.pick_list ? (int)*S.pick_count : n;
    if (i < m) {
        const int p = S.pick_list ? S.pick_list[i] : S.dense_lo + i;
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

__global__ void __launch_bounds__(64) k_zero_u32(unsigned* p, int n) {
    const int i = (int)threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_score(const ScoreIn& S, hipStream_t s) {
    if (S.dense_n > 0) {
        if (S.pick_list) {
            if (!S.pick_count) return hipErrorInvalidValue;
            // the count starts at zero with a fill kernel (no memset node if a caller captures this)
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, S.pick_count, 1);
            if (S.pre_list) {
                if (!S.pre_count) return hipErrorInvalidValue;
                // the early list holds at most dense_n positions
                hipLaunchKernelGGL(k_score_filter, dim3((S.dense_n + 255) / 256), dim3(256), 0, s, S);
            } else {
                hipLaunchKernelGGL(k_score_pick, dim3((S.dense_n + 255) / 256), dim3(256), 0, s, S);
            }
        }
        hipLaunchKernelGGL(k_score_dense, dim3((S.dense_n + 127) / 128), dim3(128), 0, s, S, S.dense_n);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (S.n_cx > 0) hipLaunchKernelGGL(k_score_cx, dim3((S.n_cx + 63) / 64), dim3(64), 0, s, S);
    return hipGetLastError();
}

hipError_t launch_score_prepick(const ScoreIn& S, hipStream_t s) {
    if (S.dense_n <= 0) return hipSuccess;
    if (!S.pick_list || !S.pick_count) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, S.pick_count, 1);
    hipLaunchKernelGGL(k_score_prepick, dim3((S.dense_n + 255) / 256), dim3(256), 0, s, S);
    return hipGetLastError();
}

hipError_t launch_fisher_table(const int* tab, int n, double* out, const double* lgt, int lgt_n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fisher_table, dim3((n + 255) / 256), dim3(256), 0, s, tab, n, out, lgt, lgt_n);
    return hipGetLastError();
}

// ---------------------------------------------------------------- the allocation-free Fisher entry
// rv_fisher_tables_noalloc works in this file-scope device array instead of engine buffers: the
// log-gamma table, then the tables in, then the p-values out, for at most FISHER_NOALLOC_MAX_TABLES tables
// per call.  Each GPU holds its own copy until the process ends; every call on one GPU uses it, so
// calls that overlap on one GPU race.  Nothing in it has to start at zero: a call writes the inputs
// it reads and every output it returns.
constexpr size_t FNA_LGT = 0;
constexpr size_t FNA_IN = FNA_LGT + sizeof(double) * FISHER_LGT_N;
constexpr size_t FNA_OUT = FNA_IN + sizeof(int) * 4 * FISHER_NOALLOC_MAX_TABLES;
constexpr size_t FNA_BYTES = FNA_OUT + sizeof(double) * 3 * FISHER_NOALLOC_MAX_TABLES;
static __device__ __align__(256) unsigned char scratch[FNA_BYTES];

namespace {
constexpr int FNA_MAX_DEVICES = 64;
std::mutex fna_mu;
unsigned char* fna_ptr[FNA_MAX_DEVICES];  // the device address of `scratch`, per device (nullptr: not set up)
double fna_host_lgt[FISHER_LGT_N];        // the host-built table, computed once (static storage)
bool fna_host_ready = false;
}  // namespace

// Set-up on the current device (once): the scratch array's device address, and the log-gamma table
// copied into it.
static hipError_t fisher_noalloc_setup(unsigned char*& base, const char** err) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *err = "no current HIP device";
        return e;
    }
    if (dev < 0 || dev >= FNA_MAX_DEVICES) {
        *err = "device index beyond the scratch table";
        return hipErrorInvalidDevice;
    }
    std::lock_guard<std::mutex> lk(fna_mu);
    if (!fna_ptr[dev]) {
        if (!fna_host_ready) {
            fisher_lgamma_table(fna_host_lgt, FISHER_LGT_N);
            fna_host_ready = true;
        }
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *err = "hipGetSymbolAddress(scratch) failed";
            return e != hipSuccess ? e : hipErrorInvalidSymbol;
        }
        e = hipMemcpy((unsigned char*)p + FNA_LGT, fna_host_lgt, sizeof fna_host_lgt, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            *err = "log-gamma table upload failed";
            return e;
        }
        fna_ptr[dev] = (unsigned char*)p;
    }
    base = fna_ptr[dev];
    return hipSuccess;
}

hipError_t fisher_tables_noalloc(const int* tab, int n, double* out, const char** err) {
    unsigned char* base = nullptr;
    hipError_t e = fisher_noalloc_setup(base, err);
    if (e != hipSuccess) return e;
    int* din = (int*)(base + FNA_IN);
    double* dout = (double*)(base + FNA_OUT);
    const double* lgt = (const double*)(base + FNA_LGT);
    if ((e = hipMemcpy(din, tab, sizeof(int) * 4 * (size_t)n, hipMemcpyHostToDevice)) != hipSuccess) {
        *err = "table upload failed";
        return e;
    }
    if ((e = launch_fisher_table(din, n, dout, lgt, FISHER_LGT_N, nullptr)) != hipSuccess) {
        *err = "Fisher kernel launch failed";
        return e;
    }
    if ((e = hipMemcpy(out, dout, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost)) != hipSuccess) {
        *err = "result download failed";
        return e;
    }
    return hipSuccess;
}

}  // namespace rv
