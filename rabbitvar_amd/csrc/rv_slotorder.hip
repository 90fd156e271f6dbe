// alignedVariants' robin_hood slot order on the device, one table growth (epoch) at a time.
//
// robin_hood keeps its keys sorted by home slot (an overflow area instead of wrap-around) and,
// within a home, by the info byte's hash bits descending, equal bytes in placement order.  An
// epoch re-inserts the previous table's keys in that table's slot order, then takes its own
// insertions (some erased right after: transient keys).  Per epoch:
//   1. sort the steps by (transient, home), stable: kept keys by home in placement order first;
//   2. a packed slot is index + the running max of (home - index); runs are separated by a free
//      slot;
//   3. each run that can reach a displacement of 6 is simulated insertion by insertion, one
//      wavefront per run, to find the first step at which each info-byte overflow displacement
//      (6, 14, 30, 62, 126) is reached (every set present in the epoch is a subset of its kept
//      keys plus one transient key, so runs evolve independently);
//   4. with the halving steps known, each key's place within its home follows from the hash bits
//      at its placement level, then placement order.
// The host computes the growth schedule and reads back the five first steps per epoch.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "rv_kernels.h"

namespace rv {

namespace {

__device__ inline unsigned long long so_mix(int32_t k) {
    const unsigned long long x = (unsigned long long)(long long)k;
    return __umul64hi(x, 0xde5fb9d2630458e9ull) + x * 0xde5fb9d2630458e9ull;
}

__device__ inline int32_t so_key(const SoEpoch& E, uint32_t s) {
    return s < (uint32_t)E.np ? E.prev[s] : E.ops[s - (uint32_t)E.np];
}

// displacement at which an info byte overflows at increments 32, 16, 8, 4, 2: 6, 14, 30, 62, 126
__device__ inline int so_disp(int level) { return (8 << level) - 2; }

__device__ inline int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) {
        const int w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}

}  // namespace

// step s -> sort key (transient << lcap | home) and value s
__global__ void __launch_bounds__(256) k_so_keys(SoEpoch E, uint32_t* skey, uint32_t* sval) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= E.S) return;
    const int32_t k = so_key(E, (uint32_t)s);
    const uint32_t t = (s >= E.np && E.tr[s - E.np]) ? 1u : 0u;
    skey[s] = (t << E.lcap) | (uint32_t)(so_mix(k) & E.mask);
    sval[s] = (uint32_t)s;
}

// v[i] = home_i - i over the kept keys (their sort keys are their homes)
__global__ void __launch_bounds__(256) k_so_diag(const uint32_t* skey, int n, int32_t* v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = (int32_t)skey[i] - i;
}

// run starts: a key whose home lies past the slot after the previous key's slot
__global__ void __launch_bounds__(256) k_so_runflag(const uint32_t* skey, const int32_t* pm, int n, int32_t* flag,
                                                     SoRuns R) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const long long prev_slot = i > 0 ? (long long)(i - 1) + pm[i - 1] : -2;
    flag[i] = (long long)skey[i] > prev_slot + 1 ? 1 : 0;
    R.maxd[i] = 0;
    R.tlo[i] = 0x7fffffff;
    R.tn[i] = 0;
}

__global__ void __launch_bounds__(256) k_so_runs(const uint32_t* skey, const int32_t* pm, const int32_t* flag,
                                                  const int32_t* rid, int n, SoRuns R) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int r = rid[i] - 1;
    const int slot = i + pm[i];
    const int home = (int)skey[i];
    if (flag[i]) {
        R.lo[r] = i;
        R.a[r] = home;
    }
    if (i == n - 1 || flag[i + 1]) {
        R.hi[r] = i + 1;
        R.b[r] = slot;
        if (i == n - 1) *R.n_runs = r + 1;
    }
    atomicMax(&R.maxd[r], slot - home);
}

// transient keys (sorted after the kept ones) whose home lies inside a run
__global__ void __launch_bounds__(256) k_so_trans(const uint32_t* skey, int n_keep, int S, uint32_t mask, SoRuns R) {
    const int j = n_keep + blockIdx.x * 256 + threadIdx.x;
    if (j >= S) return;
    const int home = (int)(skey[j] & mask);
    const int nr = *R.n_runs;
    int lo = 0, hi = nr;  // first run with a > home
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (R.a[mid] <= home)
            lo = mid + 1;
        else
            hi = mid;
    }
    const int r = lo - 1;
    if (r >= 0 && home <= R.b[r]) {
        atomicMin(&R.tlo[r], j);
        atomicAdd(&R.tn[r], 1);
    }
}

// One wavefront per run that can reach a displacement of 6: its events (kept keys and transient
// keys, sorted by step in LDS) through a slot model of the run.  Each insertion goes after the
// keys of smaller or equal home and shifts the rest up to the first free slot; a transient key is
// removed right after.  first[l] = the smallest step at which displacement so_disp(l) is reached.
__global__ void __launch_bounds__(64) k_so_simulate(const uint32_t* skey, const uint32_t* sval, uint32_t mask,
                                                     SoRuns R, unsigned* first, unsigned* too_big) {
    __shared__ unsigned long long ev[SO_EV_MAX];
    __shared__ int occ[SO_LEN_MAX + 2];
    __shared__ unsigned fl[5];
    const int lane = threadIdx.x;
    const int nr = *R.n_runs;
    for (int r = blockIdx.x; r < nr; r += gridDim.x) {
        const int nt = R.tn[r];
        if (R.maxd[r] + (nt > 0 ? 1 : 0) < so_disp(0)) continue;
        const int lo = R.lo[r], hi = R.hi[r], a = R.a[r];
        const int len = R.b[r] - a + 1;
        const int E = hi - lo + nt;
        if (E > SO_EV_MAX || len > SO_LEN_MAX) {
            if (lane == 0) atomicOr(too_big, 1u);
            continue;
        }
        int N = 64;
        while (N < E) N <<= 1;
        const int tlo = nt > 0 ? R.tlo[r] : 0;
        for (int e = lane; e < N; e += 64) {
            unsigned long long x = ~0ull;
            if (e < hi - lo) {
                const int i = lo + e;
                x = ((unsigned long long)sval[i] << 12) | ((unsigned long long)((int)skey[i] - a) << 1);
            } else if (e < E) {
                const int j = tlo + (e - (hi - lo));
                x = ((unsigned long long)sval[j] << 12) | ((unsigned long long)((int)(skey[j] & mask) - a) << 1) | 1ull;
            }
            ev[e] = x;
        }
        for (int s = lane; s < len + 2; s += 64) occ[s] = -1;
        __syncthreads();
        // bitonic sort by step (the high bits)
        for (int k = 2; k <= N; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int t = lane; t < N; t += 64) {
                    const int u = t ^ j;
                    if (u > t) {
                        const unsigned long long p = ev[t], q = ev[u];
                        const bool up = (t & k) == 0;
                        if ((p > q) == up) {
                            ev[t] = q;
                            ev[u] = p;
                        }
                    }
                }
                __syncthreads();
            }
        int maxd = 0, reached = 0;
        for (int e = 0; e < E && reached < 5; e++) {
            const unsigned long long x = ev[e];
            const unsigned step = (unsigned)(x >> 12);
            const int u = (int)((x >> 1) & 2047);
            const bool transient = (x & 1ull) != 0;
            // p: the first slot at or after u holding no key of smaller or equal home
            int p = -1;
            for (int base = u; p < 0; base += 64) {
                const int idx = base + lane;
                const int o = idx <= len + 1 ? occ[idx] : -1;
                const unsigned long long m = __ballot(o < 0 || o > u);
                if (m) p = base + __ffsll((long long)m) - 1;
            }
            // q: the first free slot at or after p; displacements of the keys shifted up
            int q = -1, md = p - u;
            for (int base = p; q < 0; base += 64) {
                const int idx = base + lane;
                const int o = idx <= len + 1 ? occ[idx] : -1;
                const unsigned long long m = __ballot(o < 0);
                const int qq = m ? base + __ffsll((long long)m) - 1 : 0x7fffffff;
                const int d = (idx < qq && o >= 0) ? idx + 1 - o : 0;
                const int w = wave_max(d);
                md = w > md ? w : md;
                if (m) q = qq;
            }
            md = md > maxd ? md : maxd;
            __syncthreads();
            if (!transient) {
                for (int top = q - 1; top >= p; top -= 64) {
                    const int idx = top - lane;
                    const int o = idx >= p ? occ[idx] : 0;
                    __syncthreads();
                    if (idx >= p) occ[idx + 1] = o;
                    __syncthreads();
                }
                if (lane == 0) occ[p] = u;
                __syncthreads();
                maxd = md;
            }
            while (reached < 5 && md >= so_disp(reached)) {
                if (lane == 0) fl[reached] = step;
                reached++;
            }
        }
        __syncthreads();
        if (lane == 0)
            for (int l = 0; l < reached; l++) atomicMin(&first[l], fl[l]);
        __syncthreads();
    }
}

// The halving steps of an epoch from the first steps reaching each displacement: a flag raised at
// step t halves the increment before step t + 1 (when there is one).  Reaching the last level's
// displacement (the table would grow early) sets *declined.
__global__ void k_so_halve(const unsigned* first, int S, SoHalve* H, unsigned* declined) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int nh = 0;
    for (int l = 0; l < 4; l++)
        if (first[l] != 0xffffffffu && first[l] + 1u < (unsigned)S) H->at[nh++] = first[l] + 1u;
    H->nh = nh;
    if (first[4] != 0xffffffffu) *declined = 1u;
}

// Each kept key's place within its home group: ranked by the level key then the step; groups of
// more than SO_GROUP_MAX keys flag too_big.
__global__ void __launch_bounds__(256) k_so_final(const uint32_t* skey, const uint32_t* sval, SoEpoch E,
                                                   const SoHalve* Hp, int32_t* out, unsigned* too_big) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= E.n_keep) return;
    const SoHalve H = *Hp;
    const uint32_t home = skey[i];
    int g0 = i;
    while (g0 > 0 && skey[g0 - 1] == home && i - g0 < SO_GROUP_MAX) g0--;
    int g1 = i + 1;
    while (g1 < E.n_keep && skey[g1] == home && g1 - i < SO_GROUP_MAX) g1++;
    const int32_t key = so_key(E, sval[i]);
    if (g1 - g0 == 1) {
        out[i] = key;
        return;
    }
    if (g1 - g0 >= SO_GROUP_MAX) {
        atomicOr(too_big, 1u);
        return;
    }
    auto level_key = [&](uint32_t step, int32_t k) {
        const uint32_t hb = (uint32_t)(so_mix(k) >> 59);
        int ph = 0;
        while (ph < H.nh && H.at[ph] <= step) ph++;
        uint32_t c = (~(hb >> H.nh)) & ((1u << (5 - H.nh)) - 1u);
        for (int l = H.nh - 1; l >= 0; l--) {
            const uint32_t late = ph > l ? 1u : 0u;
            const uint32_t bit = late ? 0u : ((~hb >> l) & 1u);
            c = (c << 2) | (late << 1) | bit;
        }
        return c;
    };
    const uint32_t mine = level_key(sval[i], key);
    int rank = 0;
    for (int j = g0; j < g1; j++) {
        if (j == i) continue;
        const uint32_t sj = sval[j];
        const uint32_t cj = level_key(sj, so_key(E, sj));
        if (cj < mine || (cj == mine && sj < sval[i])) rank++;
    }
    out[g0 + rank] = key;
}

// ---- alignedVariants' operations: the processing order's positions whose status is nonzero, in
// order (status 2: erased right after its insertion)

// record-list positions override their dense rows' status (one status per window position)
__global__ void __launch_bounds__(256) k_so_status_merge(uint8_t* status, int dense_lo, int dense_n,
                                                          const int32_t* cx_pos, int ncx) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= ncx) return;
    const int i = cx_pos[k] - dense_lo;
    if (i >= 0 && i < dense_n) status[i] = status[dense_n + k];
}

namespace {
__device__ inline int so_status(const uint8_t* st, int dense_lo, int dense_n, int32_t p) {
    const long long i = (long long)p - dense_lo;
    return (i >= 0 && i < dense_n) ? st[i] : 0;
}
}  // namespace

// selected positions per block of SO_CB processing entries
__global__ void __launch_bounds__(256) k_so_count(const int32_t* proc, int n, const uint8_t* st, int dense_lo,
                                                   int dense_n, int32_t* bcount) {
    __shared__ int wsum[4];
    const int base = blockIdx.x * SO_CB;
    int c = 0;
    for (int j = 0; j < SO_CB / 256; j++) {
        const int i = base + j * 256 + (int)threadIdx.x;
        if (i < n && so_status(st, dense_lo, dense_n, proc[i])) c++;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// the selected positions written in processing order from the block offsets
__global__ void __launch_bounds__(256) k_so_compact(const int32_t* proc, int n, const uint8_t* st, int dense_lo,
                                                     int dense_n, const int32_t* boff, int32_t* ops, uint8_t* tr) {
    __shared__ int wcnt[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = blockIdx.x * SO_CB;
    int run = boff[blockIdx.x];
    for (int j = 0; j < SO_CB / 256; j++) {
        const int i = base + j * 256 + (int)threadIdx.x;
        const int32_t p = i < n ? proc[i] : 0;
        const int v = i < n ? so_status(st, dense_lo, dense_n, p) : 0;
        const unsigned long long m = __ballot(v != 0);
        if (lane == 0) wcnt[wave] = __popcll(m);
        __syncthreads();
        int woff = 0;
        for (int w = 0; w < wave; w++) woff += wcnt[w];
        const int total = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        if (v) {
            const int at = run + woff + __popcll(m & ((1ull << lane) - 1ull));
            ops[at] = p;
            tr[at] = v == 2 ? 1 : 0;
        }
        __syncthreads();
        run += total;
    }
}

hipError_t launch_so_ops(const int32_t* proc, int n, uint8_t* status, int dense_lo, int dense_n, const int32_t* cx_pos,
                         int ncx, int32_t* bcount, int32_t* boff, void* tmp, size_t tmp_bytes, int32_t* ops, uint8_t* tr,
                         hipStream_t s) {
    hipError_t e;
    if (ncx > 0) {
        hipLaunchKernelGGL(k_so_status_merge, dim3((unsigned)((ncx + 255) / 256)), dim3(256), 0, s, status, dense_lo,
                           dense_n, cx_pos, ncx);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const int nb = (n + SO_CB - 1) / SO_CB;
    if ((e = hipMemsetAsync(bcount + nb, 0, sizeof(int32_t), s)) != hipSuccess) return e;
    if (nb > 0) {
        hipLaunchKernelGGL(k_so_count, dim3((unsigned)nb), dim3(256), 0, s, proc, n, status, dense_lo, dense_n, bcount);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    size_t tb = tmp_bytes;
    if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, bcount, boff, nb + 1, s)) != hipSuccess) return e;
    if (nb > 0) {
        hipLaunchKernelGGL(k_so_compact, dim3((unsigned)nb), dim3(256), 0, s, proc, n, status, dense_lo, dense_n, boff,
                           ops, tr);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t so_ops_scan_bytes(int nb, size_t& tmp_bytes) {
    return hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, (const int32_t*)nullptr, (int32_t*)nullptr, nb + 1,
                                            (hipStream_t)0);
}

// ---- the printer's record order: records by (their position's index in the slot order, position,
// allele order); positions without an alignedVariants key sort after all others

__global__ void __launch_bounds__(256) k_so_rank_scatter(const int32_t* order, int n_keep, int dense_lo, int dense_n,
                                                          int32_t* rank_of) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n_keep) return;
    const long long i = (long long)order[r] - dense_lo;
    if (i >= 0 && i < dense_n) rank_of[i] = r;
}

__global__ void __launch_bounds__(256) k_so_rank_keys(const OutRec* recs, int n, const int32_t* rank_of, int dense_lo,
                                                       int dense_n, long long pos_base, unsigned long long* key,
                                                       int32_t* idx, unsigned* bad) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int32_t p = recs[j].position, ord = recs[j].order;
    const long long i = (long long)p - dense_lo;
    const int32_t r = (i >= 0 && i < dense_n) ? rank_of[i] : -1;
    const unsigned long long rk = r < 0 ? 0xffffffull : (unsigned long long)r;
    const long long off = (long long)p - pos_base;
    if (r >= 0xffffff || off < 0 || off >= (1ll << 30) || ord < 0 || ord >= 1024) atomicOr(bad, 1u);
    key[j] = (rk << 40) | ((unsigned long long)(off & ((1ll << 30) - 1)) << 10) | (unsigned long long)(ord & 1023);
    idx[j] = j;
}

hipError_t launch_so_rank(const int32_t* order, int n_keep, const OutRec* recs, int n, int dense_lo, int dense_n,
                          long long pos_base, int32_t* rank_of, unsigned long long* key, int32_t* idx, unsigned* bad,
                          hipStream_t s) {
    hipError_t e;
    if (dense_n > 0 && (e = hipMemsetAsync(rank_of, 0xff, sizeof(int32_t) * (size_t)dense_n, s)) != hipSuccess) return e;
    if (n_keep > 0) {
        hipLaunchKernelGGL(k_so_rank_scatter, dim3((unsigned)((n_keep + 255) / 256)), dim3(256), 0, s, order, n_keep,
                           dense_lo, dense_n, rank_of);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (n > 0) {
        hipLaunchKernelGGL(k_so_rank_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, recs, n, rank_of, dense_lo,
                           dense_n, pos_base, key, idx, bad);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_so_epoch_keys(const SoEpoch& E, uint32_t* skey, uint32_t* sval, hipStream_t s) {
    if (E.S <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_so_keys, dim3((unsigned)((E.S + 255) / 256)), dim3(256), 0, s, E, skey, sval);
    return hipGetLastError();
}

hipError_t launch_so_runs(const uint32_t* skey, int n_keep, int S, uint32_t mask, int32_t* v, int32_t* pm,
                          int32_t* flag, int32_t* rid, const SoRuns& R, void* tmp, size_t tmp_bytes, hipStream_t s) {
    if (n_keep <= 0) return hipSuccess;
    const dim3 g((unsigned)((n_keep + 255) / 256));
    hipLaunchKernelGGL(k_so_diag, g, dim3(256), 0, s, skey, n_keep, v);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t tb = tmp_bytes;
    e = hipcub::DeviceScan::InclusiveScan(tmp, tb, v, pm, hipcub::Max(), n_keep, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_so_runflag, g, dim3(256), 0, s, skey, pm, n_keep, flag, R);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    tb = tmp_bytes;
    e = hipcub::DeviceScan::InclusiveSum(tmp, tb, flag, rid, n_keep, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_so_runs, g, dim3(256), 0, s, skey, pm, flag, rid, n_keep, R);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (S > n_keep) {
        hipLaunchKernelGGL(k_so_trans, dim3((unsigned)((S - n_keep + 255) / 256)), dim3(256), 0, s, skey, n_keep, S, mask,
                           R);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_so_simulate(const uint32_t* skey, const uint32_t* sval, uint32_t mask, const SoRuns& R, int max_runs,
                              unsigned* first, unsigned* too_big, hipStream_t s) {
    if (max_runs <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)(max_runs < 8192 ? max_runs : 8192);
    hipLaunchKernelGGL(k_so_simulate, dim3(blocks), dim3(64), 0, s, skey, sval, mask, R, first, too_big);
    return hipGetLastError();
}

hipError_t launch_so_final(const uint32_t* skey, const uint32_t* sval, const SoEpoch& E, const unsigned* first, SoHalve* H,
                           int32_t* out, unsigned* too_big, unsigned* declined, hipStream_t s) {
    hipLaunchKernelGGL(k_so_halve, dim3(1), dim3(64), 0, s, first, E.S, H, declined);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || E.n_keep <= 0) return e;
    hipLaunchKernelGGL(k_so_final, dim3((unsigned)((E.n_keep + 255) / 256)), dim3(256), 0, s, skey, sval, E,
                       (const SoHalve*)H, out, too_big);
    return hipGetLastError();
}

hipError_t so_sort_pairs(void* tmp, size_t& tmp_bytes, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                         uint32_t* vout, int n, int end_bit, hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, n, 0, end_bit, s);
}

hipError_t so_scan_bytes(int n, size_t& tmp_bytes) {
    size_t a = 0, b = 0;
    hipError_t e = hipcub::DeviceScan::InclusiveScan(nullptr, a, (const int32_t*)nullptr, (int32_t*)nullptr, hipcub::Max(),
                                                     n, (hipStream_t)0);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::InclusiveSum(nullptr, b, (const int32_t*)nullptr, (int32_t*)nullptr, n, (hipStream_t)0);
    tmp_bytes = a > b ? a : b;
    return e;
}

}  // namespace rv
