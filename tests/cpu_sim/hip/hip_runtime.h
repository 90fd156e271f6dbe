// Test harness only: host stand-ins for the HIP qualifiers and device atomics, so that the
// engine's per-item device bodies (rv_pileup_body.h, rv_score_body.h) can be compiled with g++ and
// run sequentially on the CPU.  Sequential execution gives one valid interleaving of the atomics.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <algorithm>
using std::exp;
using std::log;
#define __host__
#define __device__
#define __global__
#define __shared__
#define __forceinline__ inline
#define __launch_bounds__(x)
inline int atomicAdd(int* p, int v) { int o = *p; *p += v; return o; }
inline unsigned atomicAdd(unsigned* p, unsigned v) { unsigned o = *p; *p += v; return o; }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) { auto o = *p; *p += v; return o; }
inline int atomicSub(int* p, int v) { int o = *p; *p -= v; return o; }
inline unsigned atomicMin(unsigned* p, unsigned v) { unsigned o = *p; if (v < o) *p = v; return o; }
inline unsigned long long atomicMin(unsigned long long* p, unsigned long long v) { auto o = *p; if (v < o) *p = v; return o; }
inline unsigned atomicOr(unsigned* p, unsigned v) { unsigned o = *p; *p |= v; return o; }
inline unsigned atomicAnd(unsigned* p, unsigned v) { unsigned o = *p; *p &= v; return o; }
inline unsigned atomicCAS(unsigned* p, unsigned c, unsigned v) { unsigned o = *p; if (o == c) *p = v; return o; }
inline int atomicMax(int* p, int v) { int o = *p; if (v > o) *p = v; return o; }
// one lane: nothing to order between lanes
#define __builtin_amdgcn_fence(order, scope) ((void)0)
// a wavefront of one lane
constexpr int warpSize = 1;
inline unsigned long long __ballot(int pred) { return pred ? 1ull : 0ull; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline unsigned long long __builtin_amdgcn_read_exec() { return 1ull; }
inline unsigned __lane_id() { return 0u; }
inline int __builtin_amdgcn_readlane(int v, int) { return v; }
using std::min;
