// Learned realignment screen: per-candidate features, and a small MLP (32 -> 64 -> 64 -> 1) that
// predicts whether the full realigner will change any count at a candidate.
//
// The reference has no counterpart: its realigner (src/VariationRealigner.cpp:135-161) visits every
// candidate.  In exact mode the engine does the same and the screen plays no part.  In the opt-in
// screened mode the realigner skips candidates the network scores below a threshold; the call set
// is then compared against the exact one (recall and precision).
//
// Arithmetic of the network (the CPU restatement in oracle/mlp.py follows it step for step):
//   operands of every product are BF16 (round to nearest even from FP32), sums are FP32;
//   z1 = bf16(X) bf16(W1) + b1, h1 = bf16(relu(z1)); z2 = h1 bf16(W2) + b2, h2 = bf16(relu(z2));
//   y = h2 . bf16(w3) + b3; p = 1 / (1 + exp(-y)).
// Training loss: mean over the batch of wt * BCE(p, t), wt = pos_weight for positives.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "rv_common.h"

namespace rv {

constexpr int SCR_F = 32;  // features per candidate
constexpr int SCR_H = 64;  // hidden width
// parameter vector (FP32 master copy): W1[F][H], b1[H], W2[H][H], b2[H], w3[H], b3
constexpr int SCR_W1 = 0;
constexpr int SCR_B1 = SCR_W1 + SCR_F * SCR_H;
constexpr int SCR_W2 = SCR_B1 + SCR_H;
constexpr int SCR_B2 = SCR_W2 + SCR_H * SCR_H;
constexpr int SCR_W3 = SCR_B2 + SCR_H;
constexpr int SCR_B3 = SCR_W3 + SCR_H;
constexpr int SCR_NPARAM = SCR_B3 + 1;

// Candidate kinds, in the realigner's stage order: adjustMNP, realigndel, realignins (by
// descriptor) and realignlgins30 (by 5' soft-clip cluster).
enum : int { SC_MNP = 0, SC_DEL = 1, SC_INS = 2, SC_LGINS = 3, SC_NKIND = 4 };

struct ScreenCand {
    int32_t kind;
    int32_t position;
    int32_t count;  // the candidate map's count (clip count for a cluster)
    int32_t len;    // MNV / deletion / insertion length; clip consensus columns for a cluster
};

// Neighbourhood tables: per table, sorted distinct positions and prefix sums of their counts
// (n + 1 entries).  Tables: MNP, deletion, insertion, 5' clip, 3' clip, and every candidate.
enum : int { ST_MNP = 0, ST_DEL = 1, ST_INS = 2, ST_SC5 = 3, ST_SC3 = 4, ST_ALL = 5, ST_N = 6 };
struct ScreenIn {
    const ScreenCand* cand;
    int n;
    const int32_t* tpos[ST_N];
    const int64_t* tcum[ST_N];
    int tn[ST_N];
    const DenseRow* rows;
    int dense_lo, dense_n;
    const char* ref;  // ref[k - ref_lo] for k in [ref_lo, ref_hi]
    int ref_lo, ref_hi;
    float* feat;  // [n][SCR_F]
};

RV_HD int scr_lower(const int32_t* a, int n, int32_t x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int m = (lo + hi) >> 1;
        if (a[m] < x)
            lo = m + 1;
        else
            hi = m;
    }
    return lo;
}
// (number of table positions, sum of their counts) in [a, b]
RV_HD void scr_range(const ScreenIn& S, int t, int a, int b, int& k, int64_t& c) {
    const int i = scr_lower(S.tpos[t], S.tn[t], a), j = scr_lower(S.tpos[t], S.tn[t], b + 1);
    k = j - i;
    c = S.tcum[t][j] - S.tcum[t][i];
}
RV_HD float scr_log1p(double v) { return (float)log1p(v > 0 ? v : 0.0); }

RV_HD char scr_ref(const ScreenIn& S, int k) { return (k >= S.ref_lo && k <= S.ref_hi) ? S.ref[k - S.ref_lo] : 0; }

RV_HD const DenseRow* scr_row(const ScreenIn& S, int p) {
    const int i = p - S.dense_lo;
    return (i >= 0 && i < S.dense_n) ? &S.rows[i] : nullptr;
}

// The 32 features of candidate i.  Every value is a count, a fraction or log1p of a count.
RV_HD void screen_features(const ScreenIn& S, int i) {
    const ScreenCand c = S.cand[i];
    float* f = S.feat + (size_t)i * SCR_F;
    for (int k = 0; k < SCR_F; k++) f[k] = 0.f;
    const int p = c.position;
    f[c.kind & 3] = 1.f;  // 0-3: kind
    f[4] = scr_log1p(c.count);
    f[5] = scr_log1p(c.len);
    // pileup at p and around it
    const char rb = scr_ref(S, p);
    const int rbi = base_index(rb);
    double cov = 0, all = 0, nonref = 0, lo = 0, nr_hi = 0, nr_f = 0, nr_r = 0;
    if (const DenseRow* r = scr_row(S, p)) {
        cov = r->cov;
        for (int b = 0; b < 4; b++) {
            const DenseBase& d = r->b[b];
            if (!(d.flags & 1u)) continue;
            all += d.cnt;
            lo += d.lo;
            if (b != rbi) {
                nonref += d.cnt;
                nr_hi += d.hi;
                nr_f += d.fwd;
                nr_r += d.rev;
            }
        }
    }
    f[6] = scr_log1p(cov);
    f[7] = (float)(c.count / (cov > 1 ? cov : 1.0));
    f[8] = (float)(nonref / (all > 1 ? all : 1.0));
    f[27] = (float)((nr_f > nr_r ? nr_f - nr_r : nr_r - nr_f) / (nr_f + nr_r > 1 ? nr_f + nr_r : 1.0));
    f[29] = (float)(lo / (all > 1 ? all : 1.0));
    f[30] = (float)(nr_hi / (nonref > 1 ? nonref : 1.0));
    // mismatch density and depth over +-10 and +-50
    for (int w = 0; w < 2; w++) {
        const int rad = w == 0 ? 10 : 50;
        double nr = 0, tot = 0, cv = 0;
        for (int q = p - rad; q <= p + rad; q++) {
            const DenseRow* r = scr_row(S, q);
            if (!r) continue;
            const int qb = base_index(scr_ref(S, q));
            cv += r->cov;
            for (int b = 0; b < 4; b++) {
                if (!(r->b[b].flags & 1u)) continue;
                tot += r->b[b].cnt;
                if (b != qb) nr += r->b[b].cnt;
            }
        }
        f[9 + w] = (float)(nr / (tot > 1 ? tot : 1.0));
        if (w == 1) f[11] = scr_log1p(cv / (2 * rad + 1)) - f[6];
    }
    // other candidates nearby
    int k;
    int64_t s;
    scr_range(S, ST_MNP, p - 150, p + 150, k, s);
    f[12] = scr_log1p(k);
    scr_range(S, ST_DEL, p - 150, p + 150, k, s);
    f[13] = scr_log1p(k);
    scr_range(S, ST_INS, p - 150, p + 150, k, s);
    f[14] = scr_log1p(k);
    scr_range(S, ST_SC5, p - 150, p + 150, k, s);
    f[15] = scr_log1p(k);
    f[17] = scr_log1p((double)s);
    scr_range(S, ST_SC3, p - 150, p + 150, k, s);
    f[16] = scr_log1p(k);
    f[18] = scr_log1p((double)s);
    scr_range(S, ST_SC5, p, p, k, s);
    f[19] = scr_log1p((double)s);
    scr_range(S, ST_SC3, p, p, k, s);
    f[20] = scr_log1p((double)s);
    scr_range(S, ST_DEL, p - 10, p + 10, k, s);
    f[21] = scr_log1p(k);
    scr_range(S, ST_INS, p - 10, p + 10, k, s);
    f[22] = scr_log1p(k);
    scr_range(S, ST_MNP, p - 10, p + 10, k, s);
    f[23] = scr_log1p(k);
    scr_range(S, ST_ALL, p - 2, p + 2, k, s);
    f[31] = scr_log1p(k);
    // distance to the nearest other candidate position
    {
        const int32_t* a = S.tpos[ST_ALL];
        const int n = S.tn[ST_ALL];
        const int j = scr_lower(a, n, p);
        int best = 1000;
        if (j > 0 && p - a[j - 1] < best) best = p - a[j - 1];
        int jn = (j < n && a[j] == p) ? j + 1 : j;
        if (jn < n && a[jn] - p < best) best = a[jn] - p;
        f[28] = scr_log1p(best);
    }
    // reference context: homopolymer run through p, longest tandem repeat (unit 2-6) from p, GC
    if (rb) {
        int run = 1;
        for (int q = p + 1; q <= p + 50 && scr_ref(S, q) == rb; q++) run++;
        for (int q = p - 1; q >= p - 50 && scr_ref(S, q) == rb; q--) run++;
        f[24] = scr_log1p(run);
        int best = 0;
        for (int u = 2; u <= 6; u++) {
            int len = u;
            while (len < 60 && scr_ref(S, p + len) != 0 && scr_ref(S, p + len) == scr_ref(S, p + len - u)) len++;
            if (len > best) best = len;
        }
        f[25] = scr_log1p(best);
    }
    {
        int gc = 0, m = 0;
        for (int q = p - 20; q <= p + 20; q++) {
            const char ch = scr_ref(S, q);
            if (!ch) continue;
            m++;
            if (ch == 'G' || ch == 'C') gc++;
        }
        f[26] = m ? (float)gc / (float)m : 0.f;
    }
}

// BF16 helpers (round to nearest even; inputs here are finite)
RV_HD uint16_t f2bf(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
RV_HD float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}

}  // namespace rv
