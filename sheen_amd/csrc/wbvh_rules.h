// wbvh_rules.h — the arithmetic of the wide-BVH build, shared by the host
// builder (wbvh_build.cpp) and the device builder (wbvh_gpu.hip) so that both
// make the same choices bit for bit: centroids, bin indices, SAH costs in
// double, the area-greedy collapse and the child-box quantisation.
// Both sides compile with -ffp-contract=off, so every expression below rounds
// the same way on the CPU and the GPU.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "wbvh.h"

#if defined(__HIPCC__)
#define WB_HD __host__ __device__
#else
#define WB_HD
#endif


struct WbBox {
    float mn[3], mx[3];
};

WB_HD inline void wb_reset(WbBox& b) {
    b.mn[0] = b.mn[1] = b.mn[2] = INFINITY;
    b.mx[0] = b.mx[1] = b.mx[2] = -INFINITY;
}
// fminf/fmaxf ignore a NaN operand, as the host's std::fmin/std::fmax do
WB_HD inline void wb_grow(WbBox& b, const WbBox& o) {
    for (int a = 0; a < 3; a++) { b.mn[a] = fminf(b.mn[a], o.mn[a]); b.mx[a] = fmaxf(b.mx[a], o.mx[a]); }
}

// Input bounds are canonicalised: -0 becomes +0 (x + 0 rounds to +0 under
// round-to-nearest), so a box corner has one bit pattern whichever of two
// equal values a min/max keeps.
WB_HD inline float wb_canon(float x) { return x + 0.0f; }

// Centroid in double, then rounded: exact enough for the +-1e30 bounds of
// lathes, whose float sum would overflow nothing but lose the centre.
WB_HD inline float wb_centroid(float lo, float hi) { return (float)(0.5 * ((double)lo + (double)hi)); }

// Bin of a centroid offset already scaled by bins / extent: truncation,
// clamped to [0, bins); NaN and negative offsets land in bin 0.
WB_HD inline int wb_bin(float x) {
    if (!(x >= 0.0f)) return 0;
    if (x >= (float)WBVH_BINS) return WBVH_BINS - 1;
    return (int)x;
}

// Surface area in double; 0 for an empty or inverted box.
WB_HD inline double wb_area(const WbBox& b) {
    double dx = (double)b.mx[0] - b.mn[0], dy = (double)b.mx[1] - b.mn[1], dz = (double)b.mx[2] - b.mn[2];
    if (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

// Binned SAH choice over per-bin boxes and counts of the three axes.
// bins(a, b) -> const WbBox&, cnt(a, b) -> int; axes whose centroid extent is
// not positive are skipped.  Splits with an empty side are not candidates;
// the cost is area(left)·n_left + area(right)·n_right in double; the first
// strict minimum wins, axes in order x, y, z and, per axis, split planes from
// the top bin down.  axis = -1 when no split has both sides non-empty.
template <class BinF, class CntF>
WB_HD inline void wb_sah(const float* cmn, const float* cmx, BinF bins, CntF cnt, int& best_axis, int& best_bin) {
    best_axis = -1;
    best_bin = -1;
    double best_cost = INFINITY;
    for (int a = 0; a < 3; a++) {
        const float ext = cmx[a] - cmn[a];
        if (!(ext > 0.0f)) continue;
        WbBox lb[WBVH_BINS];
        int lc[WBVH_BINS];
        WbBox acc;
        wb_reset(acc);
        int ac = 0;
        for (int b = 0; b < WBVH_BINS - 1; b++) {
            wb_grow(acc, bins(a, b));
            ac += cnt(a, b);
            lb[b] = acc;
            lc[b] = ac;
        }
        wb_reset(acc);
        ac = 0;
        for (int b = WBVH_BINS - 1; b >= 1; b--) {
            wb_grow(acc, bins(a, b));
            ac += cnt(a, b);
            const int l = lc[b - 1];
            if (l == 0 || ac == 0) continue;
            const double c = wb_area(lb[b - 1]) * l + wb_area(acc) * ac;
            if (c < best_cost) { best_cost = c; best_axis = a; best_bin = b; }
        }
    }
}

// Bin of one centroid coordinate in a node whose centroid range on that
// axis starts at lo with scale k = bins / extent.
WB_HD inline int wb_bin_of(float c, float lo, float k) { return wb_bin((c - lo) * k); }
WB_HD inline float wb_bin_scale(float ext) { return (float)WBVH_BINS / ext; }

// Child planes are integers 0..WBVH_QMAX stored as fp16 bit patterns (exact:
// at most 11 significant bits), so the traversal feeds them to a
// mixed-precision FMA without a conversion.
WB_HD inline uint16_t wb_q_bits(uint32_t q) {
    if (q == 0) return 0;
    int e = 0;
    while ((q >> (e + 1)) != 0) e++;  // floor(log2 q), q <= 2048
    const uint32_t m = (e <= 10) ? (q << (10 - e)) : (q >> (e - 10));
    return (uint16_t)(((uint32_t)(e + 15) << 10) | (m & 0x3FFu));
}
// Value of such a pattern (zero or a normal fp16).
WB_HD inline float wb_q_value(uint16_t h) {
    if (h == 0) return 0.0f;
    const int e = (int)(h >> 10) - 15;
    const uint32_t m = (h & 0x3FFu) | 0x400u;
    return (e >= 10) ? (float)(m << (e - 10)) : (float)(m >> (10 - e));
}

// Exponent byte e (scale 2^(e-127)) such that extent / scale <= 2032:
// 2^k is the power of two just above extent / 2032.
WB_HD inline uint8_t wb_pick_exp(float ext) {
    if (!(ext > 0.0f)) return 1;
    if (!(ext <= 3.40282347e38f)) return 254;
    const double d = (double)ext / 2032.0;  // normal: ext >= 2^-149
    uint64_t bits;
    memcpy(&bits, &d, 8);
    const int k = (int)((bits >> 52) & 0x7FF) - 1022;  // d = f * 2^k, f in [0.5, 1)
    int e = k + 127;
    if (e < 1) e = 1;
    if (e > 254) e = 254;
    return (uint8_t)e;
}
WB_HD inline float wb_exp_scale(uint8_t e) {
    const uint32_t bits = (uint32_t)e << 23;
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

// Quantises nc child boxes into wn (origin, exponents, q).  A decoded bound
// fmaf(q, scale, origin) always contains the float bound; an axis whose
// children do not fit the 0..WBVH_QMAX grid at its exponent takes the next
// exponent.  Returns false when no exponent fits.
WB_HD inline bool wb_quantize(WNode& wn, const WbBox* cb, int nc) {
    WbBox nb;
    wb_reset(nb);
    for (int i = 0; i < nc; i++) wb_grow(nb, cb[i]);
    const float org[3] = {nb.mn[0], nb.mn[1], nb.mn[2]};
    wn.ox = org[0];
    wn.oy = org[1];
    wn.oz = org[2];
    uint8_t ex[3];
    for (int a = 0; a < 3; a++) ex[a] = wb_pick_exp(nb.mx[a] - nb.mn[a]);
    for (int a = 0; a < 3; a++) {
        for (;;) {
            const float sc = wb_exp_scale(ex[a]);
            bool ok = true;
            for (int i = 0; i < WBVH_WIDTH; i++) {
                if (i >= nc) { wn.q[2 * a][i] = wb_q_bits(WBVH_QMAX); wn.q[2 * a + 1][i] = 0; continue; }
                double lo = floor(((double)cb[i].mn[a] - (double)org[a]) / (double)sc);
                double hi = ceil(((double)cb[i].mx[a] - (double)org[a]) / (double)sc);
                lo = fmin(fmax(lo, 0.0), (double)(WBVH_QMAX + 1));  // also maps NaN to 0
                hi = fmin(fmax(hi, 0.0), (double)(WBVH_QMAX + 1));
                int64_t ql = (int64_t)lo, qh = (int64_t)hi;
                while (ql > 0 && fmaf((float)ql, sc, org[a]) > cb[i].mn[a]) ql--;
                while (qh < WBVH_QMAX && fmaf((float)qh, sc, org[a]) < cb[i].mx[a]) qh++;
                if (ql > WBVH_QMAX || qh > WBVH_QMAX || fmaf((float)ql, sc, org[a]) > cb[i].mn[a] ||
                    fmaf((float)qh, sc, org[a]) < cb[i].mx[a]) {
                    ok = false;
                    break;
                }
                wn.q[2 * a][i] = wb_q_bits((uint32_t)ql);
                wn.q[2 * a + 1][i] = wb_q_bits((uint32_t)qh);
            }
            if (ok) break;
            if (ex[a] >= 254) return false;
            ex[a]++;
        }
    }
    wn.ex = ex[0];
    wn.ey = ex[1];
    wn.ez = ex[2];
    return true;
}

// Leaves of the binary subtree at i, left to right, into out[0..cap); returns
// their number, or cap + 1 once the subtree is known to hold more than cap
// (every pending subtree holds at least one leaf, so the stack never exceeds
// cap entries).
template <class NodeF>
WB_HD inline int wb_treelet(int i, NodeF node, int* out, int cap) {
    int stk[WBVH_WIDTH + 1];
    int sp = 0, n = 0;
    stk[sp++] = i;
    while (sp > 0) {
        const int v = stk[--sp];
        WbBox box;
        int l, r, cnt;
        node(v, box, l, r, cnt);
        if (cnt > 0) {
            out[n++] = v;
            continue;
        }
        if (n + sp + 2 > cap) return cap + 1;
        stk[sp++] = r;
        stk[sp++] = l;
    }
    return n;
}

// One wide node's children from its binary node.
// * A leaf: the node itself.
// * A subtree of at most 8 leaves: all of its leaves (left to right), so the
//   bottom wide nodes are full instead of holding the two leaves of the
//   binary tree's last level.
// * Otherwise its two children; then repeatedly the child of largest area
//   that holds more than 8 leaves is replaced by its left child and its right
//   child appended (first in slot order on ties), until 8 slots or none is
//   left; then repeatedly the child of largest area whose leaves fit in the
//   slots left (a subtree of at most 8 leaves) is replaced by its leaves.
// node(i) -> (box, left, right, count); returns the number of slots.
template <class NodeF>
WB_HD inline int wb_collapse(int root, NodeF node, int* ch) {
    int l, r, cnt;
    WbBox box;
    node(root, box, l, r, cnt);
    if (cnt > 0) {
        ch[0] = root;
        return 1;
    }
    int n = wb_treelet(root, node, ch, WBVH_WIDTH);
    if (n <= WBVH_WIDTH) return n;
    int leaves[WBVH_WIDTH], tmp[WBVH_WIDTH];
    n = 0;
    ch[n] = l; leaves[n++] = wb_treelet(l, node, tmp, WBVH_WIDTH);
    ch[n] = r; leaves[n++] = wb_treelet(r, node, tmp, WBVH_WIDTH);
    while (n < WBVH_WIDTH) {
        int best = -1;
        double ba = -1.0;
        for (int k = 0; k < n; k++) {
            if (leaves[k] <= WBVH_WIDTH) continue;
            node(ch[k], box, l, r, cnt);
            const double a = wb_area(box);
            if (a > ba) { ba = a; best = k; }
        }
        if (best < 0) break;
        node(ch[best], box, l, r, cnt);
        ch[best] = l; leaves[best] = wb_treelet(l, node, tmp, WBVH_WIDTH);
        ch[n] = r; leaves[n++] = wb_treelet(r, node, tmp, WBVH_WIDTH);
    }
    for (;;) {
        int best = -1;
        double ba = -1.0;
        for (int k = 0; k < n; k++) {
            if (leaves[k] < 2 || n - 1 + leaves[k] > WBVH_WIDTH) continue;
            node(ch[k], box, l, r, cnt);
            const double a = wb_area(box);
            if (a > ba) { ba = a; best = k; }
        }
        if (best < 0) break;
        const int m = wb_treelet(ch[best], node, tmp, WBVH_WIDTH);
        ch[best] = tmp[0]; leaves[best] = 1;
        for (int k = 1; k < m; k++) { ch[n] = tmp[k]; leaves[n++] = 1; }
    }
    return n;
}
