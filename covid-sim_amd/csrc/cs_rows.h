// cs_rows.h -- the cell tables of the spatial destination choice: max_trans,
// the per-row cumulative tables UpdateProbs builds (full or checkpointed)
// and the destination search over them.  Device functions only, shared by
// the infection step (cs_kernels.h) and the population generator
// (cs_popgen.hip), whose place assignment draws destination cells from the
// same tables built over capacity slots (src/SetupModel.cpp:2384-2392).
#pragma once
#include "cs_model.h"

namespace cs {

// bit 31 of an InvCDF entry: the bucket stops at its cell without a scan
#define CS_INV_PURE 0x80000000u

// max_trans[l][m] = (float)num(dist2_cc_min(l, m)) (src/Kernels.cpp:48-63),
// l and m populated cells (CellLookup order); their grid coordinates come
// from d.cxy.  INL inlines the distance (loops over many independent pairs)
template <bool INL = false>
__device__ __forceinline__ float max_trans_of(const Dev& d, int l, int m)
{
    const int2 a = d.cxy[l], b = d.cxy[m];
    return (float)kernel_num(d, dist2_cc_min_xy<INL>(d.P, a.x, a.y, b.x, b.y));
}

// row j of realisation r from d.S0 (src/CovidSim.cpp:2780-2800): summed in
// the reference's order so the float results are identical
__device__ inline void update_row(const Dev& d, int r, int j)
{
    const size_t i = (size_t)r * d.NC + j;
    const int NC = d.NC;
    float* ct = d.cum_trans + i * NC;
    const float* mt = d.max_trans + (size_t)j * NC;
    const int32_t* S0 = d.S0 + (size_t)r * NC;
    float t;
    ct[0] = ((float)S0[0]) * mt[0];
    t = ((float)d.cn[0]) * mt[0];
    for (int m = 1; m < NC; m++) {
        ct[m] = ct[m - 1] + ((float)S0[m]) * mt[m];
        t += ((float)d.cn[m]) * mt[m];
    }
    float tp = ct[NC - 1];
    for (int m = 0; m < NC; m++) ct[m] /= tp;
    d.tot_prob[i] = tp / t;
    int32_t* inv = d.invcdf + i * 1025;
    int m = 0;
    for (int k = 0; k <= 1024; k++) {
        while (ct[m] * 1024 < ((float)k)) m++;
        // bit 31: every u of bucket k (u < (k+1)/1024) stops at m, because
        // cum_trans[m] already reaches the bucket's upper end, so the scan
        // in select_dest need not read cum_trans (the same cell either way)
        const bool pure = k < 1024 && (double)ct[m] >= (double)(k + 1) / 1024.0;
        inv[k] = pure ? (int32_t)((uint32_t)m | CS_INV_PURE) : m;
    }
}

// update_row by one whole wave: the same float sums, added one term at a time
// in cell order (each lane's term read into the running sums in turn, so
// every lane holds the same sums), with the row's loads, the normalisation
// and InvCDF spread over the lanes.  InvCDF[k] is the first m with
// cum_trans[m] * 1024 >= k (exact in float), so cell m is InvCDF's for
// floor(1024 cum_trans[m - 1]) < k <= floor(1024 cum_trans[m]) and the lane
// holding m writes those k.  A row of no weight (tp = 0: NaN prefixes)
// leaves every InvCDF entry at 0, as the sequential loop does.
__device__ inline void update_row_wave(const Dev& d, int r, int j)
{
    const int lane = (int)(threadIdx.x & 63);
    const size_t i = (size_t)r * d.NC + j;
    const int NC = d.NC;
    float* ct = d.cum_trans + i * NC;
    const float* mt = d.max_trans + (size_t)j * NC;
    const int32_t* S0 = d.S0 + (size_t)r * NC;
    float c = 0.f, t = 0.f;  // 0 + x == x: the first terms are the reference's ct[0] and t
    for (int m0 = 0; m0 < NC; m0 += 64) {
        const int m = m0 + lane;
        float xs = 0.f, xt = 0.f;
        if (m < NC) {
            const float w = mt[m];
            xs = ((float)S0[m]) * w;
            xt = ((float)d.cn[m]) * w;
        }
        const int cnt = min(64, NC - m0);
        float mine = 0.f;
        for (int k = 0; k < cnt; k++) {
            c = c + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xs), k));
            t = t + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xt), k));
            if (lane == k) mine = c;
        }
        if (m < NC) ct[m] = mine;
    }
    const float tp = c;
    for (int m = lane; m < NC; m += 64) ct[m] = ct[m] / tp;
    if (lane == 0) d.tot_prob[i] = tp / t;
    int32_t* inv = d.invcdf + i * 1025;
    if (!(tp / tp == 1.0f)) {
        for (int k = lane; k <= 1024; k += 64) inv[k] = 0;
        return;
    }
    for (int m = lane; m < NC; m += 64) {
        const float v = ct[m];
        const int klo = m ? (int)__builtin_floorf(ct[m - 1] * 1024.0f) + 1 : 0;
        const int khi = min(1024, (int)__builtin_floorf(v * 1024.0f));
        for (int k = klo; k <= khi; k++) {
            const bool pure = k < 1024 && (double)v >= (double)(k + 1) / 1024.0;
            inv[k] = pure ? (int32_t)((uint32_t)m | CS_INV_PURE) : m;
        }
    }
}

// destination cell: the cell where the reference's InvCDF start and
// cum_trans scan stop (src/Sweep.cpp:757-762), the first l from
// InvCDF[k] = InvCDF[floor(1024 u)] with cum_trans[l] >= u.  It lies at or
// before InvCDF[k + 1]'s cell, whose prefix reaches (k + 1) / 1024 > u (the
// InvCDF build compares cum_trans * 1024, exact in float, with k + 1), and
// the prefixes rise monotonically, so a binary search over that bracket finds
// the scan's cell with log2 of its reads: buckets over a long tail of cells
// of small weight would make the scan long.  *scanned counts the cum_trans
// entries read.
__device__ __forceinline__ int select_dest(const Dev& d, size_t rc, double u, unsigned* scanned = nullptr)
{
    const int k = (int)__builtin_floor(u * 1024);
    const int32_t* inv = d.invcdf + rc * 1025;
    const int32_t v = inv[k];
    int lo = (int)((uint32_t)v & ~CS_INV_PURE);
    unsigned reads = 1;
    if (v >= 0) {  // not a pure bucket
        const float* ct = d.cum_trans + rc * d.NC;
        int hi = (int)((uint32_t)inv[k + 1] & ~CS_INV_PURE);
        reads = 0;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            reads++;
            if ((double)ct[mid] >= u) hi = mid;
            else lo = mid + 1;
        }
    }
    if (scanned) *scanned += reads;
    return lo;
}

// row j of realisation r with checkpointed tables: update_row's sums (the
// same float terms added in cell order), max_trans evaluated term by term;
// the running prefix is kept only at the last cell of every block of ck
// cells (the last block's entry is the row total), so a destination search
// can redo any block's prefixes exactly from the entry before it
__device__ inline void update_row_ck(const Dev& d, int r, int j)
{
    const size_t i = (size_t)r * d.NC + j;
    const int NC = d.NC, B = d.ck;
    const int32_t* S0 = d.S0 + (size_t)r * NC;
    float* pre = d.ck_pre + i * d.ck_nb;
    float c = 0.f, t = 0.f;  // 0 + x == x: the first term is the reference's ct[0] and t
    for (int m0 = 0, b = 0; m0 < NC; m0 += B, b++) {
        const int m1 = min(NC, m0 + B);
        int m = m0;
        // eight terms evaluated together (independent loads and lookups in
        // flight: a row rebuild has few rows per SIMD to hide latency with),
        // added in cell order
        for (; m + 8 <= m1; m += 8) {
            float mt[8];
#pragma unroll
            for (int q = 0; q < 8; q++) mt[q] = max_trans_of<true>(d, j, m + q);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                c = c + ((float)S0[m + q]) * mt[q];
                t = t + ((float)d.cn[m + q]) * mt[q];
            }
        }
        for (; m < m1; m++) {
            const float mt = max_trans_of<true>(d, j, m);
            c = c + ((float)S0[m]) * mt;
            t = t + ((float)d.cn[m]) * mt;
        }
        pre[b] = c;
    }
    d.tot_prob[i] = c / t;
}

// row i's bucket index (ck_inv, the brackets of select_dest_ck), by one
// whole wave after the row's checkpoints are written: inv[k] = the first
// block b whose normalised last prefix v_b = pre[b] / tp reaches k / 1024.
// Block b is that first block exactly for floor(1024 v_{b-1}) < k <=
// floor(1024 v_b) (k / 1024 and 1024 v are exact in double), so the lane
// reading block b writes those k; the last block's v is tp / tp = 1, so
// every k up to 1024 is written once.  A row of no weight (tp = 0: NaN
// ratios) gets the last block for every k, where the search without the
// index ends.
__device__ inline void ck_index_row(const Dev& d, size_t i)
{
    const int nb = d.ck_nb;
    const int lane = (int)(threadIdx.x & 63);
    const float* pre = d.ck_pre + i * nb;
    int32_t* inv = d.ck_inv + i * 1025;
    const float tp = pre[nb - 1];
    if (!((double)(tp / tp) >= 1.0)) {
        for (int k = lane; k <= 1024; k += 64) inv[k] = nb - 1;
        return;
    }
    for (int b = lane; b < nb; b += 64) {
        const int klo = b ? (int)__builtin_floor(1024.0 * (double)(pre[b - 1] / tp)) + 1 : 0;
        const int khi = min(1024, (int)__builtin_floor(1024.0 * (double)(pre[b] / tp)));
        for (int k = klo; k <= khi; k++) inv[k] = b;
    }
}

// destination cell with checkpointed tables: the first cell whose
// normalised prefix reaches u, which is where the InvCDF start and the
// cum_trans scan stop (src/Sweep.cpp:757-762: every cell before
// InvCDF[floor(1024 u)] has a prefix below floor(1024 u) / 1024 <= u).  A
// binary search over the blocks' last prefixes finds the block (within the
// bracket of bucket k = floor(1024 u) when the row's bucket index is kept:
// blocks before ck_inv[k] end below k / 1024 <= u, block ck_inv[k + 1] ends
// at or above (k + 1) / 1024 > u), whose prefixes are then redone from the
// entry before it in the order k_update_rows_ck summed them, so each is the
// flat table's value.  *mt is max_trans of (source, destination), the
// acceptance test's divisor.
__device__ __forceinline__ int select_dest_ck(const Dev& d, int r, int c, double u, float* mt, unsigned* scanned = nullptr)
{
    const int NC = d.NC, B = d.ck, nb = d.ck_nb;
    const size_t row = (size_t)r * NC + c;
    const float* pre = d.ck_pre + row * nb;
    const float tp = pre[nb - 1];
    int lo = 0, hi = nb - 1;  // the last block's prefix is tp: tp / tp = 1 > u
    if (d.ck_inv) {
        const int k = (int)__builtin_floor(u * 1024);
        lo = d.ck_inv[row * 1025 + k];
        hi = d.ck_inv[row * 1025 + k + 1];
    }
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((double)(pre[mid] / tp) >= u) hi = mid;
        else lo = mid + 1;
    }
    float s = lo ? pre[lo - 1] : 0.f;
    const int32_t* S0 = d.S0 + (size_t)r * NC;
    const int m1 = min(NC, lo * B + B) - 1;
    int m = lo * B;
    float w;
    for (;; m++) {
        w = max_trans_of(d, c, m);  // the distance out of line: measured faster in the draw kernel
        s = s + ((float)S0[m]) * w;
        if (m == m1 || (double)(s / tp) >= u) break;  // the block's last prefix reaches u
    }
    if (scanned) *scanned += (unsigned)(m - lo * B + 1);
    *mt = w;
    return m;
}

} // namespace cs
