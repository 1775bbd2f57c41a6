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
// from d.cxy
__device__ __forceinline__ float max_trans_of(const Dev& d, int l, int m)
{
    const int2 a = d.cxy[l], b = d.cxy[m];
    return (float)kernel_num(d, dist2_cc_min_xy(d.P, a.x, a.y, b.x, b.y));
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

// destination cell: InvCDF then linear scan of cum_trans (src/Sweep.cpp:757-762)
__device__ __forceinline__ int select_dest(const Dev& d, size_t rc, double u, unsigned* scanned = nullptr)
{
    const int32_t v = d.invcdf[rc * 1025 + (int)__builtin_floor(u * 1024)];
    int l = (int)((uint32_t)v & ~CS_INV_PURE);
    const int l0 = l;
    if (v >= 0) {  // not a pure bucket: the reference's scan
        const float* ct = d.cum_trans + rc * d.NC;
        while (ct[l] < u) l++;
    }
    // the reference reads cum_trans[l0 .. l] (one entry for a pure bucket)
    if (scanned) *scanned += (unsigned)(l - l0 + 1);
    return l;
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
        for (int m = m0; m < m1; m++) {
            const float mt = max_trans_of(d, j, m);
            c = c + ((float)S0[m]) * mt;
            t = t + ((float)d.cn[m]) * mt;
        }
        pre[b] = c;
    }
    d.tot_prob[i] = c / t;
}

// destination cell with checkpointed tables: the first cell whose
// normalised prefix reaches u, which is where the InvCDF start and the
// cum_trans scan stop (src/Sweep.cpp:757-762: every cell before
// InvCDF[floor(1024 u)] has a prefix below floor(1024 u) / 1024 <= u).  A
// binary search over the blocks' last prefixes finds the block, whose
// prefixes are then redone from the entry before it in the order
// k_update_rows_ck summed them, so each is the flat table's value.  *mt is
// max_trans of (source, destination), the acceptance test's divisor.
__device__ __forceinline__ int select_dest_ck(const Dev& d, int r, int c, double u, float* mt, unsigned* scanned = nullptr)
{
    const int NC = d.NC, B = d.ck, nb = d.ck_nb;
    const float* pre = d.ck_pre + ((size_t)r * NC + c) * nb;
    const float tp = pre[nb - 1];
    int lo = 0, hi = nb - 1;  // the last block's prefix is tp: tp / tp = 1 > u
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
        w = max_trans_of(d, c, m);
        s = s + ((float)S0[m]) * w;
        if (m == m1 || (double)(s / tp) >= u) break;  // the block's last prefix reaches u
    }
    if (scanned) *scanned += (unsigned)(m - lo * B + 1);
    *mt = w;
    return m;
}

} // namespace cs
