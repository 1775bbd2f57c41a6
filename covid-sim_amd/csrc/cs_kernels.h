// cs_kernels.h -- the device kernels of one infection step.
//
// One launch covers every realisation: grids are [realisation][cell] blocks
// or grid-stride loops over a device-side count, so a whole timestep needs
// no host round trip and can be captured in a HIP graph.
//
// Deterministic by construction (same result for any thread schedule):
//  * every random decision draws from its own counter-based stream;
//  * infection events carry key = infector<<32 | pass<<27 | sub; for each
//    infectee the smallest key performs DoInfect and the largest key's
//    infector/type remain recorded (src/Sweep.cpp:860-880 processes its
//    queues in order: first entry infects, later entries overwrite the
//    recorded infector);
//  * cell lists S|L|I|R+D are permuted by a batched move: the sorted
//    positions of the leaving batch below hi-k are paired in order with the
//    sorted free positions in [hi-k, hi).
#pragma once
#include "cs_rows.h"

namespace cs {

#define CS_EV_FULL 0xffffffffu
// bound on the spatial rejection loop (src/Sweep.cpp:753-854 has none)
#define CS_MAX_ATTEMPTS (1 << 20)
// spatial draws per work-list chunk (a wave's 64 lanes share them)
#define CS_DRAW_CHUNK 64
// InvCDF entry flag: the bucket's draws all stop at the entry's cell

__device__ __forceinline__ uint64_t ev_key(int infector, int pass, uint32_t sub)
{
    return ((uint64_t)(uint32_t)infector << 32) | ((uint64_t)pass << 27) | (uint64_t)(sub & 0x7ffffff);
}

__device__ __forceinline__ void add_event(const Dev& d, uint64_t key, size_t gp_infectee, int infector, int type)
{
    const unsigned sh = blockIdx.x & (CS_EV_SHARDS - 1u), scap = d.ev_cap / CS_EV_SHARDS;
    const unsigned idx = atomicAdd(&d.ev_count[sh * CS_EV_STRIDE], 1u);
    if (idx < scap) {
        const size_t e = (size_t)sh * scap + idx;
        d.ev_key[e] = key;
        d.ev_gp[e] = (long long)gp_infectee;
        d.ev_infector[e] = infector;
        d.ev_type[e] = (int16_t)type;
    } else {
        atomicOr(d.status, (unsigned)FLAG_OVERFLOW);
    }
}

// halt: set when a step overflowed a work buffer; every later kernel of the
// step (and of steps already queued behind it) then does nothing, so the
// state stays that of the step's start and the step can be re-run
__device__ __forceinline__ bool halted(const Dev& d) { return (*(volatile unsigned*)d.status & FLAG_OVERFLOW) != 0; }

__device__ __forceinline__ void cadd(unsigned long long* a, long long v) { atomicAdd(a, (unsigned long long)v); }

// cadd from every active lane, merged: the lanes that add to one address
// make one atomic between them (a wave's lanes share their realisation's
// counters, so most of a wave's adds land on a few addresses).  v is the
// same in every lane; integer sums, so every total equals that of one add
// per lane.  Every active lane stays in the loop until the wave-uniform set
// of lanes still to add is empty: each pass takes the lowest such lane's
// address (a shuffle from that lane) and removes every lane with the same
// address, the lead included.  The adds do not return, so passes for many
// distinct addresses cost no memory round trips.
__device__ __forceinline__ void cadd_merged(unsigned long long* a, long long v)
{
    const uint64_t me = (uint64_t)(uintptr_t)a;
    const unsigned lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    for (unsigned long long todo = __ballot(1); todo;) {
        const int lead = __builtin_ctzll(todo);
        const uint64_t la = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(me >> 32), lead, 64) << 32) | (uint32_t)__shfl((int)(uint32_t)me, lead, 64);
        const unsigned long long same = __ballot(me == la);
        if (lane == (unsigned)lead) atomicAdd(a, (unsigned long long)(v * (long long)__popcll(same)));
        todo &= ~same;
    }
}

// lists p for its cell's batched move.  One returning atomic per lane: a
// per-wave merge of lanes of many cells (events, due lists) serialises the
// atomics' latency, and where a wave holds one cell the compiler already
// merges the lanes' adds into one.  The move kernel finds the cells with a
// pending move from the counts themselves.
__device__ __forceinline__ void batch_add(const Dev& d, int r, int c, int p)
{
    const int rc = r * d.NC + c;
    const int idx = atomicAdd(&d.bcount[rc], 1);
    d.batch[(size_t)r * d.N + d.cstart[c] + idx] = p;
}

// A work-list kernel gives each wave whole chunks, so the chunk and every
// value derived from it (realisation, cell, stream seed, slot base) is the
// same in all lanes.  readfirstlane states that to the compiler, which then
// keeps those values in scalar registers and reads them with scalar loads.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned uni(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ uint64_t uni(uint64_t v)
{
    return ((uint64_t)uni((unsigned)(v >> 32)) << 32) | uni((unsigned)v);
}
__device__ __forceinline__ unsigned grid_wave() { return uni((blockIdx.x * blockDim.x + threadIdx.x) >> 6); }
__device__ __forceinline__ unsigned grid_nwaves() { return (gridDim.x * blockDim.x) >> 6; }
__device__ __forceinline__ Chunk chunk_at(const Chunk* l, unsigned ch)
{
    const Chunk e = l[ch];
    return Chunk{uni(e.rc), uni(e.first)};
}

// this step's latent persons are all on its due list (not overflowed)
__device__ __forceinline__ bool due_listed(const Dev& d)
{
    return d.due_on && d.due_ovf[d.sc->ts & (CS_DUE_RING - 1u)] == 0u;
}

__device__ __forceinline__ unsigned* due_count(const Dev& d, unsigned b, unsigned sh) { return d.due_n + ((size_t)b * CS_DUE_SHARDS + sh) * CS_DUE_STRIDE; }

// a returning atomicAdd(a, 1) from every lane with on set, merged: the lanes
// that add to one address make one atomic between them (one instruction for
// all the wave's addresses) and take consecutive results in lane order.
// Every lane of the wave must reach the call.
__device__ __forceinline__ unsigned slot_merged(bool on, unsigned* a)
{
    const uint64_t me = (uint64_t)(uintptr_t)a;
    const unsigned lane = threadIdx.x & 63u;
    unsigned long long mine = 0;
    for (unsigned long long todo = __ballot(on); todo;) {
        const int lead = __builtin_ctzll(todo);
        const uint64_t la = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(me >> 32), lead, 64) << 32) | (uint32_t)__shfl((int)(uint32_t)me, lead, 64);
        const unsigned long long same = __ballot(on && me == la);
        if (on && me == la) mine = same;
        todo &= ~same;
    }
    const int leader = on ? __builtin_ctzll(mine) : (int)lane;
    unsigned base = 0;
    if (on && (int)lane == leader) base = atomicAdd(a, (unsigned)__popcll(mine));
    base = (unsigned)__shfl((int)base, leader, 64);
    return base + (unsigned)__popcll(mine & ((1ull << lane) - 1ull));
}

// this block's staged copy of realisation r's counters
__device__ __forceinline__ Counters* cstage(const Dev& d, int r)
{
    return d.cstage + (size_t)(blockIdx.x & (CS_TALLY_COPIES - 1u)) * d.R + r;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
    for (int w = 32; w > 0; w >>= 1) v += __shfl_xor(v, w, 64);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v)
{
    for (int w = 32; w > 0; w >>= 1) v += __shfl_xor(v, w, 64);
    return v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
    for (int w = 32; w > 0; w >>= 1) {
        const unsigned long long o = __shfl_xor(v, w, 64);
        v = o > v ? o : v;
    }
    return v;
}

// exclusive scan over a 256-thread block; *total gets the block sum
__device__ __forceinline__ unsigned block_excl_scan(unsigned v, unsigned* total)
{
    __shared__ unsigned wsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned x = v;
    for (int o = 1; o < 64; o <<= 1) {
        unsigned y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    unsigned before = 0, all = 0;
    for (int k = 0; k < 4; k++) {
        if (k < w) before += wsum[k];
        all += wsum[k];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// Work lists of the step's phases, one thread per (realisation, cell):
//   seg 0: the latent segment [S, S+L) in chunks of 64 (also tallies L+I)
//   seg 1: the infectious segment [S+L, S+L+I) in chunks, and the list of
//          cells with I > 0; with zero_draws, every cell's draw count is cleared
//   seg 2: the spatial draws [0, spK) in chunks of CS_DRAW_CHUNK
// One atomic per block reserves the block's slots; the order of the lists
// is irrelevant (every consumer's result is independent of it).
__global__ void __launch_bounds__(256) k_compact(Dev d, int seg, int zero_draws)
{
    const long long nrc = (long long)d.R * d.NC;
    const long long rc = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int n = 0;
    unsigned long long act = 0;
    if (rc < nrc) {
        if (seg == 0) {
            act = (unsigned long long)d.cL[rc] + d.cI[rc];
            // the latent chunks, unless this step's due list serves instead
            n = due_listed(d) ? 0 : d.cL[rc];
        }
        else if (seg == 1) n = d.cI[rc];
        else n = d.spK[rc];
        if (zero_draws) d.spK[rc] = 0;
    }
    const unsigned per = (seg == 2) ? CS_DRAW_CHUNK : 64u;  // entries per chunk
    const unsigned nch = (unsigned)(((unsigned)n + per - 1) / per);
    unsigned tot_ch, tot_cells;
    const unsigned ex_ch = block_excl_scan(nch, &tot_ch);
    const unsigned ex_cell = (seg == 1) ? block_excl_scan(n > 0 ? 1u : 0u, &tot_cells) : 0u;
    __shared__ unsigned base_ch, base_cell;
    if (threadIdx.x == 0) {
        base_ch = tot_ch ? atomicAdd(d.wl + (seg == 2 ? WL_DRAWS : WL_SEG), tot_ch) : 0u;
        base_cell = (seg == 1 && tot_cells) ? atomicAdd(d.wl + WL_CELLS, tot_cells) : 0u;
    }
    __syncthreads();
    Chunk* out = (seg == 2) ? d.drawl : d.segl;
    const unsigned cap = (seg == 2) ? d.draw_cap : d.seg_cap;
    for (unsigned k = 0; k < nch; k++) {
        const unsigned idx = base_ch + ex_ch + k;
        if (idx < cap) out[idx] = Chunk{(int32_t)rc, (int32_t)(k * per)};
        else atomicOr(d.status, (unsigned)FLAG_OVERFLOW);
    }
    if (seg == 1 && n > 0) d.cells[base_cell + ex_cell] = (int32_t)rc;
    if (seg == 0) {
        act = wave_sum(act);
        if ((threadIdx.x & 63) == 0 && act) wtally(d.wc->n_active, act);
    }
}

// no claim yet: the smallest key starts at the largest value, the largest at 0
__global__ void k_claim_init(ClaimRec* c, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        c[i] = ClaimRec{~0ull, 0ull};
}

__global__ void k_zero_counts(unsigned* p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// this step's scalars (src/Sweep.cpp:294-301) from t, or from the device
// clock when the step is part of a replayed graph
__global__ void k_step_begin(Dev d, double t, int from_clock)
{
    if (threadIdx.x != 0 || halted(d)) return;
    const cs_params* P = d.P;
    if (from_clock) t = d.clock->t;
    StepScalars s;
    s.t = t;
    s.ts = (unsigned)(uint16_t)(P->time_steps_per_day * t);
    s.fp = P->model_time_step / (1 - P->false_positive_rate);
    s.seas = P->do_seasonality ? P->seasonality[((int)t) % CS_DAYS_PER_YEAR] : 1.0;
    s.beta_s = s.seas * s.fp * P->local_beta;
    s.beta_h = P->do_households ? s.seas * s.fp * P->household_trans : 0;
    *d.sc = s;
}

// the clock after a completed step (not after a halted one)
__global__ void k_step_end(Dev d, int sample_done)
{
    if (threadIdx.x != 0 || halted(d)) return;
    d.clock->t += d.P->model_time_step;
    d.clock->step++;
    if (sample_done) d.clock->sample++;
}

// ---------------------------------------------------------------- tables

// the whole table, into out ([NC][NC])
__global__ void k_build_max_trans(Dev d, float* out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)d.NC * d.NC;
    if (i >= n) return;
    int l = (int)(i / d.NC), m = (int)(i % d.NC);
    out[i] = max_trans_of<true>(d, l, m);
}

// UpdateProbs(0), S0 part (src/CovidSim.cpp:2756-2768); only realisations
// with sel[r] != 0 (all when sel == nullptr)
__global__ void k_update_s0(Dev d, const int* sel)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)d.R * d.NC) return;
    int r = (int)(i / d.NC), j = (int)(i % d.NC);
    if (sel && !sel[r]) return;
    d.tot_prob[i] = 0;
    int s0 = d.cS[i] + d.cL[i] + d.cI[i];
    if (d.P->do_death) {
        s0 += d.cn[j] / 5;
        if ((d.cn[j] < 100) || (s0 > d.cn[j])) s0 = d.cn[j];
    }
    d.S0[i] = s0;
}

// UpdateProbs(0), per-row sequential float prefix (src/CovidSim.cpp:2780-2800),
// a wave per (realisation, source cell) row (update_row_wave)
__global__ void __launch_bounds__(256) k_update_rows_wave(Dev d, const int* sel)
{
    const size_t rows = (size_t)d.R * d.NC, waves = ((size_t)gridDim.x * blockDim.x) >> 6;
    for (size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < rows; i += waves) {
        const int r = (int)(i / d.NC), j = (int)(i % d.NC);
        if (!sel || sel[r]) update_row_wave(d, r, j);
    }
}

// the same with checkpointed tables
__global__ void k_update_rows_ck(Dev d, const int* sel)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)d.R * d.NC) return;
    int r = (int)(i / d.NC), j = (int)(i % d.NC);
    if (sel && !sel[r]) return;
    update_row_ck(d, r, j);
}

// the rows' bucket index after k_update_rows_ck (ck_index_row): a wave per
// (realisation, source cell)
__global__ void __launch_bounds__(256) k_ck_index(Dev d, const int* sel)
{
    const size_t rows = (size_t)d.R * d.NC, waves = ((size_t)gridDim.x * blockDim.x) >> 6;
    for (size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < rows; i += waves)
        if (!sel || sel[i / d.NC]) ck_index_row(d, i);
}

// cum_trans, InvCDF and max_trans of realisation r expanded from the
// checkpoints (each block's prefixes redone from the entry before it, as the
// destination search redoes them), for cs_get_cell_tables
__global__ void k_expand_rows(Dev d, int r, float* mt_out, float* ct_out, int32_t* inv_out)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.NC) return;
    const int NC = d.NC, B = d.ck, nb = d.ck_nb;
    const size_t rc = (size_t)r * NC + j;
    const float* pre = d.ck_pre + rc * nb;
    const int32_t* S0 = d.S0 + (size_t)r * NC;
    const float tp = pre[nb - 1];
    float* ct = ct_out + (size_t)j * NC;
    for (int b = 0; b < nb; b++) {
        float c = b ? pre[b - 1] : 0.f;
        for (int m = b * B; m < min(NC, b * B + B); m++) {
            const float mt = max_trans_of(d, j, m);
            c = c + ((float)S0[m]) * mt;
            ct[m] = c / tp;
            if (mt_out) mt_out[(size_t)j * NC + m] = mt;
        }
    }
    int32_t* inv = inv_out + (size_t)j * 1025;
    int m = 0;
    for (int k = 0; k <= 1024; k++) {
        while (m < NC - 1 && ct[m] * 1024 < ((float)k)) m++;  // the last prefix is 1: the bound never binds
        inv[k] = m;
    }
}

// RunModel's "S fell by more than 0.2" check (src/CovidSim.cpp:1515-1521).
// It reads State.S, which only RecordSample refreshes (:3163): the S of
// the last sample, not the current one.
__global__ void k_sdrop(Dev d, int* sel)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d.R) return;
    if (halted(d)) { sel[r] = 0; return; }
    double prop = d.samp_prop[r];
    int up = (d.prev_prop[r] - prop) > 0.2;
    if (up) d.prev_prop[r] = prop;
    sel[r] = up;
}

// ------------------------------------------------------- list movement

#define CS_MOVE_CAP 4096       // moves one cell's LDS arrays hold (k_list_move)
#define CS_MOVE_BIG_BLOCKS 64  // blocks of k_list_move_big, each with its own bitmap

// Move each cell's batch to the tail of its segment.  mode 0: S->L,
// 1: L->I, 2: I->R/D.  One block per (realisation, cell).  The batch is
// paired like this: holes = sorted batch positions below bnd = hi - k,
// free = sorted positions in [bnd, hi) not holding a batch member; the
// j-th hole swaps with the j-th free position.
// one cell's batched move (see the file comment); every thread of the block calls it
__device__ void list_move_cell(const Dev& d, int mode, int rc, int32_t* pos, int32_t* fpos, uint8_t* occ, int32_t* part,
                               int* dead_count_p, int* bad_p)
{
    int& dead_count = *dead_count_p;
    int& bad = *bad_p;
    const int r = rc / d.NC, c = rc % d.NC;
    const int k = d.bcount[rc];
    if (k > d.move_cap) {  // more than the LDS arrays take: k_list_move_big moves this cell
        if (threadIdx.x == 0) d.mv_over[atomicAdd(d.mv_over_n, 1u)] = rc;
        return;
    }
    const size_t rN = (size_t)r * d.N;
    const size_t base = rN + d.cstart[c];
    const int S = d.cS[rc], L = d.cL[rc], I = d.cI[rc];
    const int hi = (mode == 0) ? S : (mode == 1 ? S + L : S + L + I);
    const int bnd = hi - k;
    if (threadIdx.x == 0) { dead_count = 0; bad = 0; }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        pos[i] = d.listpos[rN + d.batch[base + i]];
        occ[i] = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        if (pos[i] < 0 || pos[i] >= hi) { bad = 1; continue; }
        if (pos[i] >= bnd) occ[pos[i] - bnd] = 1;
    }
    __syncthreads();
    if (bad) {  // a list position outside the segment: report, touch nothing
        if (threadIdx.x == 0) atomicOr(d.status, (unsigned)FLAG_BATCH);
        return;
    }
    // free positions of the window in ascending order (chunked exclusive scan)
    const int per = (k + blockDim.x - 1) / blockDim.x;
    const int a0 = min(k, (int)threadIdx.x * per), a1 = min(k, a0 + per);
    int cnt = 0;
    for (int i = a0; i < a1; i++) cnt += 1 - occ[i];
    // exclusive scan of the threads' counts: within each wave by shuffles,
    // then the totals of the waves before (the same prefixes as a serial pass)
    {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        int incl = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (lane == 63) part[w] = incl;
        __syncthreads();
        int run = incl - cnt;
        for (int q = 0; q < w; q++) run += part[q];
        for (int i = a0; i < a1; i++)
            if (!occ[i]) fpos[run++] = bnd + i;
    }
    __syncthreads();
    // holes: batch positions below bnd, paired with the free slot of equal rank
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        int ph = pos[i];
        if (ph >= bnd) {  // already in the window
            if (mode == 1) d.shot[base + ph] = d.hot[rN + d.batch[base + i]];  // its slot copy, as it joins the infectious
            continue;
        }
        int rank = 0;
        for (int j = 0; j < k; j++) rank += (pos[j] < ph);
        int q = fpos[rank];
        int pa = d.perm[base + ph], pb = d.perm[base + q];
        d.perm[base + ph] = pb;
        d.listpos[rN + pb] = ph;
        d.perm[base + q] = pa;
        d.listpos[rN + pa] = q;
        if (mode == 1) d.shot[base + q] = d.hot[rN + pa];
        if (mode == 2) {  // an infectious person stays: its gate and slot copy move with it
            d.gate[base + ph] = d.gate[base + q];
            d.shot[base + ph] = d.shot[base + q];
        }
    }
    if (mode == 1)  // [bnd, hi) now heads the infectious list: its new timers are all to be looked at
        for (int i = threadIdx.x; i < k; i += blockDim.x) d.gate[base + bnd + i] = 0;
    if (mode == 2) {
        int nd = 0;
        for (int i = threadIdx.x; i < k; i += blockDim.x) nd += is_dead_code(d.inf[rN + d.batch[base + i]]);
        if (nd) atomicAdd(&dead_count, nd);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (mode == 0) { d.cS[rc] = S - k; d.cL[rc] = L + k; }
        else if (mode == 1) { d.cL[rc] = L - k; d.cI[rc] = I + k; }
        else { d.cI[rc] = I - k; d.cD[rc] += dead_count; d.cR[rc] += k - dead_count; }
        d.bcount[rc] = 0;
    }
}

// the same move by one wave for a cell with 0 < k <= 64 moves: lane i holds
// batch entry i; ballots and shuffles stand in for the LDS arrays and the
// barriers, and the pairing rule is the same (the j-th hole by position takes
// the j-th free window position in ascending order)
__device__ void list_move_cell_wave(const Dev& d, int mode, int rc, int k)
{
    const int lane = threadIdx.x & 63;
    const int r = rc / d.NC, c = rc % d.NC;
    const size_t rN = (size_t)r * d.N;
    const size_t base = rN + d.cstart[c];
    const int S = d.cS[rc], L = d.cL[rc], I = d.cI[rc];
    const int hi = (mode == 0) ? S : (mode == 1 ? S + L : S + L + I);
    const int bnd = hi - k;
    const bool mine = lane < k;
    const int pos = mine ? d.listpos[rN + d.batch[base + lane]] : 0;
    if (__ballot(mine && (pos < 0 || pos >= hi))) {  // a list position outside the segment: report, touch nothing
        if (lane == 0) atomicOr(d.status, (unsigned)FLAG_BATCH);
        return;
    }
    // window positions [bnd, hi) taken by batch entries (bit j: bnd + j)
    unsigned long long occ = 0;
    for (unsigned long long m = __ballot(mine && pos >= bnd); m; m &= m - 1)
        occ |= 1ull << (__shfl(pos, __builtin_ctzll(m), 64) - bnd);
    const unsigned long long freem = ~occ & ((k == 64) ? ~0ull : ((1ull << k) - 1ull));
    // holes: batch entries below the window; rank among them by position
    const bool hole = mine && pos < bnd;
    int rank = 0;
    for (unsigned long long m = __ballot(hole); m; m &= m - 1)
        rank += (__shfl(pos, __builtin_ctzll(m), 64) < pos) ? 1 : 0;
    if (hole) {
        unsigned long long f = freem;
        for (int t = 0; t < rank; t++) f &= f - 1;
        if (f == 0) {
            atomicOr(d.status, (unsigned)FLAG_BATCH);  // holes and free positions disagree: cannot happen for a valid batch
        } else {
            const int ph = pos, q = bnd + __builtin_ctzll(f);
            const int pa = d.perm[base + ph], pb = d.perm[base + q];
            d.perm[base + ph] = pb;
            d.listpos[rN + pb] = ph;
            d.perm[base + q] = pa;
            d.listpos[rN + pa] = q;
            if (mode == 1) d.shot[base + q] = d.hot[rN + pa];  // its slot copy, as it joins the infectious
            if (mode == 2) {  // an infectious person stays: its gate and slot copy move with it
                d.gate[base + ph] = d.gate[base + q];
                d.shot[base + ph] = d.shot[base + q];
            }
        }
    }
    if (mode == 1 && mine && pos >= bnd) d.shot[base + pos] = d.hot[rN + d.batch[base + lane]];  // already in the window
    if (mode == 1 && mine) d.gate[base + bnd + lane] = 0;  // [bnd, hi) now heads the infectious list
    const unsigned long long nd = wave_sum((unsigned long long)((mode == 2 && mine) ? is_dead_code(d.inf[rN + d.batch[base + lane]]) : 0));
    if (lane == 0) {
        if (mode == 0) { d.cS[rc] = S - k; d.cL[rc] = L + k; }
        else if (mode == 1) { d.cL[rc] = L - k; d.cI[rc] = I + k; }
        else { d.cI[rc] = I - k; d.cD[rc] += (int)nd; d.cR[rc] += k - (int)nd; }
        d.bcount[rc] = 0;
    }
}

__global__ void __launch_bounds__(256) k_list_move(Dev d, int mode)
{
    if (halted(d)) return;
    __shared__ int32_t pos[CS_MOVE_CAP];
    __shared__ int32_t fpos[CS_MOVE_CAP];
    __shared__ uint8_t occ[CS_MOVE_CAP];
    __shared__ int32_t part[256 / 64];  // per-wave totals of the free-slot scan (blocks of 256)
    __shared__ int dead_count, bad;
    // each block owns a contiguous range of (realisation, cell) pairs and
    // reads their pending-move counts 64 at a time (every wave the same 64,
    // so the ballots are the same in all of them; the barrier after the read
    // keeps any wave from clearing a count another has yet to read).  Cells
    // with at most 64 moves go to the block's waves in turn, a wave each;
    // larger ones take the whole block.
    const unsigned rcn = (unsigned)d.R * (unsigned)d.NC;
    const unsigned per = (rcn + gridDim.x - 1) / gridDim.x;
    const unsigned c0 = min(rcn, blockIdx.x * per), c1 = min(rcn, c0 + per);
    const int w = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
    for (unsigned b = c0; b < c1; b += 64) {
        const unsigned rc = b + (threadIdx.x & 63);
        const int kk = (rc < c1) ? d.bcount[rc] : 0;
        __syncthreads();
        int t = 0;
        for (unsigned long long m = __ballot(kk > 0 && kk <= 64 && kk <= d.move_cap); m; m &= m - 1, t++) {
            if (t % nw != w) continue;
            const int j = __builtin_ctzll(m);
            list_move_cell_wave(d, mode, (int)(b + j), __shfl(kk, j, 64));
        }
        for (unsigned long long m = __ballot(kk > 64 || kk > d.move_cap); m; m &= m - 1) {
            __syncthreads();  // the previous cell's use of the shared arrays is over
            list_move_cell(d, mode, (int)(b + __builtin_ctzll(m)), pos, fpos, occ, part, &dead_count, &bad);
        }
    }
}

// one cell's batched move in global memory, for a cell with more moves than
// the LDS arrays hold; the pairing rule is list_move_cell's (the j-th batch
// position below the window, in ascending order, swaps with the j-th window
// position not holding a batch member).  The batch's segment positions are
// marked in a bitmap of the segment (bm, W words) whose per-word exclusive
// bit counts (pre) give each hole its rank, set bits below it, and each free
// window position its rank, found by binary search
__device__ void list_move_cell_big(const Dev& d, int mode, int rc, uint32_t* bm, int32_t* pre, int32_t* part,
                                   int* dead_count_p, int* bad_p)
{
    int& dead_count = *dead_count_p;
    int& bad = *bad_p;
    const int r = rc / d.NC, c = rc % d.NC;
    const int k = d.bcount[rc];
    const size_t rN = (size_t)r * d.N;
    const size_t base = rN + d.cstart[c];
    const int S = d.cS[rc], L = d.cL[rc], I = d.cI[rc];
    const int hi = (mode == 0) ? S : (mode == 1 ? S + L : S + L + I);
    const int bnd = hi - k;
    const int W = (hi + 31) >> 5;
    if (threadIdx.x == 0) { dead_count = 0; bad = (W > d.mv_words || bnd < 0) ? 1 : 0; }
    __syncthreads();
    if (bad) {  // a segment longer than the cell (cannot happen for a valid state)
        if (threadIdx.x == 0) atomicOr(d.status, (unsigned)FLAG_BATCH);
        return;
    }
    for (int w = threadIdx.x; w < W; w += blockDim.x) bm[w] = 0u;
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        const int p = d.listpos[rN + d.batch[base + i]];
        if (p < 0 || p >= hi) { bad = 1; continue; }
        atomicOr(&bm[p >> 5], 1u << (p & 31));
    }
    __syncthreads();
    if (bad) {  // a list position outside the segment: report, touch nothing
        if (threadIdx.x == 0) atomicOr(d.status, (unsigned)FLAG_BATCH);
        return;
    }
    {   // pre[w] = set bits in words below w (chunked exclusive scan, as list_move_cell's)
        const int per = (W + blockDim.x - 1) / blockDim.x;
        const int a0 = min(W, (int)threadIdx.x * per), a1 = min(W, a0 + per);
        int cnt = 0;
        for (int w = a0; w < a1; w++) cnt += __popc(bm[w]);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        int incl = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (lane == 63) part[wv] = incl;
        __syncthreads();
        int run = incl - cnt;
        for (int q = 0; q < wv; q++) run += part[q];
        for (int w = a0; w < a1; w++) { pre[w] = run; run += __popc(bm[w]); }
    }
    __syncthreads();
    // set bits below position x (x < hi)
    auto below = [&](int x) { return pre[x >> 5] + __popc(bm[x >> 5] & ((1u << (x & 31)) - 1u)); };
    const int sb = below(bnd);
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        const int ph = d.listpos[rN + d.batch[base + i]];  // written only by this thread below
        if (ph >= bnd) {  // already in the window
            if (mode == 1) d.shot[base + ph] = d.hot[rN + d.batch[base + i]];  // its slot copy, as it joins the infectious
            continue;
        }
        const int j = below(ph);  // every batch position below ph is a hole: ph's rank among the holes
        // the j-th free window position: the largest q in [bnd, hi) with
        // g(q) = (q - bnd) - (below(q) - sb) <= j, g counting the free
        // positions in [bnd, q)
        int lo = bnd, up = hi - 1;
        while (lo < up) {
            const int mid = (lo + up + 1) >> 1;
            if ((mid - bnd) - (below(mid) - sb) <= j) lo = mid;
            else up = mid - 1;
        }
        const int q = lo;
        const int pa = d.perm[base + ph], pb = d.perm[base + q];
        d.perm[base + ph] = pb;
        d.listpos[rN + pb] = ph;
        d.perm[base + q] = pa;
        d.listpos[rN + pa] = q;
        if (mode == 1) d.shot[base + q] = d.hot[rN + pa];  // its slot copy, as it joins the infectious
        if (mode == 2) {  // an infectious person stays: its gate and slot copy move with it
            d.gate[base + ph] = d.gate[base + q];
            d.shot[base + ph] = d.shot[base + q];
        }
    }
    if (mode == 1)  // [bnd, hi) now heads the infectious list: its new timers are all to be looked at
        for (int i = threadIdx.x; i < k; i += blockDim.x) d.gate[base + bnd + i] = 0;
    if (mode == 2) {
        int nd = 0;
        for (int i = threadIdx.x; i < k; i += blockDim.x) nd += is_dead_code(d.inf[rN + d.batch[base + i]]);
        if (nd) atomicAdd(&dead_count, nd);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (mode == 0) { d.cS[rc] = S - k; d.cL[rc] = L + k; }
        else if (mode == 1) { d.cL[rc] = L - k; d.cI[rc] = I + k; }
        else { d.cI[rc] = I - k; d.cD[rc] += dead_count; d.cR[rc] += k - dead_count; }
        d.bcount[rc] = 0;
    }
}

// the cells k_list_move queued (more moves than its LDS arrays hold), a
// block per cell in turn, each block with its own bitmap; the last block to
// finish empties the queue for the next launch
__global__ void __launch_bounds__(256) k_list_move_big(Dev d, int mode)
{
    __shared__ int32_t part[256 / 64];
    __shared__ int dead_count, bad;
    const unsigned n = halted(d) ? 0u : d.mv_over_n[0];
    uint32_t* bm = d.mv_bm + (size_t)blockIdx.x * 2 * d.mv_words;
    int32_t* pre = (int32_t*)(bm + d.mv_words);
    for (unsigned it = blockIdx.x; it < n; it += gridDim.x) {
        list_move_cell_big(d, mode, d.mv_over[it], bm, pre, part, &dead_count, &bad);
        __syncthreads();  // the bitmap and the shared words are free again
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(d.mv_over_n + 1, 1u) == gridDim.x - 1) {  // every block has read the length
            d.mv_over_n[0] = 0u;
            d.mv_over_n[1] = 0u;
        }
    }
}

// ------------------------------------------------------------ DoInfect

// DoInfect (src/Update.cpp:100-195) for infectee p of realisation r, with
// the first event's infector and type; the list swap is batched.
// the counter updates of one DoInfect (StateT fields, src/Update.cpp:100-195)
struct InfectTally {
    int kind, ag, kw;        // cumItype, cumIa, cumI_keyworker indices
    unsigned lat;            // latent time (the due list p joins)
    int c;                   // p's cell (the batch p joins)
    double rad2;
    long long tg, si;        // generation and serial intervals
    int ntg;                 // 1 when the infector is a person
};

// DoInfect (src/Update.cpp:100-195) without the counter updates, which it
// returns in *tl; false when p is not susceptible (no change)
__device__ bool do_infect_state(const Dev& d, int r, int p, double t, unsigned ts, int infector, int type, InfectTally* tl)
{
    const cs_params* P = d.P;
    size_t gp = (size_t)r * d.N + p;
    if (d.inf[gp] != CS_SUSCEPTIBLE) return false;
    d.inf[gp] = CS_LATENT;
    d.t_inf[gp] = (uint16_t)ts;
    const float2 loc = d.ploc[p];
    double x = (loc.x - d.loc_init[2 * r]);
    double y = (loc.y - d.loc_init[2 * r + 1]);
    tl->rad2 = x * x + y * y;
    tl->kind = type % CS_INFECT_TYPE_MASK;
    tl->ag = age_group(d, p);
    tl->kw = keyw(d, p);
    uint16_t lat;
    if (P->do_latent) {
        Stream s(d.seed[r], cw0(ts, PS_INFECT), (uint32_t)p, 0);
        double q;
        int i = (int)__builtin_floor((q = s.next() * CS_CDF_RES));
        q -= ((double)i);
        const double* c = P->icdf[CS_ICDF_LATENT];
        lat = (uint16_t)__builtin_floor(0.5 + (t - P->latent_period * clog(q * c[i + 1] + (1.0 - q) * c[i])) * P->time_steps_per_day);
    } else
        lat = (uint16_t)(t * P->time_steps_per_day);
    d.t_lat[gp] = lat;
    tl->lat = lat;
    tl->c = d.pc[p];
    tl->tg = tl->si = 0;
    tl->ntg = 0;
    if (infector >= 0) {
        size_t gi = (size_t)r * d.N + infector;
        tl->tg = ((int)ts) - ((int)d.t_inf[gi]);
        tl->si = ((int)lat) - ((int)d.t_lat[gi]);
        tl->ntg = 1;
    }
    return true;
}

// the indexed counters of a tally (per lane)
__device__ __forceinline__ void tally_indexed(Counters* cn, const InfectTally& tl)
{
    cadd(&cn->cumItype[tl.kind], 1);
    cadd(&cn->cumIa[tl.ag], 1);
    cadd(&cn->cumI_keyworker[tl.kw], 1);
}

__device__ __forceinline__ void tally_scalar(Counters* cn, unsigned long long n, double rad2, unsigned long long rad2max_bits,
                                             long long tg, long long si, unsigned long long ntg)
{
    cadd(&cn->cumI, (long long)n);
    atomicAdd(&cn->sumRad2, rad2);
    atomicMax(&cn->maxRad2_bits, rad2max_bits);
    if (ntg) {
        cadd(&cn->cumTG, tg);
        cadd(&cn->cumSI, si);
        cadd(&cn->nTG, (long long)ntg);
    }
}

// gp at index idx of its shard of the due list of the step that ends its
// latency; a full shard marks the bucket overflowed
__device__ __forceinline__ void due_put(const Dev& d, unsigned b, unsigned sh, unsigned idx, long long gp)
{
    if (idx < d.due_scap) d.due[((size_t)b * CS_DUE_SHARDS + sh) * d.due_scap + idx] = gp;
    else if (d.due_ovf[b] == 0u) atomicOr(&d.due_ovf[b], 1u);
}

__device__ void do_infect(const Dev& d, int r, int p, double t, unsigned ts, int infector, int type)
{
    InfectTally tl;
    if (!do_infect_state(d, r, p, t, ts, infector, type, &tl)) return;
    batch_add(d, r, tl.c, p);
    if (d.due_on) {
        const unsigned b = tl.lat & (CS_DUE_RING - 1u), sh = blockIdx.x & (CS_DUE_SHARDS - 1u);
        due_put(d, b, sh, atomicAdd(due_count(d, b, sh), 1u), (long long)r * d.N + p);
    }
    Counters* cn = d.cnt + r;
    tally_indexed(cn, tl);
    tally_scalar(cn, 1, tl.rad2, double_to_bits(tl.rad2), tl.tg, tl.si, (unsigned long long)tl.ntg);
    if (d.ad_cnt) atomicAdd(ad_counter(d, r, p, CS_TSU_INC_I), 1ull);  // cumI_adunit (src/Update.cpp:159)
}

// ---------------------------------------------------------- InfectSweep



// spatial beta of realisation r (src/Sweep.cpp:300)
__device__ __forceinline__ double beta_s_of(const Dev& d, const StepScalars& sc, int r)
{
    return d.rbeta ? sc.seas * sc.fp * d.rbeta[r] : sc.beta_s;
}

// the reference's infector binary search (src/Sweep.cpp:716-740)
__device__ __forceinline__ int select_infector(const float* ci, int n, double u, unsigned* probes = nullptr)
{
    if (n == 1) return 0;
    int idx, step;
    idx = step = n / 2;
    int keep = 1;
    unsigned np = 0;
    do {
        np++;
        if (step > 1) step /= 2;
        if ((idx > 0) && (__builtin_fabs((double)ci[idx - 1]) >= u)) {
            idx -= step;
            if (idx == 0) keep = 0;
        } else if ((idx < n - 1) && (__builtin_fabs((double)ci[idx]) < u)) {
            idx += step;
            if (idx == n - 1) keep = 0;
        } else keep = 0;
    } while (keep);
    if (probes) *probes += np;
    return idx;
}

// SampleWithoutReplacement replacement: all / Floyd (k <= 32) / selection
#define CS_FLOYD_MAX 32

// Household pass and spatial infectiousness of one infectious person per
// lane (src/Sweep.cpp:314-377, :649-690); a wave takes a chunk of up to 64
// infectious persons of one cell.  The place passes are a kernel of their
// own so that each keeps few registers (many waves in flight hide the
// latency of the scattered person reads).
// 8 waves per SIMD (78 SGPRs, no spills): more scattered reads in flight
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_infect_households(Dev d)
{
    const StepScalars sc = *d.sc;
    const unsigned nch = min(d.wl[WL_SEG], d.seg_cap);
    const cs_params* P = d.P;
    const unsigned ts = sc.ts;
    const int lane = threadIdx.x & 63;
    unsigned long long w_hh_inf = 0, w_hh_mem = 0, w_hh_scan = 0;
    for (unsigned ch = grid_wave(); ch < nch; ch += grid_nwaves()) {
        const Chunk e = chunk_at(d.segl, ch);
        const int rc = e.rc;
        const int j = e.first + lane;
        if (j >= uni(d.cI[rc])) continue;
        const int r = rc / d.NC, c = rc % d.NC;
        const uint64_t seed = uni(d.seed[r]);
        const size_t rN = (size_t)r * d.N;
        const size_t sbase = rN + uni(d.cstart[c] + d.cS[rc] + d.cL[rc]);  // slot of infectious index 0
        const int i = d.perm[sbase + j];
        const Hot h = d.shot[sbase + j];  // i's record, by slot: one coalesced load
        const int gen = h.itype / CS_INFECT_TYPE_MASK;
        // household (src/Sweep.cpp:331-377)
        if (sc.beta_h > 0) {
            const unsigned fl = h.flags;
            const int nh = (int)(fl >> HOT_NH_SHIFT);  // the household's size, from i's own record
            if (nh > 1) {
                w_hh_inf++;
                w_hh_mem += nh;
                // the susceptible members first: with none, nothing below can
                // infect, so the household's hazard is not needed; and none
                // will be susceptible again, so later steps skip the scan
                // (and read no household record)
                unsigned smask = 0;
                int f0 = 0;
                if (!(fl & HOT_HH_DONE)) {
                    w_hh_scan += nh;
                    f0 = i - (int)((h.pflags >> PF_HPOS_SHIFT) & PF_HPOS_MASK);  // the household's first member
                    for (int m = f0; m < f0 + nh; m++)
                        if (d.inf[rN + m] == CS_SUSCEPTIBLE) smask |= 1u << (m - f0);
                    if (!smask) d.shot[sbase + j].flags = (uint8_t)(fl | HOT_HH_DONE);
                }
                if (smask) {
                    double HI = sc.beta_h * house_inf_t(d, r, i, HotRec{h}, ts);
                    bool absent = false;
                    // a member absent from a closed place (src/Sweep.cpp:348-355); the
                    // member's absence is tested first (cheaper, same result)
                    for (int m = f0; m < f0 + nh && !absent; m++) {
                        if (!host_absent(d, rN + m, ts)) continue;
                        for (int pt = 0; pt < d.NPT && !absent; pt++) {
                            int lk = d.plinks[(size_t)m * CS_MAX_PLACE_TYPES + pt];
                            if (lk >= 0) absent = place_closed(d, r, pt, lk, ts);
                        }
                    }
                    if (absent) HI *= P->efficacies[2][CS_HOUSE];
                    for (unsigned sm = smask; sm; sm &= sm - 1) {
                        const int m = f0 + __builtin_ctz(sm);
                        double foi = HI * house_susc_ag(d, r, m, ts, (int)h.age / CS_AGE_GROUP_WIDTH);
                        Stream s(seed, cw0(ts, PS_HH), (uint32_t)i, (uint32_t)(m - f0));
                        if (s.next() < foi)
                            add_event(d, ev_key(i, KP_HH, (uint32_t)(m - f0)), rN + m, i, 1 + CS_INFECT_TYPE_MASK * (1 + gen));
                    }
                }
            }
        }
        // spatial infectiousness (src/Sweep.cpp:649-690); the closed flag is the sign
        if (beta_s_of(d, sc, r) > 0) {
            double SI = spatial_inf_t(d, r, i, HotRec{h}, ts);
            bool flag = false;  // a linked place is closed and i is absent (tested first: same result)
            if (P->do_places && h.t_as <= ts && h.t_ae > ts)
                for (int pt = 0; pt < d.NPT && !flag; pt++) {
                    int lk = d.plinks[(size_t)i * CS_MAX_PLACE_TYPES + pt];
                    if (lk >= 0) flag = place_closed(d, r, pt, lk, ts);
                }
            if (flag) {
                SI *= P->efficacies[2][CS_SPATIAL];
                d.si_buf[sbase + j] = -SI;
            } else
                d.si_buf[sbase + j] = SI;
        }
        }
    w_hh_inf = wave_sum(w_hh_inf);
    w_hh_mem = wave_sum(w_hh_mem);
    w_hh_scan = wave_sum(w_hh_scan);
    if ((threadIdx.x & 63) == 0 && (w_hh_inf | w_hh_mem)) {
        wtally(d.wc->n_hh_inf, w_hh_inf);
        wtally(d.wc->n_hh_member, w_hh_mem);
        wtally(d.wc->n_hh_scan, w_hh_scan);
    }
}

// the position of the q-th set bit of m (q < popcount(m))
__device__ __forceinline__ int nth_set(unsigned long long m, int q)
{
    int pos = 0, c;
    c = __popc((unsigned)m);
    if (q >= c) { q -= c; m >>= 32; pos += 32; }
    c = __popc((unsigned)(m & 0xffffu));
    if (q >= c) { q -= c; m >>= 16; pos += 16; }
    c = __popc((unsigned)(m & 0xffu));
    if (q >= c) { q -= c; m >>= 8; pos += 8; }
    c = __popc((unsigned)(m & 0xfu));
    if (q >= c) { q -= c; m >>= 4; pos += 4; }
    c = __popc((unsigned)(m & 0x3u));
    if (q >= c) { q -= c; m >>= 2; pos += 2; }
    if (q >= (int)(m & 1u)) pos += 1;
    return pos;
}

// the stream of (a, b, c) from its value number pos on (two values per
// Philox block)
__device__ __forceinline__ Stream stream_at(uint64_t key, uint32_t a, uint32_t b, uint32_t c, unsigned pos)
{
    Stream t(key, a, b, c);
    t.c3 = pos >> 1;
    if (pos & 1u) (void)t.next();
    return t;
}

// Within-group and between-group place passes (src/Sweep.cpp:381-640) in two
// kernels.  k_infect_places takes one infectious person per lane: for each
// place link and pass, the capped infectiousness and the binomial number of
// contacts; the rare links with contacts go to the place contact list with
// the stream position after the count.  k_place_contacts takes one contact
// list entry per lane: the members sampled (Floyd's sample in LDS, one
// column per thread, or selection sampling) and their infection tests.  The
// first kernel carries none of the second's state, so it runs with neither
// LDS nor spills.
__global__ void __launch_bounds__(256) k_infect_places(Dev d)
{
    const StepScalars sc = *d.sc;
    const unsigned nch = min(d.wl[WL_SEG], d.seg_cap);
    const cs_params* P = d.P;
    const unsigned ts = sc.ts;
    unsigned w_link = 0;  // per lane; summed per wave in 64 bits
    if (!P->do_places) return;
    const unsigned sh = blockIdx.x & (CS_EV_SHARDS - 1u), scap = d.pc_cap / CS_EV_SHARDS;
    for (unsigned ch = grid_wave(); ch < nch; ch += grid_nwaves()) {
        const Chunk e = chunk_at(d.segl, ch);
        const int rc = e.rc;
        const int j = e.first + (int)(threadIdx.x & 63);
        if (j >= uni(d.cI[rc])) continue;
        const int r = rc / d.NC, c = rc % d.NC;
        const uint64_t seed = uni(d.seed[r]);
        const size_t rN = (size_t)r * d.N;
        const size_t sbase = rN + uni(d.cstart[c] + d.cS[rc] + d.cL[rc]);  // slot of infectious index 0
        const int i = d.perm[sbase + j];
        const Hot h = d.shot[sbase + j];  // i's record, by slot: one coalesced load
        if ((h.flags & HOT_NO_LINKS) || (h.t_as <= ts && h.t_ae > ts)) continue;
        for (int pt = 0; pt < d.NPT; pt++) {
            const int lk = d.plinks[(size_t)i * CS_MAX_PLACE_TYPES + pt];
            if (lk < 0) continue;
            w_link++;
            double PI = sc.fp * sc.seas * place_inf_t(d, r, i, HotRec{h}, pt, ts);
            const int mstart = d.pl_ms[pt][lk];
            for (int pass = 0; pass < 2; pass++) {
                // Pc is both the capped place infectiousness and the
                // sampling probability: the reference keeps two copies that
                // differ only under digital contact tracing
                // (src/Sweep.cpp:425-447), which this path does not serve
                int gsize, goff;
                double Pc;
                if (pass == 0) {
                    if (pt == P->hotel_place_type) continue;
                    const int g2 = d.pglinks[(size_t)i * CS_MAX_PLACE_TYPES + pt];
                    Pc = PI;
                    if (Pc > 1) Pc = 1;
                    const int gix = d.pl_gs0[pt][lk] + g2;
                    goff = d.pl_gst[pt][gix];
                    gsize = d.pl_gsz[pt][gix];
                } else {
                    PI *= P->place_type_prop_between_group_links[pt] * P->place_type_group_size_param1[pt] / ((double)d.pl_n[pt][lk]);
                    if (PI > 1) PI = 1;
                    Pc = PI;
                    goff = 0;
                    gsize = d.pl_n[pt][lk];
                }
                Stream s(seed, cw0(ts, pass == 0 ? PS_PW_SAMPLE : PS_PB_SAMPLE), (uint32_t)i, (uint32_t)pt);
                const int n = (Pc >= 1) ? gsize : binomial(s, gsize, Pc);
                if (n <= 0) continue;
                // values used: two per Philox block, less the buffered one
                const unsigned v = 2u * s.c3 - (s.idx == 1 ? 1u : 0u);
                const unsigned idx = atomicAdd(&d.pc_count[sh * CS_EV_STRIDE], 1u);
                if (idx < scap)
                    d.pcl[(size_t)sh * scap + idx] = PlContact{r, i, n, gsize, goff, mstart, Pc, v, (uint8_t)pt, (uint8_t)pass, 0};
                else
                    atomicOr(d.status, (unsigned)FLAG_OVERFLOW);
            }
        }
    }
    const unsigned long long s_link = wave_sum(w_link);
    if ((threadIdx.x & 63) == 0 && s_link) wtally(d.wc->n_place_link, s_link);
}

// a sharded list as one sequence: pre[s] entries lie in the shards before s
// (the counts capped at the shard size, as the appends are); constant
// indices only, so pre stays in registers
struct ShardSeq {
    unsigned pre[CS_EV_SHARDS + 1], scap;
    __device__ __forceinline__ ShardSeq(const unsigned* count, unsigned cap) : scap(cap / CS_EV_SHARDS)
    {
        pre[0] = 0;
        for (int s = 0; s < CS_EV_SHARDS; s++) pre[s + 1] = pre[s] + min(count[s * CS_EV_STRIDE], scap);
    }
    __device__ __forceinline__ unsigned n() const { return pre[CS_EV_SHARDS]; }
    __device__ __forceinline__ size_t slot(unsigned k) const
    {
        unsigned sh = 0, base = 0;
#pragma unroll
        for (int j = 1; j < CS_EV_SHARDS; j++)
            if (k >= pre[j]) { sh = (unsigned)j; base = pre[j]; }
        return (size_t)sh * scap + (k - base);
    }
};

// the members sampled for one place contact list entry per lane and their
// infection tests (src/Sweep.cpp:449-640); the pass's stream resumes where
// the count left it
__global__ void __launch_bounds__(256) k_place_contacts(Dev d)
{
    if (halted(d)) return;
    const StepScalars sc = *d.sc;
    __shared__ int picked[CS_FLOYD_MAX][256];
    const cs_params* P = d.P;
    const unsigned ts = sc.ts;
    const ShardSeq q(d.pc_count, d.pc_cap);
    const unsigned nq = q.n();
    unsigned w_cand = 0;
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < nq; k += gridDim.x * blockDim.x) {
        const PlContact w = d.pcl[q.slot(k)];
        const int r = w.r, i = w.i, pt = w.pt, pass = w.pass, gsize = w.gsize, goff = w.goff, mstart = w.mstart;
        int n = w.n;
        const double Pc = w.Pc;
        const size_t rN = (size_t)r * d.N, gi = rN + i;
        const uint64_t seed = d.seed[r];
        const int gen = d.itype[gi] / CS_INFECT_TYPE_MASK;
        Stream s = stream_at(seed, cw0(ts, pass == 0 ? PS_PW_SAMPLE : PS_PB_SAMPLE), (uint32_t)i, (uint32_t)pt, w.v);
        const bool all = n >= gsize;
        const bool floyd = !all && n <= CS_FLOYD_MAX;
        if (floyd) {
            int mm = 0;
            for (int j2 = gsize - n; j2 < gsize; j2++) {
                int tt = (int)(s.next() * (double)(j2 + 1));
                bool dup = false;
                for (int q2 = 0; q2 < mm; q2++) dup |= (picked[q2][threadIdx.x] == tt);
                picked[mm++][threadIdx.x] = dup ? j2 : tt;
            }
        }
        if (all) n = gsize;
        int sel = 0;
        for (int cidx = 0, scan = 0; cidx < n; cidx++) {
            int idx;
            if (all) idx = cidx;
            else if (floyd) idx = picked[cidx][threadIdx.x];
            else {
                // selection sampling over [scan, gsize)
                for (;;) {
                    double u = s.next();
                    int cur = scan++;
                    if ((double)(gsize - cur) * u < (double)(n - sel)) { idx = cur; sel++; break; }
                }
            }
            const int m = d.pl_mem[pt][mstart + goff + idx];
            w_cand++;
            double PS = place_susc(d, r, m, pt, ts);
            if (pass == 1 && care_g(d, gi) && care_g(d, rN + m) &&
                d.pglinks[(size_t)i * CS_MAX_PLACE_TYPES + pt] != d.pglinks[(size_t)m * CS_MAX_PLACE_TYPES + pt])
                PS *= P->care_home_resident_place_scaling;
            if ((pt == P->care_home_place_type) && ((!care_g(d, gi)) || (!care_g(d, rN + m)))) PS *= P->care_home_worker_group_scaling;
            const size_t gm = rN + m;
            if (d.inf[gm] == CS_SUSCEPTIBLE && !host_absent(d, gm, ts)) {
                PS *= person_susc(d, r, m, ts, i) * Pc / Pc;  // (susc * Pc) / Pc, rounded as the reference does
                Stream s2(seed, cw0(ts, pass == 0 ? PS_PW_ACCEPT : PS_PB_ACCEPT), (uint32_t)i, ((uint32_t)pt << 24) | (uint32_t)cidx);
                if ((PS == 1) || (s2.next() < PS)) {
                    if (pass == 0)
                        add_event(d, ev_key(i, KP_PW + pt, (uint32_t)cidx), gm, i, 2 + pt + CS_INFECT_TYPE_MASK * (1 + gen));
                    else
                        add_event(d, ev_key(i, KP_PB + pt, (uint32_t)cidx), gm, i,
                                  2 + pt + CS_MAX_PLACE_TYPES + CS_INFECT_TYPE_MASK * (1 + gen));
                }
            }
        }
    }
    const unsigned long long s_cand = wave_sum(w_cand);
    if ((threadIdx.x & 63) == 0 && s_cand) wtally(d.wc->n_place_cand, s_cand);
}

// Per (realisation, cell): the cumulative spatial infectiousness by slot,
// signed by the closure flag and normalised when draws follow, and the
// Poisson number of spatial draws (src/Sweep.cpp:649-708).  The running
// sum is kept in exact 64.64 fixed point, so every prefix is the exact sum
// of its terms rounded once to double, whatever the order of the parallel
// scan; the reference instead adds in double one term at a time
// (src/Sweep.cpp:680-686), so a prefix here can differ from its in the last
// bit.  Each prefix is then stored as float, as the reference does.
#define CS_SUM_BLOCK 256

// one wave per cell with I > 0; the slots are read in rounds of 64, one per
// lane (coalesced).  Pass 1 sums the cell's total (the draw count needs it),
// pass 2 scans each round across the wave with the carry of the rounds
// before: 128-bit integer adds are exact and associative, so every prefix
// is the same for any order of the adds.
__global__ void __launch_bounds__(CS_SUM_BLOCK) k_cell_spatial_sum(Dev d)
{
    const StepScalars sc = *d.sc;
    const unsigned ncells = d.wl[WL_CELLS];
    const int lane = threadIdx.x & 63;
    // the infectious persons seen, summed per wave, then per block: one
    // global add per block (an add per cell to the one work counter
    // serialises at the memory side)
    __shared__ unsigned long long blk_inf;
    if (threadIdx.x == 0) blk_inf = 0;
    __syncthreads();
    unsigned long long w_inf = 0;
    for (unsigned a = grid_wave(); a < ncells; a += grid_nwaves()) {
        const int rc = uni(d.cells[a]);
        const int I = uni(d.cI[rc]);
        const int r = rc / d.NC, c = rc % d.NC;
        const double beta_s = beta_s_of(d, sc, r);
        w_inf += (unsigned long long)I;
        if (!(beta_s > 0)) continue;
        const size_t sbase = (size_t)r * d.N + uni(d.cstart[c] + d.cS[rc] + d.cL[rc]);
        const double* si = d.si_buf + sbase;
        Fx acc = {0, 0};
        for (int j = lane; j < I; j += 64) acc = fx_add(acc, fx_from(fabs(si[j])));
        for (int o = 32; o > 0; o >>= 1) acc = fx_add(acc, Fx{__shfl_xor(acc.hi, o, 64), __shfl_xor(acc.lo, o, 64)});
        const double sis = fx_to_double(acc);
        int K = 0;
        if (lane == 0) {
            if (sis > 0) {
                Stream s(d.seed[r], cw0(sc.ts, PS_SP_COUNT), (uint32_t)c, 0);
                K = poisson(s, sis * beta_s * ((double)d.tot_prob[rc]));
            }
            d.spK[rc] = K;
        }
        K = __shfl(K, 0, 64);
        const float fs = (float)sis;
        float* ci_out = d.ci_buf + sbase;
        Fx carry = {0, 0};
        for (int b = 0; b < I; b += 64) {
            const int j = b + lane;
            const double v = (j < I) ? si[j] : 0.0;
            Fx x = fx_from(fabs(v));
            for (int o = 1; o < 64; o <<= 1) {
                const Fx y = {__shfl_up(x.hi, o, 64), __shfl_up(x.lo, o, 64)};
                if (lane >= o) x = fx_add(x, y);
            }
            if (j < I) {
                const double pv = fx_to_double(fx_add(carry, x));
                float ci = __builtin_signbit(v) ? (float)-pv : (float)pv;
                if (K > 0) {
                    if (j < I - 1) ci /= fs;
                    else ci = (ci < 0) ? -1.0f : 1.0f;
                }
                ci_out[j] = ci;
            }
            carry = fx_add(carry, Fx{__shfl(x.hi, 63, 64), __shfl(x.lo, 63, 64)});
        }
    }
    if (lane == 0 && w_inf) atomicAdd(&blk_inf, w_inf);
    __syncthreads();
    if (threadIdx.x == 0 && blk_inf) wtally(d.wc->n_inf, blk_inf);
}

// the spatial draws of one source cell per chunk of up to CS_DRAW_CHUNK
// draws, a wave per chunk (src/Sweep.cpp:712-855).  A draw takes its
// infector from value 0 of its stream (no value when the cell has one
// infectious person: v0 = 0, else 1) and makes attempts until one is
// accepted; a rejected attempt always uses exactly three values (the
// destination, the slot, the acceptance test), so attempt j reads values
// v0 + 3j .. v0 + 3j + 2 whatever came before, and the draw's outcome is its
// lowest-numbered accepted attempt.  Each round the wave's draws in progress
// share its lanes, G = 64 / (draws in progress) lanes each, lane g of a
// draw's group making attempt (attempts so far) + g: a long draw's chain of
// dependent gathers is cut G-fold once most of the chunk's draws are over.
// The group's first accepting lane finishes the draw (the susceptibility
// test reads value v0 + 3j + 3).  The work tallies count the attempts the
// sequential loop makes (those up to the accepted one).
// at most 96 VGPRs (5 waves per SIMD, no spills): each attempt is a chain of
// dependent gathers, hidden by the waves in flight and by the groups
// CK: checkpointed destination tables (max_trans evaluated where read)
template <bool CK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_spatial_draws(Dev d)
{
    const StepScalars sc = *d.sc;
    const unsigned nch = min(d.wl[WL_DRAWS], d.draw_cap);
    const cs_params* P = d.P;
    const unsigned ts = sc.ts;
    const int lane = threadIdx.x & 63;
    unsigned w_draw = 0, w_susc = 0, w_scan = 0, w_probe = 0;  // per lane; summed per wave in 64 bits
    for (unsigned ch = grid_wave(); ch < nch; ch += grid_nwaves()) {
        const Chunk e = chunk_at(d.drawl, ch);
        const size_t rc = (size_t)e.rc;
        const int kend = min(uni(d.spK[rc]), e.first + CS_DRAW_CHUNK);  // this chunk's draws: [first, kend)
        const int r = e.rc / d.NC, c = e.rc % d.NC;
        const int S = uni(d.cS[rc]), L = uni(d.cL[rc]), I = uni(d.cI[rc]);
        const size_t rN = (size_t)r * d.N;
        const size_t sbase = rN + uni(d.cstart[c]) + S + L;
        const uint64_t seed = uni(d.seed[r]);
        const uint32_t sw0 = cw0(ts, PS_SP_DRAW);
        const unsigned v0 = (I == 1) ? 0u : 1u;  // the first value of attempt 0
        // the draw this lane owns (>= kend: none); a lane whose draw is over
        // takes the chunk's next unassigned draw (next is wave-uniform)
        int k = e.first + lane;
        int next = e.first + 64;
        bool fresh = true;
        int i = 0, hh_i = 0, att = 0, ag_i = 0, fl_i = 0;  // fl_i: 1 closed, 2 care home
        float2 loc_i = make_float2(0.f, 0.f);
        for (;;) {
            const bool active = k < kend;
            const unsigned long long act = __ballot(active);
            if (act == 0) break;
            if (active && fresh) {  // the draw's infector (src/Sweep.cpp:716-742)
                Stream s0(seed, sw0, (uint32_t)c, (uint32_t)k);
                const int jj = (I == 1) ? 0 : select_infector(d.ci_buf + sbase, I, s0.next(), &w_probe);
                fl_i = (d.ci_buf[sbase + jj] < 0) ? 1 : 0;
                i = d.perm[sbase + jj];
                const PStat pi = d.pst[i];
                loc_i = pi.loc;
                hh_i = pi.hh;
                ag_i = (int)pi.age / CS_AGE_GROUP_WIDTH;  // the infector's, kept for every attempt
                fl_i |= (pi.pflags & 1) ? 2 : 0;
                att = 0;
                fresh = false;
            }
            // draw q (the q-th lane of act) takes lanes [q*G, q*G + G)
            const int nact = __popcll(act);
            const int G = 64 / nact;
            const int q = lane / G;
            const bool inq = q < nact;
            const int own = inq ? nth_set(act, q) : lane;
            const int k_o = __shfl(k, own, 64), att_o = __shfl(att, own, 64), i_o = __shfl(i, own, 64);
            const int hh_o = __shfl(hh_i, own, 64), ag_o = __shfl(ag_i, own, 64), fl_o = __shfl(fl_i, own, 64);
            const float2 loc_o = make_float2(__shfl(loc_i.x, own, 64), __shfl(loc_i.y, own, 64));
            const int j = att_o + (lane - q * G);  // this lane's attempt of the draw
            const bool tries = inq && j < CS_MAX_ATTEMPTS;
            bool accepted = false;
            unsigned scanned = 0;
            Stream t(seed, sw0, (uint32_t)c, (uint32_t)k_o);
            int l = 0, slot = 0, m = 0;
            PStat pm{};
            int8_t inf_m = 0;
            if (tries) {  // attempt j (src/Sweep.cpp:753-788)
                t = stream_at(seed, sw0, (uint32_t)c, (uint32_t)k_o, v0 + 3u * (unsigned)j);
                float mt;
                if constexpr (CK) {
                    l = select_dest_ck(d, r, c, t.next(), &mt, &scanned);
                } else {
                    l = select_dest(d, rc, t.next(), &scanned);
                    mt = d.max_trans[(size_t)c * d.NC + l];  // issued with the slot's loads
                }
                const size_t rl = (size_t)r * d.NC + l;
                slot = (int)(t.next() * ((double)d.S0[rl]));
                m = d.perm[rN + d.cstart[l] + slot];
                pm = d.pst[m];
                // the candidate's status from its list segment: the cell's
                // lists are [S | L | I | R and D] and stay as they are through
                // the event phase, so only a slot past S + L + I (a recovered
                // or dead person) needs the person's own status byte
                const int seg = d.cS[rl] + d.cL[rl] + d.cI[rl];
                inf_m = (slot < d.cS[rl]) ? (int8_t)CS_SUSCEPTIBLE
                      : (slot < seg) ? (int8_t)CS_LATENT : d.inf[rN + m];
                const double d2 = dist2_ll(P, pm.loc, loc_o);
                const double acc = kernel_num(d, d2) / mt;
                accepted = !((t.next() >= acc) || is_dead_code(inf_m));
            }
            // the first accepting lane of each group
            const unsigned long long am = __ballot(accepted);
            const unsigned long long gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
            const unsigned long long ga = inq ? (am & (gmask << (q * G))) : 0ull;
            const int first = ga ? __builtin_ctzll(ga) : 64;
            if (tries && lane <= first) {
                w_draw++;
                w_scan += scanned;
            }
            if (accepted && lane == first && ((c != l) || (pm.hh != hh_o))) {  // src/Sweep.cpp:789-854
                const size_t rl = (size_t)r * d.NC + l;
                double SS = spatial_susc(d, r, m, ts);
                if ((pm.pflags & 1) || (fl_o & 2)) SS *= P->care_home_resident_spatial_scaling;
                if (slot < d.cS[rl]) {
                    w_susc++;
                    const int ag_m = (int)pm.age / CS_AGE_GROUP_WIDTH;
                    SS *= person_susc_ag(d, rN + m, ag_m, ag_o, ts);
                    SS *= P->waifw_spatial[ag_m][ag_o];
                    if (!(fl_o & 1) && host_absent(d, rN + m, ts)) {
                        bool f2 = false;
                        for (int pt = 0; pt < d.NPT && !f2; pt++) {
                            int lk = d.plinks[(size_t)m * CS_MAX_PLACE_TYPES + pt];
                            if (lk >= 0) f2 = place_closed(d, r, pt, lk, ts);
                        }
                        if (f2) SS *= P->efficacies[2][CS_SPATIAL];
                    }
                    if ((SS == 1) || (t.next() < SS))
                        if (inf_m == CS_SUSCEPTIBLE)
                            add_event(d, ev_key(i_o, KP_SP, (uint32_t)k_o), rN + m, i_o,
                                      2 + 2 * CS_MAX_PLACE_TYPES + CS_INFECT_TYPE_MASK * (1 + d.itype[rN + i_o] / CS_INFECT_TYPE_MASK));
                }
            }
            // the owners: a draw is over once its group accepted, or its
            // attempts ran out (no live candidate: report, never spin)
            bool done = false;
            if (active) {
                const int qm = __popcll(act & ((1ull << lane) - 1ull));
                if (am & (gmask << (qm * G))) done = true;
                else {
                    att += G;
                    if (att >= CS_MAX_ATTEMPTS) {
                        atomicOr(d.status, (unsigned)FLAG_STUCK);
                        done = true;
                    }
                }
            }
            // lanes without a draw take the chunk's next ones, in lane order
            const bool need = !active || done;
            const unsigned long long want = __ballot(need);
            const unsigned below = __builtin_amdgcn_mbcnt_hi((uint32_t)(want >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)want, 0u));
            if (need) {
                k = next + (int)below;  // >= kend: the lane stays idle
                fresh = true;
            }
            next += __popcll(want);
        }
    }
    const unsigned long long s_draw = wave_sum(w_draw), s_susc = wave_sum(w_susc);
    const unsigned long long s_scan = wave_sum(w_scan), s_probe = wave_sum(w_probe);
    if ((threadIdx.x & 63) == 0 && s_draw) {
        wtally(d.wc->n_sp_draw, s_draw);
        wtally(d.wc->n_sp_susc, s_susc);
        wtally(d.wc->n_sp_scan, s_scan);
        wtally(d.wc->n_sp_probe, s_probe);
    }
}

__global__ void k_resolve_claim(Dev d)
{
    if (halted(d)) return;
    const ShardSeq q(d.ev_count, d.ev_cap);
    const unsigned n = q.n();
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const size_t e = q.slot(k);
        size_t gp = (size_t)d.ev_gp[e];
        unsigned long long key = d.ev_key[e];
        atomicMin(&d.claim[gp], key);
        atomicMax(&d.lastk[gp], key);
    }
}

__global__ void k_resolve_apply(Dev d)
{
    if (halted(d)) return;
    const StepScalars sc = *d.sc;
    const ShardSeq q(d.ev_count, d.ev_cap);
    const unsigned n = q.n();
    const int lane = threadIdx.x & 63;
    unsigned long long nn = 0;
    // whole waves step through the events, so every lane takes part in the
    // wave reductions of the counters below
    for (unsigned base = grid_wave() * 64u; base < n; base += grid_nwaves() * 64u) {
        const unsigned k = base + lane;
        const size_t e = (k < n) ? q.slot(k) : 0;
        bool won = false;
        int r = 0, p = 0;
        InfectTally tl;
        tl.c = 0;
        if (k < n) {
            const size_t gp = (size_t)d.ev_gp[e];
            const unsigned long long key = d.ev_key[e];
            r = (int)(gp / d.N);
            p = (int)(gp % d.N);
            if (key == d.claim[gp]) {
                won = do_infect_state(d, r, p, sc.t, sc.ts, d.ev_infector[e], d.ev_type[e], &tl);
                nn++;
            }
            if (key == d.lastk[gp]) {
                d.infector[gp] = d.ev_infector[e];
                d.itype[gp] = d.ev_type[e];
            }
        }
        // the winners join their cells' batches and their latent times' due
        // lists, lanes of one cell (of one bucket) sharing one atomic
        {
            const int rcw = won ? r * d.NC + tl.c : 0;
            const unsigned bi = slot_merged(won, (unsigned*)&d.bcount[rcw]);
            if (won) d.batch[(size_t)r * d.N + d.cstart[tl.c] + bi] = p;
        }
        if (d.due_on) {
            const unsigned b = won ? (tl.lat & (CS_DUE_RING - 1u)) : 0u, sh = blockIdx.x & (CS_DUE_SHARDS - 1u);
            const unsigned di = slot_merged(won, due_count(d, b, sh));
            if (won) due_put(d, b, sh, di, (long long)r * d.N + p);
        }
        // counters, staged in the block's copy of each realisation's
        // Counters (k_cnt_fold adds them up): events are appended by whole producing
        // waves, so a wave's infections nearly always share one realisation;
        // then the scalar counters take one atomic per wave instead of one
        // per infection, and the indexed ones one per distinct index
        const unsigned long long wm = __ballot(won);
        if (wm == 0) continue;
        const int r0 = __shfl(r, __ffsll((long long)wm) - 1, 64);
        if (won) {
            Counters* t = cstage(d, r);
            cadd_merged(&t->cumItype[tl.kind], 1);
            cadd_merged(&t->cumIa[tl.ag], 1);
            cadd_merged(&t->cumI_keyworker[tl.kw], 1);
            if (d.ad_cnt) cadd_merged(ad_counter(d, r, p, CS_TSU_INC_I), 1);  // cumI_adunit (src/Update.cpp:159)
        }
        if (__ballot(won && r != r0) == 0) {
            const double rad2 = wave_sum_d(won ? tl.rad2 : 0.0);
            const unsigned long long mx = wave_max_u64(won ? double_to_bits(tl.rad2) : 0ull);
            const long long tg = (long long)wave_sum(won ? (unsigned long long)tl.tg : 0ull);
            const long long si = (long long)wave_sum(won ? (unsigned long long)tl.si : 0ull);
            const unsigned long long ntg = wave_sum(won ? (unsigned long long)tl.ntg : 0ull);
            if (lane == __ffsll((long long)wm) - 1) tally_scalar(cstage(d, r0), __popcll(wm), rad2, mx, tg, si, ntg);
        } else if (won) {
            tally_scalar(cstage(d, r), 1, tl.rad2, double_to_bits(tl.rad2), tl.tg, tl.si, (unsigned long long)tl.ntg);
        }
    }
    nn = wave_sum(nn);
    if (lane == 0 && nn) wtally(d.wc->n_new, nn);
}

// the staged counters into each realisation's Counters, clearing the stage:
// block (r, y) takes words [64y, 64y + 64) of realisation r's Counters, a
// thread per 8-byte word; the copies are read together, then summed in
// order (sums wrap as the atomics do; sumRad2 is a double sum, maxRad2_bits
// a maximum)
constexpr int CS_CNT_WORDS = (int)(sizeof(Counters) / 8);
__global__ void __launch_bounds__(64) k_cnt_fold(Dev d)
{
    if (halted(d)) return;
    const int r = blockIdx.x, f = blockIdx.y * 64 + threadIdx.x;
    if (r >= d.R || f >= CS_CNT_WORDS) return;
    constexpr int WSUM = (int)(offsetof(Counters, sumRad2) / 8), WMAX = (int)(offsetof(Counters, maxRad2_bits) / 8);
    unsigned long long v[CS_TALLY_COPIES];
#pragma unroll
    for (int c = 0; c < CS_TALLY_COPIES; c++) v[c] = ((const unsigned long long*)(d.cstage + (size_t)c * d.R + r))[f];
    unsigned long long s = 0, mx = 0, any = 0;
    double sd = 0.0;
#pragma unroll
    for (int c = 0; c < CS_TALLY_COPIES; c++) {
        any |= v[c];
        if (f == WSUM) sd += __longlong_as_double((long long)v[c]);
        else if (f == WMAX) mx = (v[c] > mx) ? v[c] : mx;
        else s += v[c];
    }
    if (any == 0ull) return;
#pragma unroll
    for (int c = 0; c < CS_TALLY_COPIES; c++)
        if (v[c]) ((unsigned long long*)(d.cstage + (size_t)c * d.R + r))[f] = 0ull;
    unsigned long long* dst = (unsigned long long*)(d.cnt + r);
    if (f == WSUM) ((double*)dst)[f] += sd;
    else if (f == WMAX) dst[f] = (mx > dst[f]) ? mx : dst[f];
    else dst[f] += s;
}

__global__ void k_resolve_reset(Dev d)
{
    if (halted(d)) return;
    const ShardSeq q(d.ev_count, d.ev_cap);
    const unsigned n = q.n();
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        size_t gp = (size_t)d.ev_gp[q.slot(k)];
        d.claim[gp] = ~0ull;
        d.lastk[gp] = 0ull;
    }
}


}  // namespace cs
