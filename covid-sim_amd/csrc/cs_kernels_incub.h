// cs_kernels_incub.h -- IncubRecoverySweep (src/Sweep.cpp:883-979) on the
// device, the realisation reset (InitModel / SeedInfection), daily
// recording, and single-function kernels used by the parity tests.
#pragma once
#include "cs_kernels.h"

namespace cs {

__device__ __forceinline__ void sev_to(Counters* c, int kind, int ag)
{
    cadd_merged(&c->sev_prev[kind], 1); cadd_merged(&c->sev_prev_age[kind][ag], 1);
    cadd_merged(&c->sev_cum[kind], 1); cadd_merged(&c->sev_cum_age[kind][ag], 1);
}
__device__ __forceinline__ void sev_from(Counters* c, int kind, int ag)
{
    cadd_merged(&c->sev_prev[kind], -1); cadd_merged(&c->sev_prev_age[kind][ag], -1);
}
__device__ __forceinline__ void sev_cum_only(Counters* c, int kind, int ag)
{
    cadd_merged(&c->sev_cum[kind], 1); cadd_merged(&c->sev_cum_age[kind][ag], 1);
}

// holidays (src/Sweep.cpp:889-918) for holiday h: one thread per
// (realisation, place); members' absence windows only widen and every place
// of the holiday writes the same bounds, so concurrent writes agree.
__device__ void holiday_places(const Dev& d, int h, double ht);

// the holidays that start within this timestep (src/Sweep.cpp:889-892)
__global__ void k_holidays(Dev d)
{
    if (halted(d)) return;
    const cs_params* P = d.P;
    if (!P->do_places) return;
    const double t = d.sc->t;
    for (int h = 0; h < P->num_holidays; h++) {
        const double ht = P->holiday_start[h] + P->holidays_start_day_simtime;
        if ((t + P->model_time_step >= ht) && (t < ht)) holiday_places(d, h, ht);
    }
}

__device__ void holiday_places(const Dev& d, int h, double ht)
{
    const cs_params* P = d.P;
    const int tot = d.np_off[CS_MAX_PLACE_TYPES];
    const unsigned n = (unsigned)d.R * (unsigned)tot;  // below 2^32 (checked at allocation)
    for (unsigned g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
        int r = (int)(g / (unsigned)tot), f = (int)(g - (unsigned)r * (unsigned)tot);
        int pt = 0;
        while (pt + 1 < CS_MAX_PLACE_TYPES && f >= d.np_off[pt + 1]) pt++;
        int k = f - d.np_off[pt];
        bool close = false;
        if (P->holiday_effect[pt] < 1) {
            if (P->holiday_effect[pt] == 0) close = true;
            else {
                Stream s(d.seed[r], cw0((unsigned)h, PS_HOLIDAY), (uint32_t)k, (uint32_t)pt);
                close = s.next() >= P->holiday_effect[pt];
            }
        }
        if (!close) continue;
        int hs = (int)(ht * P->time_steps_per_day);
        int he = (int)((ht + P->holiday_duration[h]) * P->time_steps_per_day);
        size_t pi = (size_t)r * tot + f;
        if (d.pl_cs[pi] > hs) d.pl_cs[pi] = (uint16_t)hs;
        if (d.pl_ce[pi] < he) d.pl_ce[pi] = (uint16_t)he;
        const size_t rN = (size_t)r * d.N;
        for (int q = d.pl_ms[pt][k]; q < d.pl_ms[pt][k + 1]; q++) {
            const int m = d.pl_mem[pt][q];
            size_t gm = rN + m;
            if (d.t_as[gm] > hs) d.t_as[gm] = (uint16_t)hs;
            if (d.t_ae[gm] < he) d.t_ae[gm] = (uint16_t)he;
            shot_sync(d, gm, m);
        }
    }
}

// DoIncub (src/Update.cpp:408-523) without the list swap
__device__ void do_incub(const Dev& d, int r, int p, unsigned ts)
{
    const cs_params* P = d.P;
    const size_t gp = (size_t)r * d.N + p;
    if (d.inf[gp] != CS_LATENT) return;
    int ag = age_group_g(d, gp);
    if (ag >= CS_NUM_AGE_GROUPS) ag = CS_NUM_AGE_GROUPS - 1;
    Stream s(d.seed[r], cw0(ts, PS_INCUB), (uint32_t)p, 0);
    const double tspd = P->time_steps_per_day;
    float infn = (float)P->age_infectiousness[ag];
    if (P->infectiousness_sd > 0) {
        double k = 1 / (P->infectiousness_sd * P->infectiousness_sd);
        infn *= (float)gamma(s, k, k);
    }
    double ps = P->proportion_symptomatic[ag] * (treated(d, gp, ts) ? (1 - P->treat_sympt_drop) : 1)
              * (vacced(d, gp, ts) ? (1 - P->vacc_sympt_drop) : 1);
    int8_t st;
    if (s.next() < ps) {
        st = CS_INF_ALMOST_SYMPT;
        infn *= (float)(-P->sympt_infectiousness);
    } else {
        st = CS_INF_ASYMPT;
        infn *= (float)P->asympt_infectiousness;
    }
    d.inf[gp] = st;
    d.infness[gp] = infn;
    const uint16_t lat = d.t_lat[gp];
    if (!P->do_severity || st == CS_INF_ASYMPT) {
        if (P->do_infectiousness_profile)
            d.t_rod[gp] = (uint16_t)(lat + (uint16_t)(P->infectious_period * tspd));
        else
            d.t_rod[gp] = (uint16_t)(lat + icdf_choose(P->icdf[CS_ICDF_INFECTIOUS], P->infectious_period, s.next(), tspd));
    } else {
        int ct = lat + ((int)(P->latent_to_sympt_delay / P->model_time_step));
        double x = s.next();
        int fin;  // ChooseFinalDiseaseSeverity (src/Update.cpp:1377-1389)
        if (x < P->prop_ili[ag]) fin = CS_SEV_ILI;
        else if (x < P->prop_ili[ag] + P->prop_sari[ag]) fin = CS_SEV_SARI;
        else if (x < P->prop_ili[ag] + P->prop_sari[ag] + P->prop_critical[ag]) fin = CS_SEV_CRITICAL;
        else fin = CS_SEV_MILD;
        uint8_t bits = d.bits[gp];
        // DecideIfPersonDies (src/Update.cpp:388-406)
        if (fin == CS_SEV_CRITICAL) { if (s.next() < P->cfr_critical[ag] * P->cfr_critical_scale) bits |= 1; }
        else if (fin == CS_SEV_SARI) { if (s.next() < P->cfr_sari[ag] * P->cfr_sari_scale) bits |= 1; }
        else if (fin == CS_SEV_ILI) { if (s.next() < P->cfr_ili[ag] * P->cfr_ili_scale) bits |= 1; }
        if (care_g(d, gp) && (fin == CS_SEV_CRITICAL || fin == CS_SEV_SARI) && (s.next() > P->care_home_rel_prob_hosp)) {
            bits |= 1;
            fin = CS_SEV_ILI;
        }
        d.sev_final[gp] = (uint8_t)fin;
        const bool die = bits & 1;
        d.bits[gp] = bits;
#define CH(tab, mean) icdf_choose(P->icdf[tab], P->mean_by_age[mean][ag], s.next(), tspd)
        if (fin == CS_SEV_MILD) {
            d.t_rod[gp] = (uint16_t)(ct + CH(CS_ICDF_MILD_RECOV, CS_MEAN_MILD_RECOV));
        } else if (fin == CS_SEV_CRITICAL) {
            uint16_t sari = (uint16_t)(ct + CH(CS_ICDF_ILI_SARI, CS_MEAN_ILI_SARI));
            uint16_t crit = (uint16_t)(sari + CH(CS_ICDF_SARI_CRIT, CS_MEAN_SARI_CRIT));
            d.t_sari[gp] = sari;
            d.t_crit[gp] = crit;
            if (die) {
                if (P->include_stepdown_to_death == 1) {
                    uint16_t step = (uint16_t)(crit + CH(CS_ICDF_CRIT_CRITRECOV, CS_MEAN_CRIT_CRITRECOV));
                    d.t_step[gp] = step;
                    d.t_rod[gp] = (uint16_t)(step + CH(CS_ICDF_STEPDOWN_DEATH, CS_MEAN_STEPDOWN_DEATH));
                } else
                    d.t_rod[gp] = (uint16_t)(crit + CH(CS_ICDF_CRIT_DEATH, CS_MEAN_CRIT_DEATH));
            } else {
                uint16_t step = (uint16_t)(crit + CH(CS_ICDF_CRIT_CRITRECOV, CS_MEAN_CRIT_CRITRECOV));
                d.t_step[gp] = step;
                d.t_rod[gp] = (uint16_t)(step + CH(CS_ICDF_CRITRECOV_RECOV, CS_MEAN_CRITRECOV_RECOV));
            }
        } else if (fin == CS_SEV_SARI) {
            uint16_t sari = (uint16_t)(ct + CH(CS_ICDF_ILI_SARI, CS_MEAN_ILI_SARI));
            d.t_sari[gp] = sari;
            if (die) d.t_rod[gp] = (uint16_t)(sari + CH(CS_ICDF_SARI_DEATH, CS_MEAN_SARI_DEATH));
            else d.t_rod[gp] = (uint16_t)(sari + CH(CS_ICDF_SARI_RECOV, CS_MEAN_SARI_RECOV));
        } else {
            if (die) d.t_rod[gp] = (uint16_t)(ct + CH(CS_ICDF_ILI_DEATH, CS_MEAN_ILI_DEATH));
            else d.t_rod[gp] = (uint16_t)(ct + CH(CS_ICDF_ILI_RECOV, CS_MEAN_ILI_RECOV));
        }
#undef CH
    }
    const double ctrl = control_prop(d, r);
    if (st == CS_INF_ALMOST_SYMPT && ((ctrl == 1) || (s.next() < ctrl))) {
        d.bits[gp] |= 2;
        d.t_det[gp] = (uint16_t)(ts + (uint16_t)(P->latent_to_sympt_delay * tspd));
    }
}

// latent lists: DoIncub where latent_time == now (src/Sweep.cpp:925-927); a
// wave takes a chunk of up to 64 latent persons of one cell
// the persons on this step's due list whose latency ends now (the same
// transitions as the scan below, visiting only them)
__global__ void __launch_bounds__(256) k_latent_due(Dev d)
{
    if (halted(d) || !due_listed(d)) return;
    const unsigned ts = d.sc->ts;
    const unsigned b = ts & (CS_DUE_RING - 1u);
    // the bucket's shards as one sequence: pre[s] persons lie in the shards
    // before s (each count is at most due_scap: no shard overflowed);
    // constant indices only, so pre stays in registers
    unsigned pre[CS_DUE_SHARDS + 1];
    pre[0] = 0;
#pragma unroll
    for (int s = 0; s < CS_DUE_SHARDS; s++) pre[s + 1] = pre[s] + uni(*due_count(d, b, (unsigned)s));
    const unsigned n = pre[CS_DUE_SHARDS];
    const long long* list = d.due + (size_t)b * CS_DUE_SHARDS * d.due_scap;
    unsigned long long nt = 0;
    {
        // whole waves step through the sequence, every lane reaching the merged batch append
        for (unsigned k0 = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; k0 < n; k0 += gridDim.x * blockDim.x) {
            const unsigned k = k0 + (threadIdx.x & 63u);
            bool go = false;
            int r = 0, p = 0, c = 0;
            if (k < n) {
                unsigned sh = 0, base = 0;
#pragma unroll
                for (int j = 1; j < CS_DUE_SHARDS; j++)
                    if (k >= pre[j]) { sh = (unsigned)j; base = pre[j]; }
                const long long gp = list[(size_t)sh * d.due_scap + (k - base)];
                if (d.inf[gp] == CS_LATENT && d.t_lat[gp] == ts) {
                    r = (int)(gp / d.N), p = (int)(gp % d.N), c = d.pc[p];
                    do_incub(d, r, p, ts);
                    go = true;
                    nt++;
                }
            }
            const unsigned bi = slot_merged(go, (unsigned*)&d.bcount[r * d.NC + c]);
            if (go) d.batch[(size_t)r * d.N + d.cstart[c] + bi] = p;
        }
    }
    nt = wave_sum(nt);
    if ((threadIdx.x & 63) == 0 && nt) wtally(d.wc->n_transition, nt);
}

// this step's due list is consumed: clear it for the step CS_DUE_RING later
__global__ void k_due_clear(Dev d)
{
    if (halted(d) || !d.due_on) return;
    const unsigned b = d.sc->ts & (CS_DUE_RING - 1u);
    if (threadIdx.x < CS_DUE_SHARDS) *due_count(d, b, threadIdx.x) = 0u;
    if (threadIdx.x == 0) d.due_ovf[b] = 0u;
}

__global__ void __launch_bounds__(256) k_latent_scan(Dev d)
{
    if (halted(d)) return;
    const unsigned ts = d.sc->ts;
    const unsigned nch = min(d.wl[WL_SEG], d.seg_cap);
    const int lane = threadIdx.x & 63;
    unsigned long long nt = 0;
    for (unsigned ch = grid_wave(); ch < nch; ch += grid_nwaves()) {
        const Chunk e = chunk_at(d.segl, ch);
        const int rc = e.rc;
        const int q = e.first + lane;
        if (q >= uni(d.cL[rc])) continue;
        const int r = rc / d.NC, c = rc % d.NC;
        const size_t rN = (size_t)r * d.N;
        const int p = d.perm[rN + uni(d.cstart[c] + d.cS[rc]) + q];
        if (ts == d.t_lat[rN + p]) {
            do_incub(d, r, p, ts);
            batch_add(d, r, c, p);
            nt++;
        }
    }
    nt = wave_sum(nt);
    if ((threadIdx.x & 63) == 0 && nt) wtally(d.wc->n_transition, nt);
}

// DoCase (src/Update.cpp:794-894) without interventions; returns 1 when the
// case is a child who needs an adult at home (applied after the phase).  sl
// is the person's list slot (its Hot record's slot copy is updated there)
__device__ int do_case(const Dev& d, int r, int p, unsigned ts, size_t sl)
{
    const cs_params* P = d.P;
    const size_t gp = (size_t)r * d.N + p;
    int req = 0;
    if (d.inf[gp] != CS_INF_ALMOST_SYMPT) return 0;
    int ag = age_group_g(d, gp);
    if (ag >= CS_NUM_AGE_GROUPS) ag = CS_NUM_AGE_GROUPS - 1;
    Counters* cn = cstage(d, r);
    d.inf[gp] = CS_CASE;
    const int adel = P->us_case_absenteeism_delay, adur = P->us_case_absenteeism_duration;
    if (host_absent(d, gp, ts)) {
        if ((int)d.t_ae[gp] < (int)ts + adel + adur) d.t_ae[gp] = (uint16_t)(ts + adel + adur);
    } else if (P->do_real_sympt_withdrawal && P->do_places) {
        d.t_as[gp] = CS_UNSET_TIME;
        Stream s(d.seed[r], cw0(ts, PS_CASE), (uint32_t)p, 0);
        for (int j = 0; j < d.NPT; j++) {
            if ((d.plinks[(size_t)p * CS_MAX_PLACE_TYPES + j] >= 0) && (j != P->hotel_place_type) && (!host_absent(d, gp, ts)) &&
                (P->sympt_place_type_withdrawal_prop[j] > 0)) {
                if ((!care_g(d, gp)) && ((P->sympt_place_type_withdrawal_prop[j] == 1) || (s.next() < P->sympt_place_type_withdrawal_prop[j]))) {
                    d.t_as[gp] = (uint16_t)(ts + adel);
                    d.t_ae[gp] = (uint16_t)(ts + adel + adur);
                    if ((!quarantined(d, gp, ts)) && (d.plinks[(size_t)p * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] >= 0) &&
                        ((int)d.hage[gp] >= P->case_absent_child_age_cutoff))
                        cadd_merged(&cn->cumAC, 1);
                    if (P->do_households && ((int)d.hage[gp] < P->case_absent_child_age_cutoff)) {
                        if (!quarantined(d, gp, ts)) cadd_merged(&cn->cumACS, 1);
                        if (d.prob_care[gp] < P->case_absent_child_prop_adult_carers) req = 1;
                    }
                }
            }
        }
    }
    if (d.bits[gp] & 2) {
        cadd_merged(&cn->cumDC, 1);
        if (d.ad_cnt) cadd_merged(ad_counter(d, r, p, CS_TSU_INC_DC), 1);  // cumDC_adunit (src/Update.cpp:872)
        // DoDetectedCase (src/Update.cpp:525-560): microcell trigger counts
        uint32_t* tg = d.m_trig + ((size_t)r * d.NM + d.mcell[p]) * 6;
        for (int q = 0; q < 5; q++) atomicAdd(&tg[q], 1u);
        if (!P->absenteeism_place_closure && P->place_close_round_household) atomicAdd(&tg[5], 1u);
    }
    if (treated(d, gp, ts)) atomicAdd(&d.cumTC[(size_t)r * d.NC + d.pc[p]], 1);
    cadd_merged(&cn->cumC, 1);
    cadd_merged(&cn->cumCa[ag], 1);
    if (d.ad_cnt) cadd_merged(ad_counter(d, r, p, CS_TSU_INC_C), 1);  // cumC_adunit (src/Update.cpp:892)
    cadd_merged(&cn->cumC_keyworker[keyw_g(d, gp)], 1);
    if (P->do_severity) {
        if (d.sev_final[gp] == CS_SEV_MILD) {
            if (d.sev_cur[gp] == CS_SEV_ASYMPTOMATIC) { d.sev_cur[gp] = CS_SEV_MILD; sev_to(cn, CS_SK_MILD, ag); }
        } else {
            if (d.sev_cur[gp] == CS_SEV_ASYMPTOMATIC) { d.sev_cur[gp] = CS_SEV_ILI; sev_to(cn, CS_SK_ILI, ag); }
        }
    }
    shot_sync_at(d, gp, sl);  // the changed status and absence, in its slot copy
    return req;
}

// infectious lists: case onset, severity and outcome timers
// (src/Sweep.cpp:929-977); a wave takes a chunk of up to 64 infectious persons
// at most 72 VGPRs: 7 waves per SIMD, no spills
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) k_infectious_scan(Dev d)
{
    if (halted(d)) return;
    const unsigned ts = d.sc->ts;
    const cs_params* P = d.P;
    const unsigned nch = min(d.wl[WL_SEG], d.seg_cap);
    const int lane = threadIdx.x & 63;
    unsigned long long nt = 0;
    for (unsigned ch = grid_wave(); ch < nch; ch += grid_nwaves()) {
        const Chunk e = chunk_at(d.segl, ch);
        const int rc = e.rc;
        const int q = e.first + lane;
        if (q >= uni(d.cI[rc])) continue;
        const int r = rc / d.NC, c = rc % d.NC;
        const size_t rN = (size_t)r * d.N;
        Counters* cn = cstage(d, r);
        const size_t sl = rN + uni(d.cstart[c] + d.cS[rc] + d.cL[rc]) + q;  // the person's list slot
        if (d.gate[sl] > ts) continue;  // every timer of the slot's person lies after this step
        const int p = d.perm[sl];
        const size_t gp = rN + p;
        int ag = age_group_g(d, gp);
        if (ag >= CS_NUM_AGE_GROUPS) ag = CS_NUM_AGE_GROUPS - 1;
        uint16_t ct = (uint16_t)(d.t_lat[gp] + ((int)(P->latent_to_sympt_delay / P->model_time_step)));
        if (P->do_symptoms && ts == ct) {
            if (do_case(d, r, p, ts, sl)) {
                const unsigned idx = atomicAdd(d.care_count, 1u);
                if (idx < (unsigned)((size_t)d.R * d.N)) d.care_gp[idx] = (int32_t)gp;  // a child case makes at most this and one isolation request
                else atomicOr(d.status, (unsigned)FLAG_LIST);
            }
            // a detected case: DoDetectedCase's interventions follow the scan
            if (P->interventions_active && (d.bits[gp] & 2)) {
                const unsigned idx = atomicAdd(d.itv_n, 1u);
                if (idx < d.det_cap) d.det[idx] = (int32_t)gp;
                else atomicOr(d.status, (unsigned)FLAG_LIST);
            }
            nt++;
        }
        const uint16_t rod = d.t_rod[gp];
        const bool die = d.bits[gp] & 1;
        if (P->do_severity) {
            uint8_t sc = d.sev_cur[gp];
            if (ts == d.t_sari[gp] && sc == CS_SEV_ILI) { sc = CS_SEV_SARI; sev_from(cn, CS_SK_ILI, ag); sev_to(cn, CS_SK_SARI, ag); }
            if (ts == d.t_crit[gp] && sc == CS_SEV_SARI) { sc = CS_SEV_CRITICAL; sev_from(cn, CS_SK_SARI, ag); sev_to(cn, CS_SK_CRITICAL, ag); }
            if (ts == d.t_step[gp] && sc == CS_SEV_CRITICAL && !die) { sc = CS_SEV_STEPDOWN; sev_from(cn, CS_SK_CRITICAL, ag); sev_to(cn, CS_SK_CRITRECOV, ag); }
            if (ts == rod) {
                if (die) {
                    if (sc == CS_SEV_CRITICAL) { sev_from(cn, CS_SK_CRITICAL, ag); sev_cum_only(cn, CS_SK_DEATH_CRITICAL, ag); sc = CS_SEV_DEAD; }
                    else if (sc == CS_SEV_SARI) { sev_from(cn, CS_SK_SARI, ag); sev_cum_only(cn, CS_SK_DEATH_SARI, ag); sc = CS_SEV_DEAD; }
                    else if (sc == CS_SEV_ILI) { sev_from(cn, CS_SK_ILI, ag); sev_cum_only(cn, CS_SK_DEATH_ILI, ag); sc = CS_SEV_DEAD; }
                } else if (d.inf[gp] == CS_INF_ASYMPT || d.inf[gp] == CS_CASE) {
                    if (sc == CS_SEV_MILD) { sev_from(cn, CS_SK_MILD, ag); sc = CS_SEV_RECOVERED; }
                    else if (sc == CS_SEV_ILI) { sev_from(cn, CS_SK_ILI, ag); sc = CS_SEV_RECOVERED; }
                    else if (sc == CS_SEV_SARI) { sev_from(cn, CS_SK_SARI, ag); sc = CS_SEV_RECOVERED; }
                    else if (sc == CS_SEV_STEPDOWN) { sev_from(cn, CS_SK_CRITRECOV, ag); sc = CS_SEV_RECOVERED; }
                }
            }
            d.sev_cur[gp] = sc;
        }
        if (ts == rod) {
            bool recover;
            if (!die) recover = true;
            else {
                recover = false;
                if (treated(d, gp, ts)) {
                    Stream s(d.seed[r], cw0(ts, PS_RECOVER), (uint32_t)p, 0);
                    recover = s.next() < P->treat_death_drop;
                }
            }
            const int8_t st = d.inf[gp];
            if (st == CS_INF_ASYMPT || st == CS_CASE) {
                if (recover) d.inf[gp] = (st == CS_CASE) ? CS_RECOVERED_SYMP : CS_RECOVERED_ASYMP;  // DoRecover
                else {  // DoDeath
                    d.inf[gp] = (st == CS_CASE) ? CS_DEAD_SYMP : CS_DEAD_ASYMP;
                    cadd_merged(&cn->cumDa[ag], 1);
                    if (d.ad_cnt) cadd_merged(ad_counter(d, r, p, CS_TSU_INC_DEATH), 1);  // cumD_adunit (src/Update.cpp:976)
                }
                batch_add(d, r, c, p);
                nt++;
            }
        }
        // the earliest timer after this step (0xffff: none)
        unsigned g = 0xffffu;
        if (P->do_symptoms && ct > ts) g = ct;
        if (P->do_severity) {
            if (d.t_sari[gp] > ts && d.t_sari[gp] < g) g = d.t_sari[gp];
            if (d.t_crit[gp] > ts && d.t_crit[gp] < g) g = d.t_crit[gp];
            if (d.t_step[gp] > ts && d.t_step[gp] < g) g = d.t_step[gp];
        }
        if (rod > ts && rod < g) g = rod;
        d.gate[sl] = (uint16_t)g;
    }
    nt = wave_sum(nt);
    if ((threadIdx.x & 63) == 0 && nt) wtally(d.wc->n_transition, nt);
}

// child-care absence (src/Update.cpp:836-868, and case isolation's form
// :752-776), evaluated for every request of the step against the state after
// all timers, then applied.  A request is gp (DoCase's form) or ~gp (case
// isolation's form, which sets the carer's absence start even when the
// carer is already absent).
__global__ void k_care_eval(Dev d)
{
    if (halted(d)) return;
    const unsigned ts = d.sc->ts;
    const cs_params* P = d.P;
    const unsigned n = min(*d.care_count, (unsigned)((size_t)d.R * d.N));
    for (unsigned a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
        const int32_t v = d.care_gp[a];
        size_t gp = (size_t)(v >= 0 ? v : ~v);
        int r = (int)(gp / d.N), p = (int)(gp % d.N);
        const size_t rN = (size_t)r * d.N;
        int h = d.hh[p], j1 = d.hfirst[h], j2 = j1 + d.hnh[h];
        bool f = false;
        for (int j3 = j1; (j3 < j2) && !f; j3++)
            f = (!is_dead_code(d.inf[rN + j3]) && ((int)d.age[j3] >= P->case_absent_child_age_cutoff) &&
                 ((d.plinks[(size_t)j3 * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] < 0) ||
                  host_absent(d, rN + j3, ts) || quarantined(d, rN + j3, ts)));
        int kk = -1;
        if (!f) {
            for (int j3 = j1; (j3 < j2) && !f; j3++)
                if (((int)d.age[j3] >= P->case_absent_child_age_cutoff) && !is_dead_code(d.inf[rN + j3])) { kk = j3; f = true; }
            if (f) cadd_merged(&d.cnt[r].cumAA, 1);
        }
        d.care_k[a] = kk;
        d.care_abs[a] = kk >= 0 ? (uint8_t)host_absent(d, rN + kk, ts) : 0;
    }
}

__global__ void k_care_apply(Dev d)
{
    if (halted(d)) return;
    const unsigned ts = d.sc->ts;
    const cs_params* P = d.P;
    const unsigned n = min(*d.care_count, (unsigned)((size_t)d.R * d.N));
    for (unsigned a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
        int kk = d.care_k[a];
        if (kk < 0) continue;
        const int32_t v = d.care_gp[a];
        size_t gp = (size_t)(v >= 0 ? v : ~v);
        size_t gk = (gp / d.N) * d.N + kk;
        if (!d.care_abs[a] || v < 0) d.t_as[gk] = (uint16_t)(ts + P->us_case_isolation_delay);
        d.t_ae[gk] = (uint16_t)(ts + P->us_case_isolation_delay + P->us_case_isolation_duration);
        shot_sync(d, gk, kk);
    }
}

// --------------------------------------------------------------- reset

// InitModel's per-person reset (src/CovidSim.cpp:1006-1037)
__global__ void k_reset_persons(Dev d)
{
    const cs_params* P = d.P;
    const long long n = (long long)d.R * d.N;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long long)gridDim.x * blockDim.x) {
        int r = (int)(g / d.N), k = (int)(g % d.N);
        d.t_as[g] = CS_UNSET_TIME; d.t_ae[g] = 0;
        d.t_vs[g] = d.t_ts[g] = d.t_is[g] = CS_UNSET_TIME;
        d.t_te[g] = 0;
        d.bits[g] = 0; d.t_det[g] = 0; d.dct[g] = 0;
        d.inf[g] = CS_SUSCEPTIBLE;
        d.t_lat[g] = d.t_rod[g] = 0; d.t_inf[g] = 0;
        d.gate[g] = 0;
        d.hflags[g] = (uint8_t)((no_place_links(d.plk[k].lk, d.NPT) ? HOT_NO_LINKS : 0) | (d.hnh[d.hh[k]] << HOT_NH_SHIFT));
        d.hage[g] = d.age[k];
        d.hpfl[g] = d.pflags[k];
        d.infector[g] = -1; d.itype[g] = 0;
        d.qc[g] = 2; d.t_qs[g] = CS_UNSET_TIME;
        Stream s(d.seed[r], cw0(0, PS_INIT), (uint32_t)k, 0);
        (void)s.next();  // ProbAbsent (unused on this path)
        d.prob_care[g] = (float)s.next();
        float su = 1.0f;
        if (P->susceptibility_sd > 0) {
            double kk = 1 / (P->susceptibility_sd * P->susceptibility_sd);
            su *= (float)gamma(s, kk, kk);
        }
        d.susc[g] = su;
        d.infness[g] = 0.0f;
        d.t_sari[g] = d.t_crit[g] = d.t_step[g] = CS_UNSET_TIME;
        d.sev_cur[g] = d.sev_final[g] = CS_SEV_ASYMPTOMATIC;
        d.claim[g] = ~0ull;
        d.lastk[g] = 0ull;
    }
}

__global__ void k_reset_cells(Dev d)
{
    const long long n = (long long)d.R * d.NC;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long long)gridDim.x * blockDim.x) {
        int r = (int)(g / d.NC), c = (int)(g % d.NC);
        const size_t rN = (size_t)r * d.N;
        const int base = d.cstart[c];
        for (int j = 0; j < d.cn[c]; j++) {
            int k = d.cmem[base + j];
            d.perm[rN + base + j] = k;
            d.listpos[rN + k] = j;
        }
        d.cS[g] = d.cn[c];
        d.cL[g] = d.cI[g] = d.cR[g] = d.cD[g] = d.cumTC[g] = 0;
        d.bcount[g] = 0;
    }
}

__global__ void k_reset_misc(Dev d)
{
    const int tot = d.np_off[CS_MAX_PLACE_TYPES];
    const long long n = (long long)d.R * tot;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long long)gridDim.x * blockDim.x) {
        d.pl_cs[g] = CS_UNSET_TIME;
        d.pl_ce[g] = 0;
    }
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long long)gridDim.x * blockDim.x)
        d.pl_once[g] = 0;
    const long long nm = (long long)d.R * d.NM;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < nm; g += (long long)gridDim.x * blockDim.x) {
        d.m_socdist[g] = 0;
        d.m_moverest[g] = 0;
        for (int q = 0; q < 6; q++) d.m_trig[g * 6 + q] = 0;
        d.mc_state[g] = MC_UNTREATED;
        d.mc_end[g] = 0;
    }
}

// SeedInfection's random-location branch (src/CovidSim.cpp:1338-1377): one
// thread per realisation draws the seeds in sequence
__global__ void k_seed(Dev d, int n_seed)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d.R) return;
    const cs_params* P = d.P;
    const size_t rN = (size_t)r * d.N;
    d.loc_init[2 * r] = P->location_initial_infection[0];
    d.loc_init[2 * r + 1] = P->location_initial_infection[1];
    d.prev_prop[r] = 1.0;
    Stream s(d.seed[r], cw0(0, PS_SEED), 0, 0);
    int choices = 0;
    if (d.rseed_n) n_seed = d.rseed_n[r];
    for (int inf = 0; (inf < n_seed) && (choices < 10000); inf++) {
        int p, m, nm;
        do {
            p = (int)(s.next() * (double)d.N);
            m = d.mcell[p];
            nm = d.mstart[m + 1] - d.mstart[m];
        } while (nm == 0);
        int q = d.mmem[d.mstart[m] + (int)(s.next() * (double)nm)];
        if (d.inf[rN + q] == CS_SUSCEPTIBLE && (person_susc(d, r, q, 0, 0) > 0) &&
            (P->care_home_place_type < 0 || d.plinks[(size_t)q * CS_MAX_PLACE_TYPES + P->care_home_place_type] < 0)) {
            const float2 loc = d.ploc[q];
            d.loc_init[2 * r] = loc.x;
            d.loc_init[2 * r + 1] = loc.y;
            d.infector[rN + q] = -2;
            d.itype[rN + q] = CS_INFECT_TYPE_MASK - 1;
            do_infect(d, r, q, 0.0, 0, -2, CS_INFECT_TYPE_MASK - 1);
            choices = 0;
        } else {
            inf--;
            choices++;
        }
    }
}

// the trigger logic run at each sample when interventions are active
// (defined with the intervention kernels)
__device__ void itv_triggers(const Dev& d, int r, int m, double t, long long D);

// daily totals (one block per realisation)
// sample: the index to write, or -1 for the device clock's sample count
__global__ void k_record(Dev d, int sample)
{
    if (halted(d)) return;
    if (sample < 0) sample = d.clock->sample;
    int r = blockIdx.x;
    __shared__ long long part[5][256];
    long long v[5] = {0, 0, 0, 0, 0};
    for (int c = threadIdx.x; c < d.NC; c += blockDim.x) {
        size_t i = (size_t)r * d.NC + c;
        v[0] += d.cS[i]; v[1] += d.cL[i]; v[2] += d.cI[i]; v[3] += d.cR[i]; v[4] += d.cD[i];
    }
    for (int q = 0; q < 5; q++) part[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 5; q++) part[q][threadIdx.x] += part[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) d.samp_prop[r] = (double)part[0][0] / (double)d.N;
    // the next output step's CalibrationThresholdCheck: this is sample
    // number sample + 1 of the series, at the time after the step
    if (threadIdx.x == 0 && d.P->interventions_active) itv_triggers(d, r, sample + 1, d.sc->t + d.P->model_time_step, part[4][0]);
    if (threadIdx.x == 0 && sample < d.ts_max) {
        int32_t* o = d.ts_buf + ((size_t)r * d.ts_max + sample) * CS_TS_FIELDS;
        const Counters& cn = d.cnt[r];
        long long now[CS_TS_LAST];
        now[0] = (long long)cn.cumI;
        now[1] = (long long)cn.cumC;
        now[2] = part[4][0];
        for (int k = 0; k < CS_NUM_SEV_KINDS; k++) now[3 + k] = (long long)cn.sev_cum[k];
        long long* lt = d.last_tot + CS_TS_LAST * r;
        for (int q = 0; q < 5; q++) o[q] = (int32_t)part[q][0];
        o[CS_TS_INC_I] = (int32_t)(now[0] - lt[0]);
        o[CS_TS_INC_C] = (int32_t)(now[1] - lt[1]);
        o[CS_TS_INC_DEATH] = (int32_t)(now[2] - lt[2]);
        for (int k = 0; k < CS_NUM_SEV_KINDS; k++) o[CS_TS_INC_SEV + k] = (int32_t)(now[3 + k] - lt[3 + k]);
        for (int k = 0; k <= CS_SK_CRITRECOV; k++) o[CS_TS_PREV_SEV + k] = (int32_t)(long long)cn.sev_prev[k];
        for (int q = 0; q < CS_TS_LAST; q++) lt[q] = now[q];
    }
    if (sample < d.ts_max && d.ts_ext) {  // by age group and by unit
        int32_t* o = d.ts_ext + ((size_t)r * d.ts_max + sample) * d.ext_w;
        long long* le = d.last_ext + (size_t)r * (CS_TSA_CUM * CS_NUM_AGE_GROUPS + CS_TSU_FIELDS * d.NA);
        const Counters& cn = d.cnt[r];
        for (int q = threadIdx.x; q < CS_TSA_FIELDS * CS_NUM_AGE_GROUPS; q += blockDim.x) {
            const int f = q / CS_NUM_AGE_GROUPS, a = q % CS_NUM_AGE_GROUPS;
            if (f >= CS_TSA_PREV_SEV) {
                o[q] = (int32_t)(long long)cn.sev_prev_age[f - CS_TSA_PREV_SEV][a];
                continue;
            }
            const long long now = (long long)(f == CS_TSA_INC_I ? cn.cumIa[a] : f == CS_TSA_INC_C ? cn.cumCa[a] :
                                              f == CS_TSA_INC_DEATH ? cn.cumDa[a] : cn.sev_cum_age[f - CS_TSA_INC_SEV][a]);
            o[q] = (int32_t)(now - le[q]);
            le[q] = now;
        }
        if (d.ad_cnt)
            for (int q = threadIdx.x; q < CS_TSU_FIELDS * d.NA; q += blockDim.x) {
                const int f = q / d.NA, u = q % d.NA;
                const long long now = ad_total(d, r, u, f);
                long long* l = le + CS_TSA_CUM * CS_NUM_AGE_GROUPS + q;
                o[CS_TSA_FIELDS * CS_NUM_AGE_GROUPS + q] = (int32_t)(now - *l);
                *l = now;
            }
    }
}

// the slot copies of realisation r's infectious persons, from a loaded state
__global__ void k_shot_build(Dev d, int r)
{
    const size_t rN = (size_t)r * d.N;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < d.N; p += gridDim.x * blockDim.x)
        if (in_infectious_list(d.inf[rN + p])) d.shot[rN + d.cstart[d.pc[p]] + d.listpos[rN + p]] = d.hot[rN + p];
}

__global__ void k_record_init(Dev d, double t0)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d.R) return;
    long long D = 0, S = 0;
    for (int c = 0; c < d.NC; c++) {
        D += d.cD[(size_t)r * d.NC + c];
        S += d.cS[(size_t)r * d.NC + c];
    }
    d.samp_prop[r] = (double)S / (double)d.N;  // the sample RunModel takes first
    long long* lt = d.last_tot + CS_TS_LAST * r;
    lt[0] = (long long)d.cnt[r].cumI;
    lt[1] = (long long)d.cnt[r].cumC;
    lt[2] = D;
    for (int k = 0; k < CS_NUM_SEV_KINDS; k++) lt[3 + k] = (long long)d.cnt[r].sev_cum[k];
    if (d.ts_ext || d.last_ext) {
        const Counters& cn = d.cnt[r];
        long long* le = d.last_ext + (size_t)r * (CS_TSA_CUM * CS_NUM_AGE_GROUPS + CS_TSU_FIELDS * d.NA);
        for (int a = 0; a < CS_NUM_AGE_GROUPS; a++) {
            le[CS_TSA_INC_I * CS_NUM_AGE_GROUPS + a] = (long long)cn.cumIa[a];
            le[CS_TSA_INC_C * CS_NUM_AGE_GROUPS + a] = (long long)cn.cumCa[a];
            le[CS_TSA_INC_DEATH * CS_NUM_AGE_GROUPS + a] = (long long)cn.cumDa[a];
            for (int k = 0; k < CS_NUM_SEV_KINDS; k++) le[(CS_TSA_INC_SEV + k) * CS_NUM_AGE_GROUPS + a] = (long long)cn.sev_cum_age[k][a];
        }
        if (d.ad_cnt)
            for (int q = 0; q < CS_TSU_FIELDS * d.NA; q++)
                le[CS_TSA_CUM * CS_NUM_AGE_GROUPS + q] = ad_total(d, r, q % d.NA, q / d.NA);
    }
    // the series' first sample and its CalibrationThresholdCheck (src/CovidSim.cpp:1413-1414)
    if (d.P->interventions_active) itv_triggers(d, r, 0, t0, D);
}

// ----------------------------------------------- single-function kernels

__global__ void k_t_kernel_num(Dev d, const double* r2, double* out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = kernel_num(d, r2[i]);
}

__global__ void k_t_dist2(Dev d, const int32_t* a, const int32_t* b, double* out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = dist2_pp(d, a[i], b[i]);
}

__global__ void k_t_hazards(Dev d, int r, int kind, unsigned ts, const int32_t* person, const int32_t* other, double* out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x = person[i], o = other ? other[i] : 0;
    double v;
    switch (kind) {
    case 0: v = house_inf(d, r, x, ts); break;
    case 1: v = place_inf(d, r, x, o, ts); break;
    case 2: v = spatial_inf(d, r, x, ts); break;
    case 3: v = person_inf(d, (size_t)r * d.N + x, ts); break;
    case 4: v = house_susc(d, r, x, ts, o); break;
    case 5: v = place_susc(d, r, x, o, ts); break;
    case 6: v = spatial_susc(d, r, x, ts); break;
    default: v = person_susc(d, r, x, ts, o); break;
    }
    out[i] = v;
}

__global__ void k_t_choose(Dev d, int table, const double* u, const double* mean, uint16_t* out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = icdf_choose(d.P->icdf[table], mean[i], u[i], d.P->time_steps_per_day);
}

__global__ void k_t_select_infector(const float* ci, int n_inf, const double* u, int32_t* out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = select_infector(ci, n_inf, u[i]);
}

__global__ void k_t_select_dest(Dev d, int r, int c, const double* u, int32_t* out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (d.ck) {
        float mt;
        out[i] = select_dest_ck(d, r, c, u[i], &mt);
    } else {
        out[i] = select_dest(d, (size_t)r * d.NC + c, u[i]);
    }
}

__global__ void k_t_philox(const uint32_t* ck, uint32_t* out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t o[4];
    const uint32_t* a = ck + 6 * i;
    philox4x32_10(a[0], a[1], a[2], a[3], a[4], a[5], o);
    for (int q = 0; q < 4; q++) out[4 * i + q] = o[q];
}

// samplers on given streams (distribution tests): which 0 uniform,
// 1 binomial(a, b), 2 poisson(a), 3 normal, 4 gamma(a, b)
__global__ void k_t_sample(int which, uint64_t key, uint32_t c0, uint32_t c1, double a, double b, double* out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Stream s(key, c0, c1, (uint32_t)i);
    double v;
    switch (which) {
    case 0: v = s.next(); break;
    case 1: v = (double)binomial(s, (int32_t)a, b); break;
    case 2: v = (double)poisson(s, a); break;
    case 3: v = normal(s); break;
    default: v = gamma(s, a, b); break;
    }
    out[i] = v;
}

}  // namespace cs
