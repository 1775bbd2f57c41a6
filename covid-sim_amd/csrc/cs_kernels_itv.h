// cs_kernels_itv.h -- the intervention subset on the device: the trigger and
// start-time logic of RecordSample / CalibrationThresholdCheck
// (src/CovidSim.cpp:3150-3160, :3485-3689), case isolation and household
// quarantine of detected cases (DoDetectedCase, src/Update.cpp:525-792), and
// place closure by microcell on global triggers (TreatSweep,
// src/Sweep.cpp:1422-1437, :1651-1740; DoPlaceClose, src/Update.cpp:1095-1260;
// UpdateHostClosure, :1263-1279).  Every kernel here is launched only when
// P->interventions_active.
//
// Random draws come from streams of their own: household quarantine's from
// the household's (keyed by its first member), so one application per
// household per step gives the same result whichever detected case of the
// household makes it; case isolation's from the case's; Place.ProbClose and
// Person.ProbAbsent (drawn once per realisation in the reference) from
// reset-time streams, recomputed where read.
#pragma once
#include "cs_kernels_incub.h"

namespace cs {

__device__ __forceinline__ float prob_close(const Dev& d, int r, int pt, int k)
{
    Stream s(d.seed[r], cw0(0, PS_PLACE), (uint32_t)k, (uint32_t)pt);
    return (float)s.next();
}
// the first draw of the person's reset stream (k_reset_persons discards it)
__device__ __forceinline__ float prob_absent(const Dev& d, int r, int p)
{
    Stream s(d.seed[r], cw0(0, PS_INIT), (uint32_t)p, 0);
    return (float)s.next();
}

// a trigger threshold of realisation r: its own where the realisations have
// theirs (cs_set_realisation_triggers, for the three thresholds at once),
// else the run's value v
__device__ __forceinline__ double itv_thresh(const Dev& d, int r, double v) { return d.rthr ? d.rthr[r] : v; }

// RecordSample's trigger value and CalibrationThresholdCheck's alert and
// start times for realisation r at its sample m, time t; cumDC and
// cumCritical from the counters, D the sample's dead total
__device__ void itv_triggers(const Dev& d, int r, int m, double t, long long D)
{
    const cs_params* P = d.P;
    IntState& s = d.ist[r];
    const Counters& cn = d.cnt[r];
    int32_t* h = d.trig_hist + ((size_t)r * CS_TRIG_RING + (unsigned)m % CS_TRIG_RING) * 3;
    h[0] = (int32_t)cn.cumDC;
    h[1] = (int32_t)cn.sev_cum[CS_SK_CRITICAL];
    h[2] = (int32_t)D;
    const int tsi = P->triggers_sampling_interval, win = P->window_to_evaluate_trigger_alert;
    if ((m >= tsi && tsi >= CS_TRIG_RING) || (m >= win && win >= CS_TRIG_RING)) {
        atomicOr(d.status, (unsigned)FLAG_TRIG);
        return;
    }
    const int32_t* base = d.trig_hist + (size_t)r * CS_TRIG_RING * 3;
    const int iv = P->do_icu_triggers ? 1 : 0;
    int32_t trig = h[iv];
    if (m >= tsi) trig -= base[(size_t)((unsigned)(m - tsi) % CS_TRIG_RING) * 3 + iv];
    s.trig = trig;
    int32_t alert = h[0];
    if (m >= win) alert -= base[(size_t)((unsigned)(m - win) % CS_TRIG_RING) * 3];
    if (P->trigger_alert_on_deaths) {
        alert = h[2];
        if (m >= win) alert -= base[(size_t)((unsigned)(m - win) % CS_TRIG_RING) * 3 + 2];
    }
    if (alert >= P->case_or_death_threshold_before_alert) {
        s.alert = 1;
        s.ctrl = P->post_alert_control_prop_cases_id;
        // DO_OR_DONT_AMEND_START_TIME: each start time is set once
        if (trig > itv_thresh(d, r, P->case_isolation_cell_inc_thresh) && s.ci_start >= 1e10) s.ci_start = t + P->case_isolation_time_start_base;
        if (trig > itv_thresh(d, r, P->hhquar_cell_inc_thresh) && s.hq_start >= 1e10) s.hq_start = t + P->hquarantine_time_start_base;
        if (trig >= itv_thresh(d, r, P->place_close_cell_inc_thresh) && s.pc_start >= 1e10) s.pc_start = t + P->place_close_time_start_base;
    }
    // the first closure ends after its duration when nothing re-triggers it
    if (t == s.pc_start + P->place_close_duration) {
        s.pc_prev = s.pc_start;
        if (P->place_close_inc_trig == 0 && P->place_close_frac_inc_trig == 0 && itv_thresh(d, r, P->place_close_cell_inc_thresh) == 0) s.pc_start = 9e9;
    }
}

// InitModel's intervention state of every realisation
// (src/CovidSim.cpp:1216-1232): one thread per realisation
__global__ void k_reset_interventions(Dev d)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d.R) return;
    IntState s;
    s.alert = 0;
    s.trig = 0;
    s.ctrl = d.P->pre_alert_control_prop_cases_id;
    s.ci_start = s.hq_start = s.pc_start = s.pc_prev = 1e10;
    d.ist[r] = s;
    int32_t* h = d.trig_hist + (size_t)r * CS_TRIG_RING * 3;
    for (int q = 0; q < CS_TRIG_RING * 3; q++) h[q] = 0;
}

// p's case onset (DoCase) was at this step and p is detected
__device__ __forceinline__ bool detected_now(const Dev& d, size_t gp, unsigned ts)
{
    const uint16_t ct = (uint16_t)(d.t_lat[gp] + ((int)(d.P->latent_to_sympt_delay / d.P->model_time_step)));
    return ct == ts && (d.bits[gp] & 2);
}

__device__ __forceinline__ bool to_be_quarantined(const Dev& d, size_t gp, unsigned ts)
{
    return ((int)d.t_qs[gp] + (int)d.P->us_hquarantine_house_duration > (int)ts) && (d.qc[gp] < 2);
}

__device__ __forceinline__ void itv_list_add(const Dev& d, int32_t* list, unsigned* count, unsigned cap, int32_t v)
{
    const unsigned idx = atomicAdd(count, 1u);
    if (idx < cap) list[idx] = v;
    else atomicOr(d.status, (unsigned)FLAG_LIST);
}

// household quarantine of DoDetectedCase (src/Update.cpp:612-639) for each
// of the step's detected cases: the household's lowest-indexed detected case
// of the step applies it (the draws are the household's)
__global__ void k_detected_hq(Dev d)
{
    if (halted(d)) return;
    const cs_params* P = d.P;
    if (!P->do_households) return;
    const StepScalars sc = *d.sc;
    const unsigned ts = sc.ts, n = min(d.itv_n[0], d.det_cap);
    for (unsigned a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
        const size_t gp = (size_t)d.det[a];
        const int r = (int)(gp / d.N), p = (int)(gp % d.N);
        const IntState& s = d.ist[r];
        if (!(sc.t >= s.hq_start && sc.t < s.hq_start + P->hquarantine_policy_duration)) continue;
        const size_t rN = (size_t)r * d.N;
        const int h = d.hh[p], j1 = d.hfirst[h], j2 = j1 + d.hnh[h];
        bool first = true;
        for (int j = j1; j < p && first; j++) first = !detected_now(d, rN + j, ts);
        if (!first) continue;
        if (to_be_quarantined(d, rN + j1, ts) && !P->do_hq_retrigger) continue;
        const uint16_t start = (uint16_t)(ts + (uint16_t)(P->time_steps_per_day * P->hquarantine_delay));
        Stream st(d.seed[r], cw0(ts, PS_HQ), (uint32_t)j1, 0);
        const int k = (st.next() < P->hquarantine_prop_house_compliant) ? 1 : 0;
        Counters* cn = d.cnt + r;
        if (k) cadd(&cn->cumHQ, 1);
        for (int j = j1; j < j2; j++) {
            const size_t gj = rN + j;
            d.t_qs[gj] = start;
            const uint8_t c = (uint8_t)((k == 0) ? 0 : ((st.next() < P->hquarantine_prop_indiv_compliant) ? 1 : 0));
            d.qc[gj] = c;
            if (c && !host_absent(d, gj, ts)) {
                if ((int)d.age[j] >= P->case_absent_child_age_cutoff) {
                    if (d.plinks[(size_t)j * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] >= 0) cadd(&cn->cumAH, 1);
                } else
                    cadd(&cn->cumACS, 1);
            }
        }
    }
}

// case isolation of DoDetectedCase (src/Update.cpp:641-692), after every
// household quarantine of the step; a child's carer request joins the
// step's child-care requests in isolation's form (~gp)
__global__ void k_detected_ci(Dev d)
{
    if (halted(d)) return;
    const cs_params* P = d.P;
    const StepScalars sc = *d.sc;
    const unsigned ts = sc.ts, n = min(d.itv_n[0], d.det_cap);
    for (unsigned a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
        const size_t gp = (size_t)d.det[a];
        const int r = (int)(gp / d.N), p = (int)(gp % d.N);
        const IntState& s = d.ist[r];
        if (!(sc.t >= s.ci_start && sc.t < s.ci_start + P->case_isolation_policy_duration)) continue;
        Stream st(d.seed[r], cw0(ts, PS_ISO), (uint32_t)p, 0);
        if (!((P->case_isolation_prop == 1) || (st.next() < P->case_isolation_prop))) continue;
        d.t_is[gp] = (uint16_t)ts;
        if (host_absent(d, gp, ts)) {
            const int stop = (int)ts + P->us_case_absenteeism_delay + P->us_case_isolation_duration;
            if ((int)d.t_ae[gp] < stop) d.t_ae[gp] = (uint16_t)stop;
            shot_sync(d, gp, p);
            continue;
        }
        if (!P->do_real_sympt_withdrawal) continue;
        d.t_as[gp] = (uint16_t)(ts + P->us_case_isolation_delay);
        d.t_ae[gp] = (uint16_t)(ts + P->us_case_isolation_delay + P->us_case_isolation_duration);
        shot_sync(d, gp, p);
        const bool child = (int)d.age[p] < P->case_absent_child_age_cutoff;
        Counters* cn = d.cnt + r;
        if (P->do_places && !quarantined(d, gp, ts) && d.plinks[(size_t)p * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] >= 0 && !child)
            cadd(&cn->cumAC, 1);
        if (P->do_households && P->do_places && child) {
            if (!quarantined(d, gp, ts)) cadd(&cn->cumACS, 1);
            if (d.prob_care[gp] < P->case_absent_child_prop_adult_carers)
                itv_list_add(d, d.care_gp, d.care_count, (unsigned)((size_t)d.R * d.N), ~(int32_t)gp);
        }
    }
}

// DoPlaceClose with DoAnyway (src/Update.cpp:1095-1190): the place's closure
// window; true when the place closes (its members' absences follow)
__device__ __forceinline__ bool close_place(const Dev& d, int r, int pt, int k, uint16_t t_start, uint16_t t_stop)
{
    const size_t pi = (size_t)r * d.np_off[CS_MAX_PLACE_TYPES] + np_off_of(d, pt) + k;
    if (!(d.pl_ce[pi] < t_stop) || d.pl_once[pi]) return false;
    if (d.P->do_place_close_once_only) d.pl_once[pi] = 1;
    if ((double)prob_close(d, r, pt, k) >= d.P->efficacies[2][pt]) {
        if (d.pl_cs[pi] > t_start) d.pl_cs[pi] = t_start;
        d.pl_ce[pi] = t_stop;
        return true;
    }
    d.pl_cs[pi] = d.pl_ce[pi] = t_stop;
    return false;
}

// the closure window of this step's TreatSweep
struct CloseWindow {
    uint16_t t_end, t_start, t_stop;
};
__device__ __forceinline__ CloseWindow close_window(const cs_params* P, double t, unsigned ts)
{
    CloseWindow w;
    w.t_end = (uint16_t)__builtin_ceil(P->time_steps_per_day * (t + P->place_close_delay_mean + P->place_close_duration));
    w.t_start = (uint16_t)(ts + (uint16_t)(P->time_steps_per_day * P->place_close_delay_mean));
    w.t_stop = (uint16_t)(ts + (uint16_t)(P->time_steps_per_day * (P->place_close_delay_mean + P->place_close_duration)));
    return w;
}

// TreatSweep's place-closure block on global triggers, one thread per
// (realisation, microcell); the places that close are listed for the member
// passes
__global__ void k_treat_mcells(Dev d)
{
    if (halted(d)) return;
    const cs_params* P = d.P;
    const StepScalars sc = *d.sc;
    const unsigned ts = sc.ts;
    const CloseWindow w = close_window(P, sc.t, ts);
    const unsigned nm = (unsigned)d.NM, n = (unsigned)d.R * nm;  // below 2^32 (checked at allocation)
    const int tot = d.np_off[CS_MAX_PLACE_TYPES];
    for (unsigned g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
        const int r = (int)(g / nm), m = (int)(g - (unsigned)r * nm);  // 32-bit division: a 64-bit one costs the pass most of its time
        const IntState& s = d.ist[r];
        if (!(sc.t >= s.pc_start)) continue;
        if (d.mstart[m + 1] == d.mstart[m]) continue;  // unpopulated
        if (d.mc_state[g] == MC_TREATED && ts >= d.mc_end[g]) {
            d.mc_state[g] = P->do_place_close_once_only ? MC_DONT_TREAT_AGAIN : MC_UNTREATED;
            d.mc_end[g] = (uint16_t)ts;
        }
        if (P->do_places && d.mc_state[g] == MC_UNTREATED && s.trig >= itv_thresh(d, r, P->place_close_cell_inc_thresh)) {
            d.mc_end[g] = w.t_end;
            d.mc_state[g] = MC_TREATED;
            for (int q = d.mpl_start[m]; q < d.mpl_start[m + 1]; q++) {
                const int pt = d.mpl[q] >> 27, k = d.mpl[q] & ((1 << 27) - 1);
                if (pt == P->hotel_place_type) continue;
                if (close_place(d, r, pt, k, w.t_start, w.t_stop))
                    itv_list_add(d, d.close_l, d.itv_n + 1, d.close_cap, (int32_t)((long long)r * tot + np_off_of(d, pt) + k));
            }
        }
    }
}

// does a member of p's household count as a free adult for the carer search
// of a closed place (alive, adult, and no link to place type
// PlaceTypeNoAirNum or quarantined; src/Update.cpp:1214-1217)
__device__ bool household_has_free_adult_pc(const Dev& d, size_t rN, int p, unsigned ts)
{
    const cs_params* P = d.P;
    const int h = d.hh[p], j1 = d.hfirst[h], j2 = j1 + d.hnh[h];
    for (int j = j1; j < j2; j++)
        if (!is_dead_code(d.inf[rN + j]) && ((int)d.age[j] >= P->case_absent_child_age_cutoff) &&
            ((d.plinks[(size_t)j * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] < 0) || quarantined(d, rN + j, ts)))
            return true;
    return false;
}

// the members of the places closing at this step (src/Update.cpp:1192-1260),
// a wave per place, a lane per member.  pass 0: the tests and counters of
// every member, and the carers queued (in care_k); pass 1: the absences.
// The tests of pass 0 read the state before any absence of the sweep.
__global__ void __launch_bounds__(256) k_close_members(Dev d, int pass)
{
    if (halted(d)) return;
    const cs_params* P = d.P;
    if (!P->do_real_sympt_withdrawal) return;
    const StepScalars sc = *d.sc;
    const unsigned ts = sc.ts;
    const CloseWindow w = close_window(P, sc.t, ts);
    const unsigned n = min(d.itv_n[1], d.close_cap);
    const int tot = d.np_off[CS_MAX_PLACE_TYPES];
    const int lane = threadIdx.x & 63;
    const unsigned ccap = (unsigned)((size_t)d.R * d.N);
    for (unsigned a = grid_wave(); a < n; a += grid_nwaves()) {
        const int v = uni(d.close_l[a]);
        const int r = v / tot, f = v % tot;
        int pt = 0;
        while (pt + 1 < CS_MAX_PLACE_TYPES && f >= d.np_off[pt + 1]) pt++;
        const int k = f - d.np_off[pt];
        const size_t rN = (size_t)r * d.N;
        const double att = P->place_close_prop_attending[pt];
        Counters* cn = d.cnt + r;
        const int m0 = d.pl_ms[pt][k], m1 = d.pl_ms[pt][k + 1];
        for (int q = m0 + lane; q < m1; q += 64) {
            const int wm = d.pl_mem[pt][q];
            const size_t gw = rN + wm;
            if (!((att == 0) || ((double)prob_absent(d, r, wm) >= att))) continue;
            const bool child = (int)d.age[wm] < P->case_absent_child_age_cutoff;
            if (pass == 0) {
                if (!host_absent(d, gw, ts) && !quarantined(d, gw, ts) && child) {
                    cadd_merged(&cn->cumAPCS, 1);
                    if (d.prob_care[gw] < P->case_absent_child_prop_adult_carers && !household_has_free_adult_pc(d, rN, wm, ts)) {
                        const int h = d.hh[wm], j1 = d.hfirst[h], j2 = j1 + d.hnh[h];
                        int c = -1;
                        for (int j = j1; j < j2 && c < 0; j++)
                            if (((int)d.age[j] >= P->case_absent_child_age_cutoff) && !is_dead_code(d.inf[rN + j])) c = j;
                        if (c >= 0) {
                            itv_list_add(d, d.care_k, d.itv_n + 2, ccap, (int32_t)(rN + c));
                            cadd_merged(&cn->cumAPA, 1);
                        }
                    }
                }
                if (!child && d.plinks[(size_t)wm * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] >= 0) cadd_merged(&cn->cumAPC, 1);
            } else {
                // every writer of this sweep writes the same bounds, so the
                // unordered min and max give one result
                if (d.t_as[gw] > w.t_start) d.t_as[gw] = w.t_start;
                if (d.t_ae[gw] < w.t_stop) d.t_ae[gw] = w.t_stop;
                shot_sync(d, gw, (int)(gw % d.N));
            }
        }
    }
}

// UpdateHostClosure (src/Update.cpp:1263-1279): the queued carers' absences
__global__ void k_host_closure(Dev d)
{
    if (halted(d)) return;
    const cs_params* P = d.P;
    if (!P->do_real_sympt_withdrawal) return;
    const StepScalars sc = *d.sc;
    const CloseWindow w = close_window(P, sc.t, sc.ts);
    const unsigned n = min(d.itv_n[2], (unsigned)((size_t)d.R * d.N));
    for (unsigned a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
        const size_t gc = (size_t)d.care_k[a];
        if (d.t_as[gc] > w.t_start) d.t_as[gc] = w.t_start;
        if (d.t_ae[gc] < w.t_stop) d.t_ae[gc] = w.t_stop;
        shot_sync(d, gc, (int)(gc % d.N));
    }
}

}  // namespace cs
