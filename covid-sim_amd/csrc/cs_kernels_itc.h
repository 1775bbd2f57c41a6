// cs_kernels_itc.h -- InitTransmissionCoeffs (src/SetupModel.cpp:640-846)
// on the device: per-person draws, per-person contributions to the
// household, spatial and place reproduction numbers, and a two-level sum
// (fixed-order block partials, then one block) so the result does not
// depend on the schedule.
#pragma once
#include "cs_model.h"

namespace cs {

enum { ITC_HH = 0, ITC_HH_DENOM, ITC_SPATIAL, ITC_REC_DAYS, ITC_REC_STEPS, ITC_PLACE0, ITC_NSUM = ITC_PLACE0 + CS_MAX_PLACE_TYPES };
#define ITC_BLOCK 256

// per-person draws of src/SetupModel.cpp:667-690 from the stream (seed, ITC, person)
__global__ void k_itc_draw(Dev d, uint64_t seed, float* infness, uint16_t* rec)
{
    const cs_params* P = d.P;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < d.N; p += (long long)gridDim.x * blockDim.x) {
        const int a = age_group(d, (int)p);
        Stream s(seed, cw0(0, PS_ITC), (uint32_t)p, 0);
        // the susceptibility draw comes first in the stream, as in the reference
        if (P->susceptibility_sd != 0) {
            double k = 1 / (P->susceptibility_sd * P->susceptibility_sd);
            (void)gamma(s, k, k);
        }
        float in;
        if (P->infectiousness_sd == 0) in = (float)P->age_infectiousness[a];
        else {
            double k = 1 / (P->infectiousness_sd * P->infectiousness_sd);
            in = (float)(P->age_infectiousness[a] * gamma(s, k, k));
        }
        if (s.next() < P->proportion_symptomatic[a]) in *= (float)(-P->sympt_infectiousness);
        else in *= (float)P->asympt_infectiousness;
        infness[p] = in;
        double q = s.next() * CS_CDF_RES;
        int j = (int)floor(q);
        q -= (double)j;
        const double* ic = P->icdf[CS_ICDF_INFECTIOUS];
        rec[p] = (uint16_t)floor(0.5 - (P->infectious_period * clog(q * ic[j + 1] + (1.0 - q) * ic[j]) / P->model_time_step));
    }
}

__device__ __forceinline__ double itc_survive(const cs_params* P, double ps, double rate, int d0, int d1)
{
    for (int t = d0; t < d1 && t < CS_MAX_INFECTIOUS_STEPS; t++) {
        double today = 1.0 - rate * P->infectiousness[t];
        ps *= (today < 0) ? 0 : today;
    }
    return ps;
}

// per-person contributions, summed per block in LDS in a fixed tree order;
// partial[block][ITC_NSUM]
__global__ void __launch_bounds__(ITC_BLOCK) k_itc_person(Dev d, const float* infness, const uint16_t* rec, const int8_t* infstat,
                                                         const double* prop_age, double* partial)
{
    __shared__ double red[ITC_NSUM][ITC_BLOCK];
    const cs_params* P = d.P;
    const double dt = P->model_time_step;
    double acc[ITC_NSUM];
    for (int q = 0; q < ITC_NSUM; q++) acc[q] = 0;
    for (long long pp = (long long)blockIdx.x * blockDim.x + threadIdx.x; pp < d.N; pp += (long long)gridDim.x * blockDim.x) {
        const int p = (int)pp;
        const int a = age_group(d, p);
        const double inf = (double)infness[p];
        const int rp = (int)rec[p];
        const bool cp = care(d, p);
        if (P->do_households) {
            double hi = P->no_infectiousness_sd_in_hh ? ((infness[p] < 0) ? P->sympt_infectiousness : P->asympt_infectiousness) : fabs(inf);
            if (cp) hi *= P->care_home_resident_household_scaling;
            const int h = d.hh[p];
            hi *= dt * P->household_trans * P->household_denom_lookup[d.hnhr[h] - 1];
            const double ps = itc_survive(P, 1.0, hi, 0, rp);
            const int f0 = d.hfirst[h];
            for (int m = f0; m < f0 + d.hnh[h]; m++)
                if ((!infstat || infstat[m] == CS_SUSCEPTIBLE) && m != p)
                    acc[ITC_HH] += (1 - ps) * P->age_susceptibility[a] * (care(d, m) ? P->care_home_resident_household_scaling : 1.0);
            acc[ITC_HH_DENOM] += (double)(d.hnhr[h] - 1);
        }
        const double lsd = (P->latent_to_sympt_delay > rp * dt) ? rp * dt : P->latent_to_sympt_delay;
        const int nds = (int)(lsd / dt);
        {
            double si = fabs(inf) * P->relative_spatial_contact[a] * (cp ? P->care_home_resident_spatial_scaling : 1.0) * dt;
            if (prop_age) {  // P.PropAgeGroup-weighted spatial contact (src/SetupModel.cpp:731-742)
                double av = 0;
                const double* pa = prop_age + (size_t)d.padunit[p] * CS_NUM_AGE_GROUPS;
                for (int b = 0; b < CS_NUM_AGE_GROUPS; b++) av += pa[b] * P->waifw_spatial[b][a];
                si *= av;
            }
            int t;
            for (t = 0; t < nds && t < CS_MAX_INFECTIOUS_STEPS; t++) acc[ITC_SPATIAL] += si * P->infectiousness[t];
            si *= (infness[p] < 0) ? P->sympt_spatial_contact_rate : 1;
            for (; t < rp && t < CS_MAX_INFECTIOUS_STEPS; t++) acc[ITC_SPATIAL] += si * P->infectiousness[t];
        }
        if (P->do_places)
            for (int pt = 0; pt < d.NPT; pt++) {
                if (pt == P->hotel_place_type) continue;
                const int k = d.plinks[(size_t)p * CS_MAX_PLACE_TYPES + pt];
                if (k < 0) continue;
                const double pinf = fabs(inf) * dt * P->place_type_trans[pt];
                double scaled = pinf / P->place_type_group_size_param1[pt];
                double ps = itc_survive(P, 1.0, scaled, 0, nds);
                const int g = d.pglinks[(size_t)p * CS_MAX_PLACE_TYPES + pt];
                const double ngroup = (double)((long long)d.pl_gsz[pt][d.pl_gs0[pt][k] + g] - 1);
                const double sympt_rate = ((infness[p] < 0) && !cp)
                    ? (P->sympt_place_type_contact_rate[pt] * (1 - P->sympt_place_type_withdrawal_prop[pt])) : 1;
                scaled *= sympt_rate;
                ps = itc_survive(P, ps, scaled, nds, rp);
                const double ps_group = ps;
                const double n = (double)d.pl_n[pt][k];
                scaled = P->place_type_prop_between_group_links[pt] * pinf / n;
                if (cp)
                    scaled *= (1.0 - P->care_home_prop_residents) + P->care_home_prop_residents *
                              (P->care_home_worker_group_scaling * ((n - 1) - ngroup) + ngroup) / (n - 1);
                ps = itc_survive(P, 1.0, scaled, 0, nds);
                scaled *= sympt_rate;
                ps = itc_survive(P, ps, scaled, nds, rp);
                acc[ITC_PLACE0 + pt] += (1 - ps_group * ps) * ngroup + (1 - ps) * (((double)((long long)d.pl_n[pt][k] - 1)) - ngroup);
            }
        acc[ITC_REC_DAYS] += rp * dt;
        acc[ITC_REC_STEPS] += rp;
    }
    for (int q = 0; q < ITC_NSUM; q++) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = ITC_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < ITC_NSUM; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < ITC_NSUM) partial[(size_t)blockIdx.x * ITC_NSUM + threadIdx.x] = red[threadIdx.x][0];
}

// sum of the block partials, one thread per quantity in block order
__global__ void k_itc_final(const double* partial, int nblocks, double* out)
{
    const int q = threadIdx.x;
    if (q >= ITC_NSUM) return;
    double s = 0;
    for (int b = 0; b < nblocks; b++) s += partial[(size_t)b * ITC_NSUM + q];
    out[q] = s;
}

}  // namespace cs
