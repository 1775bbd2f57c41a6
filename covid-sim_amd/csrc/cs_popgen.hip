// cs_popgen.hip -- the population generator on the device: CovidSim's
// SetupPopulation, AssignHouseholdAges, AssignPeopleToPlaces and
// StratifyPlaces (src/SetupModel.cpp:848-2601) as GPU passes, and the C ABI
// of include/covidsim_hip.h (cs_popgen_*).
//
// Every random draw comes from a counter-based stream keyed by the caller's
// seed, the pass and the entity (microcell, household, person, place), so a
// population depends only on the seed, never on the launch shape or the
// order in which concurrent work lands.  The exceptions are the two places
// where the reference itself is a sequential competition for capacity: the
// random capacity decrements and the cell-gravity assignment, whose claims
// are atomic and whose outcome depends on which claimant lands first; their
// law is the reference's (see the comments at each), their draws are not.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/covidsim_hip.h"
#include "cs_rows.h"

namespace pg {

using cs::Dev;
using cs::Stream;

#define PG_AGES 85            // NUM_AGE_GROUPS * AGE_GROUP_WIDTH years
#define PG_MAX_TRIES (1 << 20)  // bound on every rejection loop (the reference has none)
#define PG_MAX_NEAREST 64     // PlaceTypeNearestNeighb above this is refused
#define PG_GRAV_TESTS 1000    // acceptance tests per person before giving up (src/SetupModel.cpp:2398)
#define PG_GRAV_DRAWS 32      // slot draws per person per round
#define PG_GRAV_ROUNDS 400

enum { PG_MCELL = 1, PG_HHSIZE, PG_HHAGE, PG_AGECORR, PG_PLCOUNT, PG_PLLOC, PG_ELIG, PG_SIZE, PG_CAPADJ,
       PG_CAPINC, PG_CAPDEC, PG_NEAR, PG_GRAV, PG_GROUP };
// status bits of the generator (bit 1, cs::FLAG_RANGE, is kernel_num's)
enum { PGF_TRIES = 1 << 8, PGF_AGE = 1 << 9, PGF_CAP = 1 << 10 };

// counter word 0 of a generator stream: the pass (and the place type)
__device__ __forceinline__ uint32_t pw(int sub, int tp) { return cs::cw0((unsigned)(sub | (tp << 8)), cs::PS_POPGEN); }

// generator constants (device copy)
struct GP {
    cs_popgen_params g;
    int32_t inv_age[1000];                 // State.InvAgeDist[0] (src/SetupModel.cpp:1101-1111)
    double prop_places[PG_AGES][CS_MAX_PLACE_TYPES];  // PropPlaces after the conditional rescaling (:1553-1561)
    double corr_f[CS_NUM_AGE_GROUPS], corr_b[CS_NUM_AGE_GROUPS];  // AgeDistCorrF/B (:1300-1324)
};

__device__ __forceinline__ int draw_age(const GP& G, Stream& s) { return G.inv_age[(int)(1000.0 * s.next())]; }

// AssignHouseholdAges (src/SetupModel.cpp:1820-1988) for a household of n,
// in the reference's order of draws (left operand first where it draws two
// in one expression); false when a rejection loop ran PG_MAX_TRIES times
__device__ bool household_ages(const GP& G, Stream& s, int n, int* a)
{
    const cs_popgen_params& P = G.g;
    int tries = 0;
#define PG_TRY if (++tries > PG_MAX_TRIES) return false
    if (!P.do_households) {
        for (int i = 0; i < n; i++) a[i] = draw_age(G, s);
        return true;
    }
    const double span_old = (double)((long long)P.old_pers_age - P.no_child_pers_age + 1);
    const double span_young = (double)((long long)P.young_and_single - P.min_adult_age);
    if (n == 1) {
        if (s.next() < P.one_pers_house_prob_old) {
            do {
                PG_TRY;
                a[0] = draw_age(G, s);
            } while ((a[0] < P.no_child_pers_age) || (s.next() > (((double)a[0]) - P.no_child_pers_age + 1) / span_old));
        } else if ((P.one_pers_house_prob_young > 0) && (s.next() < P.one_pers_house_prob_young / (1 - P.one_pers_house_prob_old))) {
            do {
                PG_TRY;
                a[0] = draw_age(G, s);
            } while ((a[0] > P.young_and_single) || (a[0] < P.min_adult_age) ||
                     (s.next() > 1 - P.young_and_single_slope * (((double)a[0]) - P.min_adult_age) / span_young));
        } else {
            do {
                PG_TRY;
            } while ((a[0] = draw_age(G, s)) < P.min_adult_age);
        }
    } else if (n == 2) {
        if (s.next() < P.two_pers_house_prob_old) {
            do {
                PG_TRY;
                a[0] = draw_age(G, s);
            } while ((a[0] < P.no_child_pers_age) || (s.next() > (((double)a[0]) - P.no_child_pers_age + 1) / span_old));
            do {
                PG_TRY;
                a[1] = draw_age(G, s);
            } while ((a[1] > a[0] + P.max_mf_partner_age_gap) || (a[1] < a[0] - P.max_fm_partner_age_gap) ||
                     (a[1] < P.no_child_pers_age) || (s.next() > (((double)a[1]) - P.no_child_pers_age + 1) / span_old));
        } else if (s.next() < P.one_child_two_pers_prob / (1 - P.two_pers_house_prob_old)) {
            do {
                PG_TRY;
            } while ((a[0] = draw_age(G, s)) > P.max_child_age);
            do {
                PG_TRY;
                a[1] = draw_age(G, s);
            } while ((a[1] > a[0] + P.max_parent_age_gap) || (a[1] < a[0] + P.min_parent_age_gap) || (a[1] < P.min_adult_age));
        } else if ((P.two_pers_house_prob_young > 0) &&
                   (s.next() < P.two_pers_house_prob_young / (1 - P.two_pers_house_prob_old - P.one_child_two_pers_prob))) {
            do {
                PG_TRY;
                a[0] = draw_age(G, s);
            } while ((a[0] < P.min_adult_age) || (a[0] > P.young_and_single) ||
                     (s.next() > 1 - P.young_and_single_slope * (((double)a[0]) - P.min_adult_age) / span_young));
            do {
                PG_TRY;
                a[1] = draw_age(G, s);
            } while ((a[1] > a[0] + P.max_mf_partner_age_gap) || (a[1] < a[0] - P.max_fm_partner_age_gap) || (a[1] < P.min_adult_age));
        } else {
            do {
                PG_TRY;
                a[0] = draw_age(G, s);
            } while (a[0] < P.min_adult_age);
            do {
                PG_TRY;
                a[1] = draw_age(G, s);
            } while ((a[1] > a[0] + P.max_mf_partner_age_gap) || (a[1] < a[0] - P.max_fm_partner_age_gap) || (a[1] < P.min_adult_age));
        }
    } else {
        int nc;
        if (n == 3) {
            if ((P.zero_child_three_pers_prob > 0) || (P.two_child_three_pers_prob > 0))
                nc = (s.next() < P.zero_child_three_pers_prob) ? 0 : ((s.next() < P.two_child_three_pers_prob) ? 2 : 1);
            else
                nc = 1;
        } else if (n == 4) {
            nc = (s.next() < P.one_child_four_pers_prob) ? 1 : 2;
        } else if (n == 5) {
            nc = (s.next() < P.three_child_five_pers_prob) ? 3 : 2;
        } else {
            nc = n - 2 - (int)(3 * s.next());
        }
        if (nc <= 0) {  // three adults
            do {
                PG_TRY;
                a[0] = draw_age(G, s);
                a[1] = draw_age(G, s);
            } while ((a[1] < P.min_adult_age) || (a[0] < P.min_adult_age));
            do {
                PG_TRY;
                a[2] = draw_age(G, s);
            } while ((a[2] >= a[1] + P.max_mf_partner_age_gap) || (a[2] < a[1] - P.max_fm_partner_age_gap));
        } else {
            int k;
            do {
                PG_TRY;
                a[0] = 0;
                for (int i = 1; i < nc; i++)
                    a[i] = a[i - 1] + 1 + (int)cs::poisson(s, (double)((long long)P.mean_child_age_gap - 1));
                const int base = draw_age(G, s);
                a[0] = base - a[(int)(s.next() * ((double)nc))];
                for (int i = 1; i < nc; i++) a[i] += a[0];
                k = (((nc == 1) && (s.next() < P.one_child_prob_youngest_under_five)) ||
                     ((nc == 2) && (s.next() < P.two_children_prob_youngest_under_five)) ||
                     ((nc > 2) && (s.next() < P.prob_youngest_child_under_five)))
                        ? 5 : P.max_child_age;
            } while ((a[0] < 0) || (a[0] > k) || (a[nc - 1] > P.max_child_age));
            int j = a[nc - 1] - a[0] - (P.max_parent_age_gap - P.min_parent_age_gap);
            if (j > 0) j += P.max_parent_age_gap;
            else j = P.max_parent_age_gap;
            do {
                PG_TRY;
                a[nc] = draw_age(G, s);
                k = a[nc - 1];
            } while ((a[nc] > a[0] + j) || (a[nc] < k + P.min_parent_age_gap) || (a[nc] < P.min_adult_age));
            // PROP_OTHER_PARENT_AWAY is 0 (src/SetupModel.cpp:1818): the test
            // always passes, but draws its uniform
            if ((n > nc + 1) && (s.next() > 0.0)) {
                do {
                    PG_TRY;
                    a[nc + 1] = draw_age(G, s);
                } while ((a[nc + 1] > a[nc] + P.max_mf_partner_age_gap) || (a[nc + 1] < a[nc] - P.max_fm_partner_age_gap) ||
                         (a[nc + 1] > a[0] + j) || (a[nc + 1] < k + P.min_parent_age_gap) || (a[nc + 1] < P.min_adult_age));
            }
            if (n > nc + 2) {
                j = ((a[nc + 1] > a[nc]) ? a[nc + 1] : a[nc]) + P.older_gen_gap;
                if (j >= PG_AGES) j = PG_AGES - 1;
                if (j < P.no_child_pers_age) j = P.no_child_pers_age;
                for (int i = nc + 2; i < n; i++) {
                    do {
                        PG_TRY;
                    } while ((a[i] = draw_age(G, s)) < j);
                }
            }
        }
    }
#undef PG_TRY
    return true;
}

__device__ __forceinline__ void flag(unsigned* st, unsigned f) { atomicOr(st, f); }

// ---------------------------------------------------------------- counts
// One level of a multinomial split as a tree of binomials: node k of level
// L covers leaves [k span, (k + 1) span) of a power-of-two range and sends
// Binomial(n, w_left / w) of its n to its left half.  W is the prefix sum of
// the leaf weights (nw + 1 entries; leaves beyond nw weigh nothing).  The
// reference's sequential conditional binomials over microcells
// (src/SetupModel.cpp:1136-1152, :1529-1535) draw the same multinomial.
__global__ void k_split(const double* W, int nw, int depth, int L, const int32_t* cin, int32_t* cout, uint64_t seed, int sub, int tp)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (1 << L)) return;
    const long long span = 1ll << (depth - L);
    const long long a = (long long)k * span, mid = a + span / 2, b = a + span;
    const double Wa = W[min(a, (long long)nw)], Wm = W[min(mid, (long long)nw)], Wb = W[min(b, (long long)nw)];
    const int32_t n = cin[k];
    int32_t nl = 0;
    if (n > 0 && Wb > Wa) {
        double p = (Wm - Wa) / (Wb - Wa);
        if (p >= 1.0) nl = n;
        else if (p > 0.0) {
            Stream s(seed, pw(sub, tp), (uint32_t)k, (uint32_t)L);
            nl = cs::binomial(s, n, p);
        }
    }
    cout[2 * k] = nl;
    cout[2 * k + 1] = n - nl;
}

// ---------------------------------------------------------------- households
// household sizes of a microcell of n persons (src/SetupModel.cpp:1272-1279):
// a household never holds more than the persons the microcell has left
__device__ __forceinline__ int hh_size(const GP& G, Stream& s, int left)
{
    int m = 1;
    if (G.g.do_households) {
        const double u = s.next();
        while ((u > G.g.household_size_cdf[m - 1]) && (m < left) && (m < CS_MAX_HOUSEHOLD_SIZE)) m++;
    }
    return m;
}

// per populated microcell (in McellLookup order): its number of households
__global__ void k_hh_count(const GP* G, const int32_t* mcl, int nmc, const int32_t* m_n, int32_t* nhh, uint64_t seed)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nmc) return;
    const int j = mcl[q], n = m_n[j];
    Stream s(seed, pw(PG_HHSIZE, 0), (uint32_t)j, 0);
    int c = 0;
    for (int k = 0; k < n; c++) k += hh_size(*G, s, n - k);
    nhh[q] = c;
}

// the same sizes again, written: households and their persons (person and
// household numbers follow McellLookup order, src/SetupModel.cpp:1264-1298)
__global__ void k_hh_write(const GP* G, const int32_t* mcl, int nmc, const int32_t* m_n, const int64_t* poff, const int64_t* hoff,
                           uint64_t seed, int32_t* p_hh, int32_t* p_mcell, int32_t* p_pcell, int32_t* h_first, uint16_t* h_nh)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nmc) return;
    const cs_popgen_params& P = G->g;
    const int j = mcl[q], n = m_n[j];
    const int pcell = ((j / P.mch) / P.nmcl) * P.nch + ((j % P.mch) / P.nmcl);
    Stream s(seed, pw(PG_HHSIZE, 0), (uint32_t)j, 0);
    int64_t p = poff[q], h = hoff[q];
    for (int k = 0; k < n;) {
        const int m = hh_size(*G, s, n - k);
        h_first[h] = (int32_t)p;
        h_nh[h] = (uint16_t)m;
        for (int i = 0; i < m; i++) {
            p_hh[p + i] = (int32_t)h;
            p_mcell[p + i] = j;
            p_pcell[p + i] = pcell;
        }
        p += m;
        h++;
        k += m;
    }
}

// per household: its location in its microcell, then its ages
// (src/SetupModel.cpp:1300-1335); one stream per household
__global__ void k_hh_ages(const GP* G, int64_t nh, const int32_t* h_first, const uint16_t* h_nh, const int32_t* p_mcell,
                          uint64_t seed, float2* h_loc, uint8_t* p_age, unsigned* status)
{
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const cs_popgen_params& P = G->g;
    const int f = h_first[h], n = h_nh[h];
    const int j = p_mcell[f];
    Stream s(seed, pw(PG_HHAGE, 0), (uint32_t)h, (uint32_t)(h >> 32));
    const double x = (double)(j / P.mch), y = (double)(j % P.mch);
    const double xh = P.mc_width * (s.next() + x);
    const double yh = P.mc_height * (s.next() + y);
    h_loc[h] = make_float2((float)xh, (float)yh);
    int a[CS_MAX_HOUSEHOLD_SIZE + 2];
    if (!household_ages(*G, s, n, a)) {
        flag(status, PGF_TRIES);
        for (int i = 0; i < n; i++) a[i] = 0;
    }
    for (int i = 0; i < n; i++) p_age[f + i] = (uint8_t)a[i];
}

// AssignHouseholdAges alone for given sizes (household k takes household
// k's stream, after its two location draws)
__global__ void k_t_hh_ages(const GP* G, int n, const int32_t* sizes, const int64_t* off, uint64_t seed, uint8_t* ages, unsigned* status)
{
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= n) return;
    Stream s(seed, pw(PG_HHAGE, 0), (uint32_t)h, 0);
    (void)s.next();
    (void)s.next();
    int a[CS_MAX_HOUSEHOLD_SIZE + 2];
    const int m = sizes[h];
    if (!household_ages(*G, s, m, a)) {
        flag(status, PGF_TRIES);
        for (int i = 0; i < m; i++) a[i] = 0;
    }
    for (int i = 0; i < m; i++) ages[off[h] + i] = (uint8_t)a[i];
}

// ages by year
__global__ void k_age_hist(int64_t n, const uint8_t* age, unsigned long long* hist)
{
    __shared__ unsigned h[PG_AGES + 1];
    for (int i = threadIdx.x; i <= PG_AGES; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[min((int)age[i], PG_AGES)], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i <= PG_AGES; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
}

// the age correction (src/SetupModel.cpp:1326-1342): one age group up or
// down with the corrections' probabilities
__global__ void k_age_correct(const GP* G, int64_t n, uint8_t* age, uint64_t seed, unsigned* status)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int a = age[i];
    const int j = a / 5;
    Stream s(seed, pw(PG_AGECORR, 0), (uint32_t)i, (uint32_t)(i >> 32));
    const double u = s.next();
    if (u < G->corr_f[j]) a += 5;
    else if (u < G->corr_f[j] + G->corr_b[j]) a -= 5;
    if (a < 0 || a >= PG_AGES) {
        flag(status, PGF_AGE);
        a = min(max(a, 0), PG_AGES - 1);
    }
    age[i] = (uint8_t)a;
}

// ---------------------------------------------------------------- places
// each microcell's place numbers and their microcell (places are numbered
// microcell by microcell, src/SetupModel.cpp:1523-1550)
__global__ void k_place_mcell(int nm, const int64_t* off, int32_t* pl_mcell)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nm) return;
    for (int64_t k = off[j]; k < off[j + 1]; k++) pl_mcell[k] = j;
}

// a place's location: uniform in its microcell
__global__ void k_place_loc(const GP* G, int npl, int tp, const int32_t* pl_mcell, uint64_t seed, float2* loc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npl) return;
    const cs_popgen_params& P = G->g;
    const int j = pl_mcell[k];
    Stream s(seed, pw(PG_PLLOC, tp), (uint32_t)k, 0);
    const double x = (double)(j / P.mch), y = (double)(j % P.mch);
    const double xh = P.mc_width * (s.next() + x);
    const double yh = P.mc_height * (s.next() + y);
    loc[k] = make_float2((float)xh, (float)yh);
}

// who may join a place of type tp (src/SetupModel.cpp:2022-2041): with
// probability PropPlaces[age][tp], if linked to no earlier type
__global__ void k_eligible(const GP* G, int64_t n, int tp, const uint8_t* age, const int32_t* links, uint64_t seed, uint8_t* flags)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double pp = G->prop_places[age[i]][tp];
    bool f = false;
    if (pp > 0) {
        Stream s(seed, pw(PG_ELIG, tp), (uint32_t)i, (uint32_t)(i >> 32));
        f = s.next() < pp;
    }
    for (int k = 0; k < tp && f; k++)
        if (links[i * CS_MAX_PLACE_TYPES + k] >= 0) f = false;
    flags[i] = f ? 1 : 0;
}

// place capacities (src/SetupModel.cpp:2070-2111): Poisson, lognormal or
// offset power law by the type's parameters, at most USHRT_MAX - 1
__global__ void k_place_size(const GP* G, int npl, int tp, uint64_t seed, int32_t* cap, unsigned* status)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npl) return;
    const cs_popgen_params& P = G->g;
    Stream s(seed, pw(PG_SIZE, tp), (uint32_t)k, 0);
    long long j;
    if (P.place_size_power[tp] == 0 && P.place_size_sd[tp] == 0) {
        j = 1 + (long long)cs::poisson(s, P.place_mean_size[tp] - 1);
    } else if (P.place_size_power[tp] == 0) {
        // gen_lognormal (src/Rand.cpp:1098-1107) until within [min, max]
        const double mu = P.place_mean_size[tp], sg = P.place_size_sd[tp];
        const double v = 1 + ((sg * sg) / (mu * mu));
        const double loc = log(mu / sqrt(v)), sc = sqrt(log(v));
        int tries = 0;
        do {
            if (++tries > PG_MAX_TRIES) {
                flag(status, PGF_TRIES);
                j = (long long)P.place_size_min[tp];
                break;
            }
            j = (long long)exp(loc + sc * cs::normal(s));
        } while (j < P.place_size_min[tp] || j > P.place_size_max[tp]);
    } else {
        const double off = P.place_size_offset[tp], pw_ = P.place_size_power[tp];
        const double sm = pow(off / (off + P.place_size_max[tp] - P.place_size_min[tp]), pw_);
        j = (long long)floor(off * pow((1 - sm) * s.next() + sm, -1 / pw_) + P.place_size_min[tp] - off);
    }
    if (j > 65534) j = 65534;
    if (j < 0) j = 0;
    cap[k] = (int32_t)j;
}

// the capacity adjustment (src/SetupModel.cpp:2188-2193): 1 + Poisson(43/40 c - 1)
__global__ void k_cap_adjust(int npl, int tp, uint64_t seed, int32_t* cap)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npl) return;
    Stream s(seed, pw(PG_CAPADJ, tp), (uint32_t)k, 0);
    const double m = ((double)cap[k]) * 43 / 40 - 1;
    long long c = 1 + (long long)cs::poisson(s, m > 0 ? m : 0.0);
    cap[k] = (int32_t)min(c, 65534ll);
}

// the total brought to its target (src/SetupModel.cpp:2194-2207): j2
// increments, each of a uniform place, or -j2 decrements, each of a uniform
// place among those at 2 or more.  Each change draws its own places and is
// applied atomically, so each lands on a uniform place of those eligible
// when it applies, as in the reference's sequential loop.
__global__ void k_cap_change(int64_t nchange, int npl, int tp, int up, uint64_t seed, int32_t* cap, unsigned* status)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nchange) return;
    Stream s(seed, pw(up ? PG_CAPINC : PG_CAPDEC, tp), (uint32_t)q, (uint32_t)(q >> 32));
    if (up) {
        const int k = min((int)(((double)npl) * s.next()), npl - 1);
        if (atomicAdd(&cap[k], 1) >= 65534) atomicSub(&cap[k], 1);
        return;
    }
    for (int tries = 0; tries < PG_MAX_TRIES; tries++) {
        const int k = min((int)(((double)npl) * s.next()), npl - 1);
        int old = atomicAdd(&cap[k], 0);
        while (old >= 2) {
            const int got = atomicCAS(&cap[k], old, old - 1);
            if (got == old) return;
            old = got;
        }
    }
    flag(status, PGF_CAP);
}

// the nearest-places choice (src/SetupModel.cpp:2250-2350): a square spiral
// over microcells from the person's own, keeping the nn places of the
// person's country with the largest kernel weight, until nn are found and
// the spiral's side reaches 4; then one of them with probability
// proportional to its weight.  kd carries the place type's kernel tables.
__global__ void k_nearest(Dev kd, const GP* G, int64_t nel, const int32_t* elig, int tp, int nn, const int32_t* p_hh,
                          const int32_t* p_mcell, const float2* h_loc, const int32_t* mc_country, const int64_t* pl_off,
                          const float2* pl_loc, uint64_t seed, int32_t* links)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nel) return;
    const cs_popgen_params& P = G->g;
    const int i = elig[q];
    const int mcw = P.mcw, mch = P.mch;
    const int ic = p_mcell[i];
    int px = ic / mch, py = ic % mch;
    const int country = mc_country[ic];
    const float2 hl = h_loc[p_hh[i]];
    int near_id[PG_MAX_NEAREST];
    double near_w[PG_MAX_NEAREST];
    int dir = 0, side = 1, k = 0, turn = 0, step = 0;  // dir: 0 right (+x), 1 up (-y), 2 left, 3 down
    while (((k < nn) || (side < 4)) && (side < mcw)) {
        if (px >= 0 && py >= 0 && px < mcw && py < mch) {
            const int c = px * mch + py;
            if (mc_country[c] == country) {
                for (int64_t pl = pl_off[c]; pl < pl_off[c + 1]; pl++) {
                    const double w = cs::kernel_num(kd, cs::dist2_ll(kd.P, hl, pl_loc[pl]));
                    if (w > 0) {
                        if (k < nn) {
                            near_id[k] = (int)pl;
                            near_w[k] = w;
                            k++;
                        } else {
                            double t = 1e10;
                            int j2 = 0;
                            for (int i2 = 0; i2 < nn; i2++)
                                if (near_w[i2] < t) {
                                    t = near_w[i2];
                                    j2 = i2;
                                }
                            if (w > t) {
                                near_w[j2] = w;
                                near_id[j2] = (int)pl;
                            }
                        }
                    }
                }
            }
        }
        switch (dir) {
        case 0: px++; break;
        case 1: py--; break;
        case 2: px--; break;
        default: py++; break;
        }
        step = (step + 1) % side;
        if (step == 0) {
            dir = (dir + 1) & 3;
            turn = (turn + 1) % 2;
            if (turn == 0) side++;
        }
    }
    if (k == 0) return;  // no place within reach: the person stays unplaced
    for (int i2 = 1; i2 < k; i2++) near_w[i2] += near_w[i2 - 1];
    const double tot = near_w[k - 1];
    Stream s(seed, pw(PG_NEAR, tp), (uint32_t)i, 0);
    const double u = s.next();
    int pick = near_id[k - 1];
    for (int i2 = 0; i2 < k; i2++)
        if (u < near_w[i2] / tot) {
            pick = near_id[i2];
            break;
        }
    links[(size_t)i * CS_MAX_PLACE_TYPES + tp] = pick;
}

// ---- the cell-gravity assignment (src/SetupModel.cpp:2353-2420) ----------
// Capacity slots: each cell lists its places of the type, each repeated by
// its capacity.  The cell tables are built from the lists' lengths
// (UpdateProbs(1): S0 = S); in every round each unplaced person draws a
// destination cell from its own cell's row, a uniform slot among the S0
// slots of that cell's list, and tests kernel(dist) / max_trans; a slot
// taken since the tables were built is a miss (the reference draws among the
// S slots still free, so P(free) = S / S0 here stands for its S / S0
// acceptance factor); an accepted slot goes to the first person whose atomic
// claim lands.  Taken slots stay in the lists, so the law holds however
// many rounds share one build; the lists are compacted and the tables
// rebuilt once half the slots of the last build are taken.  A person gives
// up after PG_GRAV_TESTS acceptance tests, as the reference's loop does
// after 1000.

__global__ void k_place_cell(const GP* G, int npl, const float2* loc, double cell_w, double cell_h, const int32_t* cell_map,
                             int32_t* key, unsigned* status)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npl) return;
    const int g = G->g.nch * ((int)(loc[k].x / cell_w)) + ((int)(loc[k].y / cell_h));
    const int c = (g >= 0 && g < G->g.ncw * G->g.nch) ? cell_map[g] : -1;
    if (c < 0) flag(status, PGF_CAP);
    key[k] = max(c, 0);
}

__global__ void k_slot_fill(int npl, const int32_t* sorted_place, const int32_t* cap, const int64_t* slot_off, int32_t* slots)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npl) return;
    const int p = sorted_place[k];
    for (int64_t s = slot_off[k], e = slot_off[k] + cap[p]; s < e; s++) slots[s] = p;
}

__global__ void k_slot_count(int npl, const int32_t* key, const int32_t* cap, int32_t* S)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npl) return;
    atomicAdd(&S[key[k]], cap[k]);
}

__global__ void k_rows(Dev gd)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= gd.NC) return;
    if (gd.ck) cs::update_row_ck(gd, 0, j);
    else cs::update_row(gd, 0, j);
}

// the rows' bucket index after checkpointed rows: a wave per row
__global__ void __launch_bounds__(256) k_rows_index(Dev gd)
{
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i < (size_t)gd.NC) cs::ck_index_row(gd, i);
}

__global__ void k_build_mt(Dev gd, float* out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)gd.NC * gd.NC) return;
    const int l = (int)(i / gd.NC), m = (int)(i % gd.NC);
    out[i] = cs::max_trans_of<true>(gd, l, m);
}

__global__ void k_grav_round(Dev gd, int64_t nact, const int32_t* act, int tp, int round, const int32_t* p_hh,
                             const int32_t* p_pcell, const int32_t* cell_map, const float2* h_loc, const float2* pl_loc,
                             const int64_t* cell_slot, int32_t* slots, int32_t* tests, uint64_t seed, int32_t* links,
                             unsigned long long* claims)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nact) return;
    const int i = act[q];
    const int c = cell_map[p_pcell[i]];
    const float2 hl = h_loc[p_hh[i]];
    int used = tests[i];
    Stream s(seed, pw(PG_GRAV, tp), (uint32_t)i, (uint32_t)round);
    for (int a = 0; a < PG_GRAV_DRAWS && used < PG_GRAV_TESTS; a++) {
        int l;
        float mt;
        if (gd.ck) {
            l = cs::select_dest_ck(gd, 0, c, s.next(), &mt);
        } else {
            l = cs::select_dest(gd, (size_t)c, s.next());
            mt = gd.max_trans[(size_t)c * gd.NC + l];
        }
        const int S0 = gd.S0[l];
        if (S0 <= 0) continue;
        const int m = min((int)(s.next() * (double)S0), S0 - 1);
        int32_t* slot = slots + cell_slot[l] + m;
        const int pid = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (pid < 0) continue;  // taken earlier in this round
        used++;
        const double acc = cs::kernel_num(gd, cs::dist2_ll(gd.P, hl, pl_loc[pid])) / mt;
        if (s.next() < acc && atomicCAS(slot, pid, -1) == pid) {
            links[(size_t)i * CS_MAX_PLACE_TYPES + tp] = pid;
            atomicAdd(claims, 1ull);
            break;
        }
    }
    tests[i] = used;
}

// each cell's slots still free, moved to the front in order
__global__ void k_slot_compact(int nc, const int64_t* cell_slot, int32_t* slots, int32_t* S)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int64_t b = cell_slot[c];
    const int n = S[c];
    int w = 0;
    for (int r = 0; r < n; r++) {
        const int v = slots[b + r];
        if (v >= 0) slots[b + w++] = v;
    }
    S[c] = w;
}

__global__ void k_active(int64_t n, const uint8_t* elig, const int32_t* links, int tp, const int32_t* tests, uint8_t* act)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    act[i] = (elig[i] && links[i * CS_MAX_PLACE_TYPES + tp] < 0 && tests[i] < PG_GRAV_TESTS) ? 1 : 0;
}

// ---------------------------------------------------------------- groups
__global__ void k_link_keys(int64_t n, const int32_t* links, int tp, uint8_t* has, int32_t* key)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int l = links[i * CS_MAX_PLACE_TYPES + tp];
    has[i] = l >= 0 ? 1 : 0;
    key[i] = l;
}

__global__ void k_count_members(int64_t n, const int32_t* links, int tp, int32_t* cnt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int l = links[i * CS_MAX_PLACE_TYPES + tp];
    if (l >= 0) atomicAdd(&cnt[l], 1);
}

// StratifyPlaces (src/SetupModel.cpp:2518-2552): 1 + Poisson(n / g - 1)
// groups (one when n < g)
__device__ __forceinline__ int n_groups(Stream& s, int n, double gsz)
{
    if (n <= 0) return 0;
    const double t = ((double)n) / gsz - 1.0;
    return (t < 0) ? 1 : 1 + (int)cs::poisson(s, t);
}

__global__ void k_groups_count(const GP* G, int npl, int tp, const int32_t* cnt, uint64_t seed, int32_t* ng)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npl) return;
    Stream s(seed, pw(PG_GROUP, tp), (uint32_t)k, 0);
    ng[k] = n_groups(s, cnt[k], G->g.place_group_size[tp]);
}

// group sizes 1 + Binomial(left, 1 / (groups left)), the members shuffled
// by the reference's swap loop, then group links by position
__global__ void k_groups_write(const GP* G, int npl, int tp, const int32_t* cnt, const int64_t* mstart, const int64_t* gstart,
                               uint64_t seed, int32_t* members, int32_t* g_start, int32_t* g_size, uint16_t* glinks)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npl) return;
    const int n = cnt[k];
    if (n <= 0) return;
    Stream s(seed, pw(PG_GROUP, tp), (uint32_t)k, 0);
    const int ng = n_groups(s, n, G->g.place_group_size[tp]);
    int32_t* mem = members + mstart[k];
    const int64_t g0 = gstart[k];
    int m = n - ng, l = 0;
    for (int g = 0; g < ng; g++) {
        const double t = 1 / ((double)(ng - g));
        const int extra = (m <= 0) ? 0 : ((t >= 1.0) ? m : cs::binomial(s, m, t));
        g_start[g0 + g] = l;
        g_size[g0 + g] = 1 + extra;
        m -= extra;
        l += 1 + extra;
    }
    for (int i = 0; i < n; i++) {
        const int j = min((int)(((double)n) * s.next()), n - 1);
        const int v = mem[j];
        mem[j] = mem[i];
        mem[i] = v;
    }
    l = 0;
    for (int g = 0; g < ng; g++)
        for (int i = 0; i < g_size[g0 + g]; i++) glinks[(size_t)mem[l++] * CS_MAX_PLACE_TYPES + tp] = (uint16_t)g;
}

__global__ void k_widen(int64_t n, const int32_t* in, int64_t* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) out[i] = (i < n) ? (int64_t)in[i] : 0;
}

// eligible persons of type tp left without a place
__global__ void k_count_unplaced(int64_t n, const uint8_t* elig, const int32_t* links, int tp, unsigned long long* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (elig[i] && links[i * CS_MAX_PLACE_TYPES + tp] < 0) atomicAdd(out, 1ull);
}

__global__ void k_fill_i32(int64_t n, int32_t* p, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_gather_i32(int64_t n, const int32_t* idx, const int32_t* src, int32_t* dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

} // namespace pg

// ======================================================================== host

namespace {

template <class T> struct DArr {
    T* p = nullptr;
    size_t n = 0;
    DArr() = default;
    DArr(const DArr&) = delete;
    DArr& operator=(const DArr&) = delete;
    ~DArr() { release(); }
    void release()
    {
        // work queued on the buffer finishes before it is freed
        if (p) (void)hipDeviceSynchronize();
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t m)
    {
        release();
        n = m;
        return hipMalloc(&p, std::max<size_t>(m, 1) * sizeof(T));
    }
};

unsigned blocks(long long n, int b = 256)
{
    long long g = (n + b - 1) / b;
    return (unsigned)std::max(1ll, g);
}

} // namespace

struct cs_popgen {
    char err[512];
    int device = 0;
    hipStream_t stream = nullptr;
    bool have = false;
    cs_popgen_params g;
    cs_popgen_sizes sz;
    int NPT = 0;
    std::vector<int32_t> cell_index, cell_n, m_adunit;
    DArr<int32_t> p_hh, p_mcell, p_pcell, links, h_first, pl_n[CS_MAX_PLACE_TYPES], pl_mcell[CS_MAX_PLACE_TYPES];
    DArr<int32_t> pl_members[CS_MAX_PLACE_TYPES], pl_gst[CS_MAX_PLACE_TYPES], pl_gsz[CS_MAX_PLACE_TYPES];
    DArr<int64_t> pl_ms[CS_MAX_PLACE_TYPES], pl_gs0[CS_MAX_PLACE_TYPES];
    DArr<uint8_t> p_age;
    DArr<uint16_t> h_nh, glinks;
    DArr<float2> h_loc, pl_loc[CS_MAX_PLACE_TYPES];
    DArr<unsigned> status;
};

namespace {

int pg_err(cs_popgen* g, int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g->err, sizeof(g->err), fmt, ap);
    va_end(ap);
    return code;
}

#define PGCHK(g, x)                                                                         \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return pg_err(g, CS_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)
#define PGALLOC(g, arr, m)                                                                                    \
    do {                                                                                                      \
        if ((arr).alloc(m) != hipSuccess) {                                                                   \
            (void)hipGetLastError();                                                                          \
            return pg_err(g, CS_ENOMEM, "device allocation of %zu elements (%s)", (size_t)(m), #arr);       \
        }                                                                                                     \
    } while (0)

template <class T> int to_dev(cs_popgen* g, DArr<T>& d, const T* h, size_t n)
{
    PGALLOC(g, d, n);
    if (n) PGCHK(g, hipMemcpy(d.p, h, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}
template <class T> int to_host(cs_popgen* g, std::vector<T>& h, const T* d, size_t n)
{
    h.resize(n);
    if (n) PGCHK(g, hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
    return 0;
}

// exclusive sums of n int32 values into n + 1 int64 (the last is the total)
int dev_scan(cs_popgen* g, const int32_t* in, int64_t* out, size_t n)
{
    DArr<int64_t> w;
    PGALLOC(g, w, n + 1);
    pg::k_widen<<<blocks((long long)n + 1), 256, 0, g->stream>>>((int64_t)n, in, w.p);
    PGCHK(g, hipGetLastError());
    void* t = nullptr;
    size_t tb = 0;
    PGCHK(g, hipcub::DeviceScan::ExclusiveSum(t, tb, w.p, out, (int)(n + 1), g->stream));
    DArr<uint8_t> tmp;
    PGALLOC(g, tmp, tb);
    PGCHK(g, hipcub::DeviceScan::ExclusiveSum((void*)tmp.p, tb, w.p, out, (int)(n + 1), g->stream));
    PGCHK(g, hipStreamSynchronize(g->stream));
    return 0;
}

// indices i < n with flag[i] != 0, in order; returns the count in *m
int dev_select(cs_popgen* g, const uint8_t* flags, size_t n, DArr<int32_t>& out, int64_t* m)
{
    if (!out.p || out.n < n) PGALLOC(g, out, n);
    DArr<int32_t> num;
    PGALLOC(g, num, 1);
    hipcub::CountingInputIterator<int32_t> it(0);
    void* t = nullptr;
    size_t tb = 0;
    PGCHK(g, hipcub::DeviceSelect::Flagged(t, tb, it, flags, out.p, num.p, (int)n, g->stream));
    DArr<uint8_t> tmp;
    PGALLOC(g, tmp, tb);
    PGCHK(g, hipcub::DeviceSelect::Flagged((void*)tmp.p, tb, it, flags, out.p, num.p, (int)n, g->stream));
    int32_t c = 0;
    PGCHK(g, hipStreamSynchronize(g->stream));
    PGCHK(g, hipMemcpy(&c, num.p, 4, hipMemcpyDeviceToHost));
    *m = c;
    return 0;
}

// (key, value) pairs sorted by key, stably
int dev_sort_pairs(cs_popgen* g, DArr<int32_t>& keys, DArr<int32_t>& vals, size_t n, int key_bits)
{
    DArr<int32_t> k2, v2;
    PGALLOC(g, k2, n);
    PGALLOC(g, v2, n);
    void* t = nullptr;
    size_t tb = 0;
    PGCHK(g, hipcub::DeviceRadixSort::SortPairs(t, tb, keys.p, k2.p, vals.p, v2.p, (int)n, 0, key_bits, g->stream));
    DArr<uint8_t> tmp;
    PGALLOC(g, tmp, tb);
    PGCHK(g, hipcub::DeviceRadixSort::SortPairs((void*)tmp.p, tb, keys.p, k2.p, vals.p, v2.p, (int)n, 0, key_bits, g->stream));
    PGCHK(g, hipStreamSynchronize(g->stream));
    std::swap(keys.p, k2.p);
    std::swap(vals.p, v2.p);
    return 0;
}

int bits_for(long long n)
{
    int b = 1;
    while ((1ll << b) <= n) b++;
    return b;
}

// multinomial of n over weights with prefix sums W (nw + 1 entries) on the device
int multinomial(cs_popgen* g, const std::vector<double>& W, int nw, int32_t n, uint64_t seed, int sub, int tp, std::vector<int32_t>& out)
{
    int depth = 0;
    while ((1ll << depth) < nw) depth++;
    DArr<double> dW;
    if (int rc = to_dev(g, dW, W.data(), W.size())) return rc;
    DArr<int32_t> a, b;
    PGALLOC(g, a, (size_t)1 << depth);
    PGALLOC(g, b, (size_t)1 << depth);
    PGCHK(g, hipMemcpy(a.p, &n, 4, hipMemcpyHostToDevice));
    for (int L = 0; L < depth; L++) {
        pg::k_split<<<blocks(1ll << L), 256, 0, g->stream>>>(dW.p, nw, depth, L, a.p, b.p, seed, sub, tp);
        PGCHK(g, hipGetLastError());
        std::swap(a.p, b.p);
    }
    PGCHK(g, hipStreamSynchronize(g->stream));
    return to_host(g, out, a.p, (size_t)nw);
}

int check_status(cs_popgen* g)
{
    unsigned st = 0;
    PGCHK(g, hipStreamSynchronize(g->stream));
    PGCHK(g, hipMemcpy(&st, g->status.p, 4, hipMemcpyDeviceToHost));
    if (st & cs::FLAG_RANGE) return pg_err(g, CS_ERANGE, "a household-place distance beyond the kernel table");
    if (st & pg::PGF_TRIES) return pg_err(g, CS_ERANGE, "a rejection loop of the generator ran %d times", PG_MAX_TRIES);
    if (st & pg::PGF_AGE) return pg_err(g, CS_ERANGE, "the age correction moved an age outside [0, %d)", PG_AGES);
    if (st & pg::PGF_CAP) return pg_err(g, CS_ERANGE, "place capacities could not reach their target (or a place lies outside the populated cells)");
    return 0;
}

// host-side constants: InvAgeDist and PropPlaces (src/SetupModel.cpp:1101-1111, :1425-1441)
void host_constants(const cs_popgen_params& p, pg::GP& G, double pp_raw[PG_AGES][CS_MAX_PLACE_TYPES])
{
    G.g = p;
    double cum[CS_NUM_AGE_GROUPS + 1];
    cum[0] = 0;
    for (int i = 1; i <= CS_NUM_AGE_GROUPS; i++) cum[i] = cum[i - 1] + p.prop_age_group[i - 1];
    for (int i = 0, j = 0; i < 1000; i++) {
        double t = ((double)i) / 1000;
        if (t >= cum[j + 1]) j++;
        t = 5 * (((double)j) + (t - cum[j]) / (cum[j + 1] - cum[j]));
        G.inv_age[i] = (int)t;
    }
    G.inv_age[999] = CS_NUM_AGE_GROUPS * 5 - 1;
    double ppc[PG_AGES][CS_MAX_PLACE_TYPES];
    for (int k = 0; k < PG_AGES; k++)
        for (int l = 0; l < CS_MAX_PLACE_TYPES; l++) {
            pp_raw[k][l] = ppc[k][l] = 0;
            if (l >= p.num_place_types) continue;
            for (int b = 0; b < 3; b++)
                if ((k < p.place_age_max[b][l]) && (k >= p.place_age_min[b][l])) pp_raw[k][l] += p.place_prop[b][l];
            if (l == p.hotel_place_type) ppc[k][l] = (l > 0) ? ppc[k][l - 1] : 0;
            else ppc[k][l] = pp_raw[k][l] + ((l > 0) ? ppc[k][l - 1] : 0);
        }
    for (int k = 0; k < PG_AGES; k++)
        for (int l = 0; l < CS_MAX_PLACE_TYPES; l++) {
            G.prop_places[k][l] = pp_raw[k][l];
            if (l >= 1 && l < p.num_place_types && l != p.hotel_place_type) {
                if (ppc[k][l - 1] < 1) G.prop_places[k][l] /= (1 - ppc[k][l - 1]);
                else if (G.prop_places[k][l] != 0) G.prop_places[k][l] = 1.0;
            }
        }
    for (int j = 0; j < CS_NUM_AGE_GROUPS; j++) G.corr_f[j] = G.corr_b[j] = 0;
}

// the reference's age-correction probabilities from the realised age groups
// (src/SetupModel.cpp:1283-1324), one administrative demography
void age_corrections(const cs_popgen_params& p, const unsigned long long* hist, pg::GP& G)
{
    double ad[CS_NUM_AGE_GROUPS] = {0}, F[CS_NUM_AGE_GROUPS + 1] = {0}, B[CS_NUM_AGE_GROUPS + 1] = {0};
    for (int a = 0; a < PG_AGES; a++) ad[a / 5] += (double)hist[a];
    double s = 0.0;
    for (int j = 0; j < CS_NUM_AGE_GROUPS; j++) s += ad[j];
    for (int j = 0; j < CS_NUM_AGE_GROUPS; j++) ad[j] /= s;
    double t = 0;
    s = 0;
    B[0] = 0;
    for (int j = 0; j < CS_NUM_AGE_GROUPS; j++) {
        s = t + ad[j] - p.prop_age_group[j] - B[j];
        if (s > 0) {
            t = F[j] = s;
            B[j + 1] = 0;
        } else {
            t = F[j] = 0;
            B[j + 1] = fabs(s);
        }
        F[j] /= ad[j];
        B[j] /= ad[j];
    }
    for (int j = 0; j < CS_NUM_AGE_GROUPS; j++) {
        G.corr_f[j] = F[j];
        G.corr_b[j] = B[j];
    }
}

int check_params(cs_popgen* g, const cs_popgen_params* p)
{
    if (p->pop_size <= 0 || p->mcw <= 0 || p->mch <= 0 || p->nmcl <= 0 || p->ncw <= 0 || p->nch <= 0)
        return pg_err(g, CS_EINVAL, "empty region or population");
    if (p->ncw * p->nmcl != p->mcw || p->nch * p->nmcl != p->mch)
        return pg_err(g, CS_EINVAL, "microcell grid (%d x %d) is not cells x nmcl", p->mcw, p->mch);
    if ((long long)p->mcw * p->mch >= (1ll << 31)) return pg_err(g, CS_EINVAL, "too many microcells");
    if ((uint64_t)p->pop_size * CS_MAX_PLACE_TYPES >= (1ull << 32)) return pg_err(g, CS_EINVAL, "population too large for 32-bit links");
    if (p->num_place_types < 0 || p->num_place_types > CS_MAX_PLACE_TYPES) return pg_err(g, CS_EINVAL, "bad number of place types");
    if (p->nsp != 0) return pg_err(g, CS_EUNSUPPORTED, "places read from a school file");
    for (int t = 0; t < p->num_place_types; t++) {
        if (p->inhibit_inter_adunit[t]) return pg_err(g, CS_EUNSUPPORTED, "InhibitInterAdunitPlaceAssignment");
        if (p->place_nearest[t] < 0 || p->place_nearest[t] > PG_MAX_NEAREST)
            return pg_err(g, CS_EUNSUPPORTED, "place type %d picks from %d nearest places (at most %d)", t, p->place_nearest[t], PG_MAX_NEAREST);
        if (t != p->hotel_place_type && !(p->place_mean_size[t] > 0)) return pg_err(g, CS_EINVAL, "place type %d: mean size", t);
        if (t != p->hotel_place_type && !(p->place_group_size[t] > 0)) return pg_err(g, CS_EINVAL, "place type %d: group size", t);
    }
    return 0;
}

// the place type's kernel (src/SetupModel.cpp:2230-2235): the region's
// kernel table size and delta, the type's kernel family and parameters
int place_kernel(cs_popgen* g, const cs_params* kp, const cs_popgen_params* pg_, int tp, cs_params& kt,
                 DArr<cs_params>& dP, DArr<double>& kl, DArr<double>& kh)
{
    kt = *kp;
    kt.kernel_type = pg_->place_kernel_type[tp];
    kt.kernel_scale = pg_->place_kernel_scale[tp];
    kt.kernel_shape = pg_->place_kernel_shape[tp];
    kt.kernel_p3 = pg_->place_kernel_p3[tp];
    kt.kernel_p4 = pg_->place_kernel_p4[tp];
    if (kt.kernel_type < 1 || kt.kernel_type > 7) return pg_err(g, CS_EINVAL, "place type %d: unknown kernel type %d", tp, kt.kernel_type);
    const size_t n = (size_t)kt.kernel_size + 1;
    std::vector<double> hl(n), hh(n);
    for (size_t i = 0; i < n; i++) {
        hl[i] = cs::kernel_fn(kt, (int)i * kt.kernel_delta) / 1.0;
        hh[i] = cs::kernel_fn(kt, (int)i * kt.kernel_delta / kt.kernel_expansion) / 1.0;
    }
    int rc;
    if ((rc = to_dev(g, dP, &kt, 1)) || (rc = to_dev(g, kl, hl.data(), n)) || (rc = to_dev(g, kh, hh.data(), n))) return rc;
    return 0;
}

} // namespace

extern "C" {

int cs_popgen_create(cs_popgen** out, int device)
{
    if (!out) return CS_EINVAL;
    cs_popgen* g = new cs_popgen();
    g->err[0] = 0;
    g->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&g->stream) != hipSuccess) {
        (void)hipGetLastError();
        delete g;
        return CS_EHIP;
    }
    if (g->status.alloc(1) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamDestroy(g->stream);
        delete g;
        return CS_ENOMEM;
    }
    *out = g;
    return 0;
}

void cs_popgen_destroy(cs_popgen* g)
{
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

const char* cs_popgen_last_error(const cs_popgen* g) { return g ? g->err : "null generator"; }

int cs_popgen_run(cs_popgen* g, const cs_popgen_params* p, const cs_params* kp, const double* mcell_density,
                  const int32_t* mcell_country, const int32_t* mcell_adunit, uint64_t seed)
{
    if (!g || !p || !kp || !mcell_density || !mcell_country || !mcell_adunit) return CS_EINVAL;
    PGCHK(g, hipSetDevice(g->device));
    g->have = false;
    int rc = check_params(g, p);
    if (rc) return rc;
    g->g = *p;
    g->NPT = p->num_place_types;
    memset(&g->sz, 0, sizeof(g->sz));
    const int NM = p->mcw * p->mch, NCT = p->ncw * p->nch, MPT = CS_MAX_PLACE_TYPES;
    const int64_t N = p->pop_size;
    hipStream_t st = g->stream;
    PGCHK(g, hipMemset(g->status.p, 0, sizeof(unsigned)));

    // constants
    pg::GP G;
    double pp_raw[PG_AGES][CS_MAX_PLACE_TYPES];
    host_constants(*p, G, pp_raw);
    DArr<pg::GP> dG;
    if ((rc = to_dev(g, dG, &G, 1))) return rc;

    // 1. microcell populations: one multinomial over the microcells' mean
    // densities (src/SetupModel.cpp:1136-1152)
    std::vector<double> W(NM + 1, 0.0);
    for (int i = 0; i < NM; i++) {
        const double w = mcell_density[i];
        if (!(w >= 0)) return pg_err(g, CS_EINVAL, "microcell %d: density %g", i, w);
        W[i + 1] = W[i] + w;
    }
    if (!(W[NM] > 0)) return pg_err(g, CS_EINVAL, "no population density");
    std::vector<int32_t> m_n;
    if ((rc = multinomial(g, W, NM, p->pop_size, seed, pg::PG_MCELL, 0, m_n))) return rc;

    // populated microcells in McellLookup order (cell by cell, src/SetupModel.cpp:1162-1184)
    std::vector<int32_t> mcl;
    std::vector<int64_t> poff, cstart;
    g->cell_index.clear();
    g->cell_n.clear();
    std::vector<int32_t> cell_map(NCT, -1);
    {
        int64_t pos = 0;
        cstart.push_back(0);
        for (int i = 0; i < NCT; i++) {
            const int k = (i / p->nch) * p->nmcl * p->mch + (i % p->nch) * p->nmcl;
            int64_t cn = 0;
            for (int l = 0; l < p->nmcl; l++)
                for (int m = 0; m < p->nmcl; m++) {
                    const int j = k + m + l * p->mch;
                    if (m_n[j] > 0) {
                        mcl.push_back(j);
                        poff.push_back(pos);
                        pos += m_n[j];
                        cn += m_n[j];
                    }
                }
            if (cn > 0) {
                cell_map[i] = (int)g->cell_index.size();
                g->cell_index.push_back(i);
                g->cell_n.push_back((int32_t)cn);
            }
        }
        if (pos != N) return pg_err(g, CS_EINVAL, "microcell populations sum to %lld, not %lld", (long long)pos, (long long)N);
    }
    const int nmc = (int)mcl.size(), NC = (int)g->cell_index.size();
    g->m_adunit.assign(mcell_adunit, mcell_adunit + NM);

    // 2. households
    DArr<int32_t> d_mcl, d_mn, d_nhh;
    DArr<int64_t> d_poff, d_hoff;
    if ((rc = to_dev(g, d_mcl, mcl.data(), mcl.size())) || (rc = to_dev(g, d_mn, m_n.data(), m_n.size())) ||
        (rc = to_dev(g, d_poff, poff.data(), poff.size())))
        return rc;
    PGALLOC(g, d_nhh, nmc);
    pg::k_hh_count<<<blocks(nmc), 256, 0, st>>>(dG.p, d_mcl.p, nmc, d_mn.p, d_nhh.p, seed);
    PGCHK(g, hipGetLastError());
    PGALLOC(g, d_hoff, nmc + 1);
    if ((rc = dev_scan(g, d_nhh.p, d_hoff.p, nmc))) return rc;
    int64_t NH = 0;
    PGCHK(g, hipMemcpy(&NH, d_hoff.p + nmc, 8, hipMemcpyDeviceToHost));
    if (NH >= (1ll << 31)) return pg_err(g, CS_EINVAL, "too many households");
    PGALLOC(g, g->p_hh, N);
    PGALLOC(g, g->p_mcell, N);
    PGALLOC(g, g->p_pcell, N);
    PGALLOC(g, g->p_age, N);
    PGALLOC(g, g->h_first, NH);
    PGALLOC(g, g->h_nh, NH);
    PGALLOC(g, g->h_loc, NH);
    pg::k_hh_write<<<blocks(nmc), 256, 0, st>>>(dG.p, d_mcl.p, nmc, d_mn.p, d_poff.p, d_hoff.p, seed, g->p_hh.p, g->p_mcell.p,
                                                g->p_pcell.p, g->h_first.p, g->h_nh.p);
    pg::k_hh_ages<<<blocks(NH), 256, 0, st>>>(dG.p, NH, g->h_first.p, g->h_nh.p, g->p_mcell.p, seed, g->h_loc.p, g->p_age.p,
                                              g->status.p);
    PGCHK(g, hipGetLastError());
    if ((rc = check_status(g))) return rc;

    // 3. the age correction
    DArr<unsigned long long> d_hist;
    PGALLOC(g, d_hist, PG_AGES + 1);
    PGCHK(g, hipMemset(d_hist.p, 0, (PG_AGES + 1) * 8));
    pg::k_age_hist<<<1024, 256, 0, st>>>(N, g->p_age.p, d_hist.p);
    std::vector<unsigned long long> hist;
    PGCHK(g, hipStreamSynchronize(st));
    if ((rc = to_host(g, hist, d_hist.p, PG_AGES + 1))) return rc;
    if (hist[PG_AGES]) return pg_err(g, CS_ERANGE, "household ages beyond %d", PG_AGES - 1);
    if (p->do_correct_age_dist) {
        age_corrections(*p, hist.data(), G);
        PGCHK(g, hipMemcpy(dG.p, &G, sizeof(G), hipMemcpyHostToDevice));
        pg::k_age_correct<<<blocks(N), 256, 0, st>>>(dG.p, N, g->p_age.p, seed, g->status.p);
        PGCHK(g, hipGetLastError());
        if ((rc = check_status(g))) return rc;
        PGCHK(g, hipMemset(d_hist.p, 0, (PG_AGES + 1) * 8));
        pg::k_age_hist<<<1024, 256, 0, st>>>(N, g->p_age.p, d_hist.p);
        PGCHK(g, hipStreamSynchronize(st));
        if ((rc = to_host(g, hist, d_hist.p, PG_AGES + 1))) return rc;
    }

    // 4. places: counts from demand, spread over microcells by population,
    // located uniformly in their microcell (src/SetupModel.cpp:1511-1552)
    PGALLOC(g, g->links, (size_t)N * MPT);
    pg::k_fill_i32<<<blocks(N * MPT), 256, 0, st>>>(N * MPT, g->links.p, -1);
    PGALLOC(g, g->glinks, (size_t)N * MPT);
    PGCHK(g, hipMemsetAsync(g->glinks.p, 0, (size_t)N * MPT * 2, st));
    std::vector<int64_t> pl_off[CS_MAX_PLACE_TYPES];
    DArr<int64_t> d_ploff[CS_MAX_PLACE_TYPES];
    {
        std::vector<double> Wn(NM + 1, 0.0);
        for (int i = 0; i < NM; i++) Wn[i + 1] = Wn[i] + (double)m_n[i];
        for (int t = 0; t < g->NPT; t++) {
            double dem = 0;
            for (int a = 0; a < PG_AGES; a++) dem += (double)hist[a] * pp_raw[a][t];
            const double np = ceil(dem / p->place_mean_size[t]);
            if (np >= (double)(1 << 27)) return pg_err(g, CS_EINVAL, "place type %d: %g places", t, np);
            const int npl = (int)np;
            g->sz.n_places[t] = npl;
            std::vector<int32_t> cnt;
            if (npl > 0) {
                if ((rc = multinomial(g, Wn, NM, npl, seed, pg::PG_PLCOUNT, t, cnt))) return rc;
            } else {
                cnt.assign(NM, 0);
            }
            pl_off[t].assign(NM + 1, 0);
            for (int i = 0; i < NM; i++) pl_off[t][i + 1] = pl_off[t][i] + cnt[i];
            if ((rc = to_dev(g, d_ploff[t], pl_off[t].data(), NM + 1))) return rc;
            PGALLOC(g, g->pl_mcell[t], npl);
            PGALLOC(g, g->pl_loc[t], npl);
            pg::k_place_mcell<<<blocks(NM), 256, 0, st>>>(NM, d_ploff[t].p, g->pl_mcell[t].p);
            pg::k_place_loc<<<blocks(npl), 256, 0, st>>>(dG.p, npl, t, g->pl_mcell[t].p, seed, g->pl_loc[t].p);
            PGCHK(g, hipGetLastError());
        }
    }

    // 5. people to places, type by type (src/SetupModel.cpp:2013-2460)
    DArr<uint8_t> d_elig;
    PGALLOC(g, d_elig, N);
    DArr<int32_t> d_country, d_cellmap, d_cidx, d_cn;
    DArr<int2> d_cxy;
    std::vector<int2> cxy(NC);
    for (int c = 0; c < NC; c++) cxy[c] = make_int2(g->cell_index[c] / kp->nch, g->cell_index[c] % kp->nch);
    if (kp->nch != p->nch) return pg_err(g, CS_EINVAL, "cs_params.nch %d and the generator's nch %d differ", kp->nch, p->nch);
    if ((rc = to_dev(g, d_country, mcell_country, NM)) || (rc = to_dev(g, d_cellmap, cell_map.data(), NCT)) ||
        (rc = to_dev(g, d_cidx, g->cell_index.data(), NC)) || (rc = to_dev(g, d_cn, g->cell_n.data(), NC)) ||
        (rc = to_dev(g, d_cxy, cxy.data(), NC)))
        return rc;
    for (int t = 0; t < g->NPT; t++) {
        if (t == p->hotel_place_type) continue;
        const int npl = g->sz.n_places[t];
        pg::k_eligible<<<blocks(N), 256, 0, st>>>(dG.p, N, t, g->p_age.p, g->links.p, seed, d_elig.p);
        PGCHK(g, hipGetLastError());
        DArr<int32_t> elig;
        int64_t nel = 0;
        if ((rc = dev_select(g, d_elig.p, N, elig, &nel))) return rc;
        if (npl == 0 || nel == 0) {
            g->sz.n_unplaced[t] = nel;
            continue;
        }
        // capacities and their adjustment
        DArr<int32_t> cap;
        PGALLOC(g, cap, npl);
        pg::k_place_size<<<blocks(npl), 256, 0, st>>>(dG.p, npl, t, seed, cap.p, g->status.p);
        pg::k_cap_adjust<<<blocks(npl), 256, 0, st>>>(npl, t, seed, cap.p);
        PGCHK(g, hipGetLastError());
        if ((rc = check_status(g))) return rc;
        std::vector<int32_t> hcap;
        if ((rc = to_host(g, hcap, cap.p, npl))) return rc;
        long long m = 0;
        for (int k = 0; k < npl; k++) m += hcap[k];
        const long long target = (long long)(((double)nel) * 1.125);
        const long long j2 = target - m;
        if (j2 != 0) {
            pg::k_cap_change<<<blocks(j2 > 0 ? j2 : -j2), 256, 0, st>>>(j2 > 0 ? j2 : -j2, npl, t, j2 > 0, seed, cap.p, g->status.p);
            PGCHK(g, hipGetLastError());
            if ((rc = check_status(g))) return rc;
        }
        // the type's kernel
        cs_params kt;
        DArr<cs_params> dP;
        DArr<double> kl, kh;
        if ((rc = place_kernel(g, kp, p, t, kt, dP, kl, kh))) return rc;
        cs::Dev kd;
        memset(&kd, 0, sizeof(kd));
        kd.P = dP.p;
        kd.kl = kl.p;
        kd.kh = kh.p;
        kd.status = g->status.p;
        if (p->place_nearest[t] > 0) {
            pg::k_nearest<<<blocks(nel, 64), 64, 0, st>>>(kd, dG.p, nel, elig.p, t, p->place_nearest[t], g->p_hh.p, g->p_mcell.p,
                                                         g->h_loc.p, d_country.p, d_ploff[t].p, g->pl_loc[t].p, seed, g->links.p);
            PGCHK(g, hipGetLastError());
            if ((rc = check_status(g))) return rc;
        } else {
            // slots: places sorted by cell (stably), each repeated by its capacity
            DArr<int32_t> key, ids;
            PGALLOC(g, key, npl);
            PGALLOC(g, ids, npl);
            pg::k_place_cell<<<blocks(npl), 256, 0, st>>>(dG.p, npl, g->pl_loc[t].p, kp->cell_width, kp->cell_height, d_cellmap.p,
                                                          key.p, g->status.p);
            PGCHK(g, hipGetLastError());
            if ((rc = check_status(g))) return rc;
            DArr<int32_t> S;
            PGALLOC(g, S, NC);
            PGCHK(g, hipMemsetAsync(S.p, 0, (size_t)NC * 4, st));
            pg::k_slot_count<<<blocks(npl), 256, 0, st>>>(npl, key.p, cap.p, S.p);
            {
                std::vector<int32_t> iota(npl);
                for (int k = 0; k < npl; k++) iota[k] = k;
                PGCHK(g, hipMemcpy(ids.p, iota.data(), (size_t)npl * 4, hipMemcpyHostToDevice));
            }
            if ((rc = dev_sort_pairs(g, key, ids, npl, bits_for(NC)))) return rc;
            DArr<int32_t> capsorted;
            PGALLOC(g, capsorted, npl);
            pg::k_gather_i32<<<blocks(npl), 256, 0, st>>>(npl, ids.p, cap.p, capsorted.p);
            DArr<int64_t> slot_off, cell_slot;
            PGALLOC(g, slot_off, npl + 1);
            PGALLOC(g, cell_slot, NC + 1);
            if ((rc = dev_scan(g, capsorted.p, slot_off.p, npl))) return rc;
            if ((rc = dev_scan(g, S.p, cell_slot.p, NC))) return rc;
            int64_t nslots = 0;
            PGCHK(g, hipMemcpy(&nslots, slot_off.p + npl, 8, hipMemcpyDeviceToHost));
            DArr<int32_t> slots;
            PGALLOC(g, slots, nslots);
            pg::k_slot_fill<<<blocks(npl), 256, 0, st>>>(npl, ids.p, cap.p, slot_off.p, slots.p);
            PGCHK(g, hipGetLastError());
            // cell tables over slots (full while they fit, else checkpointed)
            cs::Dev gd = kd;
            gd.NC = NC;
            gd.R = 1;
            gd.cidx = d_cidx.p;
            gd.cxy = d_cxy.p;
            gd.cn = d_cn.p;
            DArr<int32_t> S0;
            DArr<float> mt, ct, tot, ckp;
            DArr<int32_t> inv;
            PGALLOC(g, S0, NC);
            PGALLOC(g, tot, NC);
            gd.S0 = S0.p;
            gd.tot_prob = tot.p;
            const bool full = 2.0 * (double)NC * NC * 4 <= (double)(16ull << 30);
            if (full) {
                PGALLOC(g, mt, (size_t)NC * NC);
                PGALLOC(g, ct, (size_t)NC * NC);
                PGALLOC(g, inv, (size_t)NC * 1025);
                gd.max_trans = mt.p;
                gd.cum_trans = ct.p;
                gd.invcdf = inv.p;
                pg::k_build_mt<<<blocks((long long)NC * NC), 256, 0, st>>>(gd, mt.p);
            } else {
                gd.ck = 8;
                while ((double)NC * NC * 4 / gd.ck + (double)NC * 1025 * 4 > (double)(16ull << 30)) gd.ck *= 2;
                gd.ck_nb = (NC + gd.ck - 1) / gd.ck;
                PGALLOC(g, ckp, (size_t)NC * gd.ck_nb);
                gd.ck_pre = ckp.p;
                PGALLOC(g, inv, (size_t)NC * 1025);  // the rows' bucket index (select_dest_ck's brackets)
                gd.ck_inv = inv.p;
            }
            DArr<int32_t> tests, act;
            PGALLOC(g, tests, N);
            PGCHK(g, hipMemsetAsync(tests.p, 0, (size_t)N * 4, st));
            DArr<uint8_t> aflag;
            PGALLOC(g, aflag, N);
            int64_t nact = nel;
            PGALLOC(g, act, nel);
            PGCHK(g, hipMemcpyAsync(act.p, elig.p, (size_t)nel * 4, hipMemcpyDeviceToDevice, st));
            DArr<unsigned long long> claims;
            PGALLOC(g, claims, 1);
            bool rebuild = true;
            long long built = 0;
            for (int round = 0; round < PG_GRAV_ROUNDS && nact > 0; round++) {
                if (rebuild) {
                    if (round > 0) pg::k_slot_compact<<<blocks(NC), 256, 0, st>>>(NC, cell_slot.p, slots.p, S.p);
                    PGCHK(g, hipMemcpyAsync(S0.p, S.p, (size_t)NC * 4, hipMemcpyDeviceToDevice, st));
                    pg::k_rows<<<blocks(NC, 64), 64, 0, st>>>(gd);
                    if (gd.ck_inv) pg::k_rows_index<<<blocks((long long)NC * 64, 256), 256, 0, st>>>(gd);
                    PGCHK(g, hipMemsetAsync(claims.p, 0, 8, st));
                    std::vector<int32_t> hs;
                    if ((rc = to_host(g, hs, S.p, NC))) return rc;
                    built = 0;
                    for (int c = 0; c < NC; c++) built += hs[c];
                    if (built == 0) break;
                    rebuild = false;
                }
                pg::k_grav_round<<<blocks(nact), 256, 0, st>>>(gd, nact, act.p, t, round, g->p_hh.p, g->p_pcell.p, d_cellmap.p,
                                                                g->h_loc.p, g->pl_loc[t].p, cell_slot.p, slots.p, tests.p, seed,
                                                                g->links.p, claims.p);
                pg::k_active<<<blocks(N), 256, 0, st>>>(N, d_elig.p, g->links.p, t, tests.p, aflag.p);
                PGCHK(g, hipGetLastError());
                if ((rc = dev_select(g, aflag.p, N, act, &nact))) return rc;
                if ((rc = check_status(g))) return rc;
                unsigned long long taken = 0;
                PGCHK(g, hipMemcpy(&taken, claims.p, 8, hipMemcpyDeviceToHost));
                if ((long long)taken * 2 >= built) rebuild = true;
            }
        }
        {
            DArr<unsigned long long> un;
            PGALLOC(g, un, 1);
            PGCHK(g, hipMemsetAsync(un.p, 0, 8, st));
            pg::k_count_unplaced<<<blocks(N), 256, 0, st>>>(N, d_elig.p, g->links.p, t, un.p);
            PGCHK(g, hipGetLastError());
            unsigned long long h = 0;
            PGCHK(g, hipStreamSynchronize(st));
            PGCHK(g, hipMemcpy(&h, un.p, 8, hipMemcpyDeviceToHost));
            g->sz.n_unplaced[t] = (int64_t)h;
        }
    }
    if ((rc = check_status(g))) return rc;

    // 6. members and groups (src/SetupModel.cpp:2468-2552)
    for (int t = 0; t < g->NPT; t++) {
        const int npl = g->sz.n_places[t];
        DArr<uint8_t> has;
        DArr<int32_t> key, who, cnt, ng;
        PGALLOC(g, has, N);
        PGALLOC(g, key, N);
        pg::k_link_keys<<<blocks(N), 256, 0, st>>>(N, g->links.p, t, has.p, key.p);
        PGCHK(g, hipGetLastError());
        int64_t nm = 0;
        if ((rc = dev_select(g, has.p, N, who, &nm))) return rc;
        g->sz.n_members[t] = nm;
        DArr<int32_t> k2;
        PGALLOC(g, k2, nm);
        pg::k_gather_i32<<<blocks(nm), 256, 0, st>>>(nm, who.p, key.p, k2.p);
        if (nm > 0 && (rc = dev_sort_pairs(g, k2, who, nm, bits_for(npl)))) return rc;
        std::swap(g->pl_members[t].p, who.p);
        std::swap(g->pl_members[t].n, who.n);
        PGALLOC(g, cnt, npl);
        PGCHK(g, hipMemsetAsync(cnt.p, 0, (size_t)npl * 4, st));
        pg::k_count_members<<<blocks(N), 256, 0, st>>>(N, g->links.p, t, cnt.p);
        PGALLOC(g, g->pl_ms[t], npl + 1);
        if ((rc = dev_scan(g, cnt.p, g->pl_ms[t].p, npl))) return rc;
        PGALLOC(g, ng, npl);
        pg::k_groups_count<<<blocks(npl), 256, 0, st>>>(dG.p, npl, t, cnt.p, seed, ng.p);
        PGALLOC(g, g->pl_gs0[t], npl + 1);
        if ((rc = dev_scan(g, ng.p, g->pl_gs0[t].p, npl))) return rc;
        int64_t ngr = 0;
        PGCHK(g, hipMemcpy(&ngr, g->pl_gs0[t].p + npl, 8, hipMemcpyDeviceToHost));
        g->sz.n_groups[t] = ngr;
        PGALLOC(g, g->pl_gst[t], ngr);
        PGALLOC(g, g->pl_gsz[t], ngr);
        pg::k_groups_write<<<blocks(npl, 64), 64, 0, st>>>(dG.p, npl, t, cnt.p, g->pl_ms[t].p, g->pl_gs0[t].p, seed,
                                                          g->pl_members[t].p, g->pl_gst[t].p, g->pl_gsz[t].p, g->glinks.p);
        PGCHK(g, hipGetLastError());
        std::swap(g->pl_n[t].p, cnt.p);
        std::swap(g->pl_n[t].n, cnt.n);
    }
    PGCHK(g, hipStreamSynchronize(st));
    if ((rc = check_status(g))) return rc;
    g->sz.n_people = (int32_t)N;
    g->sz.n_households = (int32_t)NH;
    g->sz.n_cells = NC;
    g->sz.n_cells_total = NCT;
    g->sz.n_mcells = NM;
    g->have = true;
    return 0;
}

int cs_popgen_get_sizes(cs_popgen* g, cs_popgen_sizes* out)
{
    if (!g || !out) return CS_EINVAL;
    if (!g->have) return pg_err(g, CS_ESTATE, "no population generated");
    *out = g->sz;
    return 0;
}

int cs_popgen_fetch(cs_popgen* g, const cs_popgen_out* o)
{
    if (!g || !o) return CS_EINVAL;
    if (!g->have) return pg_err(g, CS_ESTATE, "no population generated");
    PGCHK(g, hipSetDevice(g->device));
    PGCHK(g, hipStreamSynchronize(g->stream));
    const size_t N = (size_t)g->sz.n_people, NH = (size_t)g->sz.n_households, NC = (size_t)g->sz.n_cells;
#define PGGET(dst, src, n) \
    if (dst) PGCHK(g, hipMemcpy(dst, src, (n) * sizeof(*(dst)), hipMemcpyDeviceToHost))
    PGGET(o->hh, g->p_hh.p, N);
    PGGET(o->mcell, g->p_mcell.p, N);
    PGGET(o->pcell, g->p_pcell.p, N);
    PGGET(o->age, g->p_age.p, N);
    PGGET(o->place_links, g->links.p, N * CS_MAX_PLACE_TYPES);
    PGGET(o->place_group_links, g->glinks.p, N * CS_MAX_PLACE_TYPES);
    PGGET(o->hh_first, g->h_first.p, NH);
    PGGET(o->hh_nh, g->h_nh.p, NH);
    if (o->hh_loc) PGCHK(g, hipMemcpy(o->hh_loc, g->h_loc.p, NH * sizeof(float2), hipMemcpyDeviceToHost));
#undef PGGET
    if (o->cell_index) memcpy(o->cell_index, g->cell_index.data(), NC * 4);
    if (o->cell_n) memcpy(o->cell_n, g->cell_n.data(), NC * 4);
    if (o->cell_member_start) {
        int64_t s = 0;
        for (size_t c = 0; c < NC; c++) {
            o->cell_member_start[c] = (int32_t)s;
            s += g->cell_n[c];
        }
        o->cell_member_start[NC] = (int32_t)s;
    }
    if (o->cell_members)
        for (size_t i = 0; i < N; i++) o->cell_members[i] = (int32_t)i;  // persons are numbered cell by cell
    if (o->mcell_adunit) memcpy(o->mcell_adunit, g->m_adunit.data(), g->m_adunit.size() * 4);
    for (int t = 0; t < g->NPT; t++) {
        const size_t npl = (size_t)g->sz.n_places[t];
        const size_t nm = (size_t)g->sz.n_members[t], ngr = (size_t)g->sz.n_groups[t];
        if (o->place_n[t] && npl) PGCHK(g, hipMemcpy(o->place_n[t], g->pl_n[t].p, npl * 4, hipMemcpyDeviceToHost));
        if (o->place_mcell[t] && npl) PGCHK(g, hipMemcpy(o->place_mcell[t], g->pl_mcell[t].p, npl * 4, hipMemcpyDeviceToHost));
        if (o->place_loc[t] && npl) PGCHK(g, hipMemcpy(o->place_loc[t], g->pl_loc[t].p, npl * 8, hipMemcpyDeviceToHost));
        if (o->place_members[t] && nm) PGCHK(g, hipMemcpy(o->place_members[t], g->pl_members[t].p, nm * 4, hipMemcpyDeviceToHost));
        if (o->place_group_start[t] && ngr) PGCHK(g, hipMemcpy(o->place_group_start[t], g->pl_gst[t].p, ngr * 4, hipMemcpyDeviceToHost));
        if (o->place_group_size[t] && ngr) PGCHK(g, hipMemcpy(o->place_group_size[t], g->pl_gsz[t].p, ngr * 4, hipMemcpyDeviceToHost));
        std::vector<int64_t> v;
        if (o->place_member_start[t]) {
            int rc = to_host(g, v, g->pl_ms[t].p, npl + 1);
            if (rc) return rc;
            for (size_t k = 0; k <= npl; k++) o->place_member_start[t][k] = (int32_t)v[k];
        }
        if (o->place_groups_start[t]) {
            int rc = to_host(g, v, g->pl_gs0[t].p, npl + 1);
            if (rc) return rc;
            for (size_t k = 0; k <= npl; k++) o->place_groups_start[t][k] = (int32_t)v[k];
        }
    }
    return 0;
}

int cs_popgen_household_ages(cs_popgen* g, const cs_popgen_params* p, uint64_t seed, const int32_t* sizes, int n, uint8_t* ages)
{
    if (!g || !p || !sizes || !ages || n < 0) return CS_EINVAL;
    PGCHK(g, hipSetDevice(g->device));
    std::vector<int64_t> off(n + 1, 0);
    for (int i = 0; i < n; i++) {
        if (sizes[i] < 1 || sizes[i] > CS_MAX_HOUSEHOLD_SIZE) return pg_err(g, CS_EINVAL, "household size %d", sizes[i]);
        off[i + 1] = off[i] + sizes[i];
    }
    pg::GP G;
    double pp_raw[PG_AGES][CS_MAX_PLACE_TYPES];
    host_constants(*p, G, pp_raw);
    DArr<pg::GP> dG;
    DArr<int32_t> ds;
    DArr<int64_t> doff;
    DArr<uint8_t> da;
    int rc;
    if ((rc = to_dev(g, dG, &G, 1)) || (rc = to_dev(g, ds, sizes, (size_t)n)) || (rc = to_dev(g, doff, off.data(), off.size())))
        return rc;
    PGALLOC(g, da, (size_t)off[n]);
    PGCHK(g, hipMemset(g->status.p, 0, 4));
    if (n) pg::k_t_hh_ages<<<blocks(n), 256, 0, g->stream>>>(dG.p, n, ds.p, doff.p, seed, da.p, g->status.p);
    PGCHK(g, hipGetLastError());
    if ((rc = check_status(g))) return rc;
    if (off[n]) PGCHK(g, hipMemcpy(ages, da.p, (size_t)off[n], hipMemcpyDeviceToHost));
    return 0;
}

} // extern "C"
