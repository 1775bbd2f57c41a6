// cs_model.h -- device-side model state and the per-person functions of
// CovidSim's infection step: kernel lookup (src/Kernels.cpp:108-129),
// distances (src/Dist.cpp), hazards (src/CalcInfSusc.cpp), state predicates
// (src/ModelMacros.h), InverseCdf::choose (src/InverseCdf.cpp:60-74).
//
// Layout in HBM.  Static population arrays are shared by all realisations;
// every dynamic per-person array is [n_real][n_people] (realisation-major),
// per-cell arrays [n_real][n_cells], per-place arrays [n_real][n_places].
// A global person index gp = r * n_people + p addresses the dynamic arrays.
//
// The fields the infection kernels read together are packed into records,
// so that a scattered visit to one person costs one cache line instead of
// one line per field: Hot and Tim per (realisation, person), PStat and
// PLinks per person, HRec per household, PlRec per place and GRec per place
// group.  Each packed field keeps an array-like accessor (Fld), so
// d.inf[gp] reads the inf field of record gp.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/covidsim_hip.h"
#include "cs_math.h"

namespace cs {

struct Counters {  // device mirror of cs_counters, updated with atomics
    unsigned long long cumI, cumC, cumDC, cumFC, cumTG, cumSI, nTG;
    unsigned long long cumAC, cumACS, cumAA;
    double sumRad2;
    unsigned long long maxRad2_bits;
    unsigned long long cumItype[CS_INFECT_TYPE_MASK];
    unsigned long long cumIa[CS_NUM_AGE_GROUPS], cumCa[CS_NUM_AGE_GROUPS], cumDa[CS_NUM_AGE_GROUPS];
    unsigned long long cumI_keyworker[2], cumC_keyworker[2];
    unsigned long long sev_prev[CS_NUM_SEV_KINDS], sev_cum[CS_NUM_SEV_KINDS];
    unsigned long long sev_prev_age[CS_NUM_SEV_KINDS][CS_NUM_AGE_GROUPS];
    unsigned long long sev_cum_age[CS_NUM_SEV_KINDS][CS_NUM_AGE_GROUPS];
    unsigned long long cumHQ, cumAH, cumAPC, cumAPA, cumAPCS;
};

// the per-transition counters of the infection and incubation phases are
// staged: CS_TALLY_COPIES copies of each realisation's Counters (a block adds
// to copy blockIdx.x % CS_TALLY_COPIES), folded into the realisation's
// Counters at the end of each phase.  Every wave adds to its realisation's
// few hot counters, and adds from the whole GPU to one line serialise at the
// memory side.
#define CS_TALLY_COPIES 16
#define CS_AD_COPIES 16  // copies of the counters by administrative unit (ad_counter)
static_assert(sizeof(Counters) % 8 == 0, "Counters is a sequence of 8-byte words");

// one realisation's copy of the P and AdUnits fields the trigger logic sets
// (src/CovidSim.cpp:3485-3689); written at samples by one thread
struct IntState {
    int alert;                 // the alert condition has held at a sample
    int trig;                  // State.trigDetectedCases of the last sample
    double ctrl;               // P.ControlPropCasesId
    double ci_start, hq_start, pc_start, pc_prev;
};
// samples of trigger history kept per realisation (cumDC, cumCritical, D at
// sample m in slot m % CS_TRIG_RING); a trigger window that reaches back
// further within the run is an error
#define CS_TRIG_RING 256
// microcell place-closure states (TreatStat, src/Model.h)
enum { MC_UNTREATED = 0, MC_TREATED = 1, MC_DONT_TREAT_AGAIN = 2 };

struct StepScalars {  // per-timestep constants of InfectSweep (src/Sweep.cpp:294-301)
    double t, fp, seas, beta_s, beta_h;
    unsigned ts;
};

struct Clock {  // the device's own step counter for replayed step graphs
    double t;
    int step, sample;
};

struct Chunk {  // up to 64 consecutive entries starting at `first` of cell rc's segment
    int32_t rc, first;
};
enum { WL_SEG = 0, WL_CELLS, WL_DRAWS, WL_N };

// event tallies for the algorithmic-bytes figure.  Every wave of a kernel
// adds its sums at the end, and adds from the whole GPU to one 128-byte line
// serialise at the memory side, so each tally has CS_WORK_COPIES copies, a
// line each: a block adds to copy blockIdx.x % CS_WORK_COPIES (wtally), and
// the host sums the copies.
#define CS_WORK_COPIES 8
struct alignas(128) WorkCell {
    unsigned long long v;
};
struct Work {
    WorkCell n_inf[CS_WORK_COPIES], n_hh_inf[CS_WORK_COPIES], n_hh_member[CS_WORK_COPIES];
    WorkCell n_place_link[CS_WORK_COPIES], n_place_cand[CS_WORK_COPIES];
    WorkCell n_sp_draw[CS_WORK_COPIES], n_sp_susc[CS_WORK_COPIES], n_new[CS_WORK_COPIES];
    WorkCell n_active[CS_WORK_COPIES], n_transition[CS_WORK_COPIES];
    WorkCell n_sp_scan[CS_WORK_COPIES], n_sp_probe[CS_WORK_COPIES];
    WorkCell n_hh_scan[CS_WORK_COPIES];
};
__device__ __forceinline__ void wtally(WorkCell* t, unsigned long long v) { atomicAdd(&t[blockIdx.x & (CS_WORK_COPIES - 1u)].v, v); }
inline unsigned long long wsum(const WorkCell* t)
{
    unsigned long long s = 0;
    for (int i = 0; i < CS_WORK_COPIES; i++) s += t[i].v;
    return s;
}

enum { FLAG_OVERFLOW = 1, FLAG_RANGE = 2, FLAG_BATCH = 4, FLAG_STUCK = 8, FLAG_TRIG = 16, FLAG_LIST = 32 };
// Hot.flags: HOT_HH_DONE, the person's household has no susceptible member
// left (none can return to susceptible, so the household pass may skip it
// for good); HOT_NO_LINKS, the person has no place link (a copy of PLinks
// the place pass can test without reading PLinks); bits HOT_NH_SHIFT and up,
// the size of the person's household (at most CS_MAX_HOUSEHOLD_SIZE), so an
// infector whose household needs no scan reads no household record
enum { HOT_HH_DONE = 1, HOT_NO_LINKS = 2, HOT_NH_SHIFT = 2 };
static_assert((CS_MAX_HOUSEHOLD_SIZE << HOT_NH_SHIFT) <= 0xff, "household size fits the flags byte");

__host__ __device__ __forceinline__ bool no_place_links(const int32_t* lk, int npt)
{
    for (int t = 0; t < npt; t++)
        if (lk[t] >= 0) return false;
    return true;
}

// latent due lists: bucket ts % CS_DUE_RING lists the persons whose latent
// time is ts, appended when they are infected (every latent time lies less
// than CS_DUE_RING steps ahead, checked from the latent-period table)
#define CS_DUE_RING 256
// a bucket is CS_DUE_SHARDS shards, a block appending to shard
// blockIdx.x % CS_DUE_SHARDS; each shard's count sits alone on a 128-byte
// line (CS_DUE_STRIDE words apart): a step's infections append to the few
// buckets of their latent times, and adds from the whole GPU to one line
// serialise at the memory side
#define CS_DUE_STRIDE 32
#define CS_DUE_SHARDS 16

// the event queue is CS_EV_SHARDS shards: a block appends to shard
// blockIdx.x % CS_EV_SHARDS, whose count sits alone on a 128-byte line and
// whose events fill its own 1/CS_EV_SHARDS of the queue (appends from the
// whole GPU to one count serialise at the memory side)
#define CS_EV_SHARDS 8
#define CS_EV_STRIDE 32

// a place link of the place passes whose binomial number of contacts is
// positive (src/Sweep.cpp:425-449), with what the member sampling needs:
// the group, the capped infectiousness and the number of values of the
// pass's stream the count used.  The list is CS_EV_SHARDS shards, as the
// event queue.
struct PlContact {
    int32_t r, i, n, gsize, goff, mstart;
    double Pc;
    uint32_t v;
    uint8_t pt, pass;
    uint16_t pad;
};
static_assert(sizeof(PlContact) == 40, "PlContact layout");

struct Hot {  // per (realisation, person): 16 bytes, one vector load
    int8_t inf;       // InfStat code
    uint8_t flags;    // HOT_HH_DONE, HOT_NO_LINKS, household size << HOT_NH_SHIFT
    int16_t itype;    // infect_type
    uint16_t t_lat, t_as, t_ae;  // latent time, absence start and end
    uint8_t age, pflags;  // copies of the person's PStat fields, read with this record
    float infness;    // infectiousness (sign: symptomatic)
};
struct Tim {  // per (realisation, person): the disease timers and severity, 16 bytes
    uint16_t t_inf, t_rod, t_sari, t_crit, t_step, t_det;  // infection, recovery/death, SARI, critical, stepdown, detected
    uint8_t sev_cur, sev_final;
    uint8_t bits;  // bit0 to_die, bit1 detected
    uint8_t pad;
};
struct ClaimRec {  // per (realisation, person): this step's smallest and largest event keys
    unsigned long long claim, lastk;
};
// PStat.pflags (copied into Hot.pflags): bit 0 care-home resident, bit 1 key
// worker, bit 2 essential (the population's flags; its bit 3, the contact
// tracing app user, has no use here, digital contact tracing being refused
// at the boundary), bits 3-6 the person's
// position in its household (its first member is p - position), bit 7 the
// household's resident count (nhr) differs from its size (else the size
// in Hot.flags gives it)
enum { PF_INPUT = 7, PF_HPOS_SHIFT = 3, PF_HPOS_MASK = 15, PF_NHR_DIFFERS = 0x80 };
static_assert(CS_MAX_HOUSEHOLD_SIZE <= PF_HPOS_MASK + 1, "a household position fits its 4 bits");
struct PStat {  // per person, shared by all realisations
    int32_t hh;
    uint8_t age, pflags;
    uint16_t pad;
    float2 loc;  // the household's location (a copy, so a distance needs no second lookup)
};
struct HRec {  // per household
    int32_t hfirst;
    uint16_t hnh, hnhr;
};
struct PLinks {  // per person: place and group link of each place type
    int32_t lk[CS_MAX_PLACE_TYPES];
    uint16_t g[CS_MAX_PLACE_TYPES];
    uint16_t pad;
};
struct PlRec {  // per place (n_places + 1 records; the last holds the ends)
    int32_t ms, gs0, n, pad;  // member start, first group, member count
};
struct GRec {  // per group of a place
    int32_t gst, gsz;  // group start within the place's members, group size
};
static_assert(sizeof(Hot) == 16 && sizeof(Tim) == 16 && sizeof(ClaimRec) == 16 && sizeof(PStat) == 16 && sizeof(HRec) == 8 && sizeof(PLinks) == 32 &&
              sizeof(PlRec) == 16 && sizeof(GRec) == 8, "record sizes");

// one field of an array of records, indexed like a plain array
template <typename Rec, typename T, size_t OFF>
struct Fld {
    Rec* base;
    __host__ __device__ __forceinline__ T& operator[](size_t i) const
    {
        using B = typename std::conditional<std::is_const<Rec>::value, const char, char>::type;
        return *reinterpret_cast<T*>(reinterpret_cast<B*>(base + i) + OFF);
    }
};
#define CS_FLD(Rec, f) Fld<Rec, decltype(((Rec*)nullptr)->f), offsetof(Rec, f)>
#define CS_CFLD(Rec, f) Fld<const Rec, const decltype(((Rec*)nullptr)->f), offsetof(Rec, f)>

// the [person][place type] link arrays, flat index i * CS_MAX_PLACE_TYPES + t
// (below 2^32: persons x place types)
struct LinkFld {
    const PLinks* base;
    __host__ __device__ __forceinline__ const int32_t& operator[](size_t i) const
    {
        const uint32_t p = (uint32_t)i / CS_MAX_PLACE_TYPES;
        return base[p].lk[(uint32_t)i - p * CS_MAX_PLACE_TYPES];
    }
};
struct GLinkFld {
    const PLinks* base;
    __host__ __device__ __forceinline__ const uint16_t& operator[](size_t i) const
    {
        const uint32_t p = (uint32_t)i / CS_MAX_PLACE_TYPES;
        return base[p].g[(uint32_t)i - p * CS_MAX_PLACE_TYPES];
    }
};

struct Dev {
    int N, NH, NC, NM, NPT, R;
    const cs_params* P;
    const double* kl;
    const double* kh;
    // population (static)
    const PStat* pst;     // [N]
    const HRec* hrec;     // [NH]
    CS_CFLD(PStat, hh) hh;
    CS_CFLD(PStat, age) age;
    CS_CFLD(PStat, pflags) pflags;
    CS_CFLD(PStat, loc) ploc;
    CS_CFLD(HRec, hfirst) hfirst;
    CS_CFLD(HRec, hnh) hnh;
    CS_CFLD(HRec, hnhr) hnhr;
    const PLinks* plk;    // [N]
    LinkFld plinks;
    GLinkFld pglinks;
    const int32_t *pc, *mcell, *cstart, *cn, *cidx, *mstart, *mmem, *cmem;
    const int2* cxy;      // [NC] grid coordinates (cidx / nch, cidx % nch) of each populated cell
    const int32_t* padunit;  // administrative unit of each person's microcell
    int np[CS_MAX_PLACE_TYPES];
    int np_off[CS_MAX_PLACE_TYPES + 1];  // offsets of each type in the per-realisation place arrays
    const int32_t* np_off_d;  // a device copy of np_off, for run-time type indices
    const PlRec* pl_rec[CS_MAX_PLACE_TYPES];
    const GRec* pl_grp[CS_MAX_PLACE_TYPES];
    CS_CFLD(PlRec, n) pl_n[CS_MAX_PLACE_TYPES];
    CS_CFLD(PlRec, ms) pl_ms[CS_MAX_PLACE_TYPES];
    CS_CFLD(PlRec, gs0) pl_gs0[CS_MAX_PLACE_TYPES];
    CS_CFLD(GRec, gst) pl_gst[CS_MAX_PLACE_TYPES];
    CS_CFLD(GRec, gsz) pl_gsz[CS_MAX_PLACE_TYPES];
    const int32_t* pl_mem[CS_MAX_PLACE_TYPES];
    const float* max_trans;
    // dynamic, per realisation
    const uint64_t* seed;
    const double* rbeta;     // [R] per-realisation LocalBeta, or null (cs_params.local_beta)
    const int32_t* rseed_n;  // [R] per-realisation seed infections, or null
    const double* rthr;      // [R] per-realisation intervention trigger thresholds, or null
    Hot* hot;             // [R][N]
    Hot* shot;            // [R][N] by list slot: the Hot record of the person at each slot of an infectious segment (else stale)
    CS_FLD(Hot, inf) inf;
    CS_FLD(Hot, flags) hflags;
    CS_FLD(Hot, age) hage;
    CS_FLD(Hot, pflags) hpfl;
    CS_FLD(Hot, itype) itype;
    CS_FLD(Hot, infness) infness;
    CS_FLD(Hot, t_lat) t_lat;
    CS_FLD(Hot, t_as) t_as;
    CS_FLD(Hot, t_ae) t_ae;
    uint16_t* gate;       // [R][N] by list slot: no timer of the infectious person in the slot fires before this step
    Tim* tim;             // [R][N]
    CS_FLD(Tim, t_inf) t_inf;
    CS_FLD(Tim, t_rod) t_rod;
    CS_FLD(Tim, t_sari) t_sari;
    CS_FLD(Tim, t_crit) t_crit;
    CS_FLD(Tim, t_step) t_step;
    CS_FLD(Tim, t_det) t_det;
    CS_FLD(Tim, sev_cur) sev_cur;
    CS_FLD(Tim, sev_final) sev_final;
    CS_FLD(Tim, bits) bits;
    uint8_t *qc, *dct;
    int32_t *listpos, *infector;
    float *susc, *prob_care;
    uint16_t *t_ts, *t_te, *t_vs, *t_is, *t_qs;
    int32_t *cS, *cL, *cI, *cR, *cD, *cumTC, *S0;
    float* tot_prob;
    float* cum_trans;     // full tables (ck == 0): [R][NC][NC] normalised prefixes
    int32_t* invcdf;      // full tables: [R][NC][1025]
    // checkpointed tables (ck > 0), for regions whose NC x NC tables do not
    // fit: max_trans is evaluated where it is read, and each row keeps its
    // sequential float prefix only at the last cell of every ck cells
    float* ck_pre;        // [R][NC][ck_nb], the last entry being the row total
    int32_t* ck_inv;      // [R][NC][1025] or null: per row and k, the first block whose normalised last prefix reaches k / 1024
    int ck, ck_nb;
    int32_t* perm;
    uint16_t *pl_cs, *pl_ce;  // [R][total places]
    uint8_t *m_socdist, *m_moverest;
    uint32_t* m_trig;
    // interventions (launched only when P->interventions_active)
    IntState* ist;        // [R]
    int32_t* trig_hist;   // [R][CS_TRIG_RING][3]
    uint8_t* mc_state;    // [R][NM] Microcell.placeclose
    uint16_t* mc_end;     // [R][NM] Microcell.place_end_time
    uint8_t* pl_once;     // [R][total places] closed under once-only (Place.control_trig at USHRT_MAX - 1)
    const int32_t* mpl_start;  // [NM + 1] places by microcell (Microcell.places) ...
    const int32_t* mpl;        // ... each type << 27 | place
    int32_t* det;         // this step's detected cases (gp)
    int32_t* close_l;     // places closing in this step's TreatSweep: r * total places + place offset
    unsigned* itv_n;      // counts: [0] det, [1] close_l, [2] carers queued by closures (in care_k)
    unsigned det_cap, close_cap;
    Counters* cnt;
    PlContact* pcl;       // [pc_cap] place contacts of this step, CS_EV_SHARDS shards
    unsigned* pc_count;   // [CS_EV_SHARDS * CS_EV_STRIDE]
    unsigned pc_cap;
    Counters* cstage;     // [CS_TALLY_COPIES][R] staged counters (zero between phases)
    double* prev_prop;
    double* samp_prop;  // State.S / PopSize as of the last sample, per realisation
    double* loc_init;  // [R][2]
    // step scratch
    double* si_buf;       // [R][N] spatial infectiousness by permutation slot (sign = closed flag)
    float* ci_buf;        // [R][N] normalised cumulative infectiousness by slot
    int32_t* spK;         // [R*NC] draws per source cell
    // work lists, rebuilt every phase: chunks of up to 64 consecutive list
    // entries of one (realisation, cell), and lists of cells
    Chunk* segl;          // persons of a latent or infectious segment
    Chunk* drawl;         // spatial draws of a source cell
    int32_t* cells;       // cells with infectious persons
    unsigned* wl;         // counts: [0] segl, [1] cells, [2] drawl
    unsigned seg_cap, draw_cap;
    uint64_t* ev_key;
    long long* ev_gp;
    int32_t* ev_infector;
    int16_t* ev_type;
    unsigned int* ev_count;  // [CS_EV_SHARDS * CS_EV_STRIDE], shard s's count at s * CS_EV_STRIDE
    unsigned int ev_cap;     // a multiple of CS_EV_SHARDS; shard s holds [s, s + 1) * ev_cap / CS_EV_SHARDS
    ClaimRec* clr;        // [R][N]
    CS_FLD(ClaimRec, claim) claim;
    CS_FLD(ClaimRec, lastk) lastk;
    // point a record array's field accessors at it, all in one place per
    // record (a field left unbound would index from address zero)
    void bind_hot(Hot* p)
    {
        hot = p;
        inf.base = p; hflags.base = p; hage.base = p; hpfl.base = p; itype.base = p; infness.base = p;
        t_lat.base = p; t_as.base = p; t_ae.base = p;
    }
    void bind_tim(Tim* p)
    {
        tim = p;
        t_inf.base = p; t_rod.base = p; t_sari.base = p; t_crit.base = p; t_step.base = p; t_det.base = p;
        sev_cur.base = p; sev_final.base = p; bits.base = p;
    }
    void bind_clr(ClaimRec* p) { clr = p; claim.base = p; lastk.base = p; }
    int32_t* batch;       // [R][N] by permutation slot range of the cell
    int32_t* bcount;      // [R*NC]
    int32_t* mv_over;     // [R*NC] cells whose moves exceed move_cap, queued by k_list_move
    unsigned* mv_over_n;  // [2] queue length, blocks of k_list_move_big done
    uint32_t* mv_bm;      // [CS_MOVE_BIG_BLOCKS][2 * mv_words] per block: position bitmap, word prefixes
    int mv_words;         // words of one bitmap: a cell's largest segment / 32, rounded up
    int move_cap;         // largest move count the LDS path takes (<= CS_MOVE_CAP)
    int32_t *care_gp, *care_k;
    uint8_t* care_abs;
    unsigned int* care_count;
    unsigned int* status;
    Work* wc;
    StepScalars* sc;      // this step's scalars (k_step_begin)
    Clock* clock;
    int32_t* ts_buf;      // [R][max_samples][CS_TS_FIELDS]
    int ts_max;
    long long* last_tot;  // [R][CS_TS_LAST] cumI, cumC, D, severity incidence at the previous sample
    // the samples by age group and by administrative unit: a row of
    // ext_w = CS_TSA_FIELDS * CS_NUM_AGE_GROUPS + CS_TSU_FIELDS * NA words
    // per (realisation, sample), differencing the cumulative counters kept
    // in last_ext ([R][CS_TSA_CUM * CS_NUM_AGE_GROUPS + CS_TSU_FIELDS * NA])
    int32_t* ts_ext;
    long long* last_ext;
    int ext_w, NA;
    unsigned long long* ad_cnt;  // [CS_AD_COPIES][R][NA][CS_TSU_FIELDS] cumI, cumC, cumDC, cumD by unit, or null
    long long* due;       // [CS_DUE_RING][CS_DUE_SHARDS][due_scap] persons (gp) by latent time
    unsigned* due_n;      // [CS_DUE_RING][CS_DUE_SHARDS] counts, CS_DUE_STRIDE words apart
    unsigned* due_ovf;    // [CS_DUE_RING] a shard of the bucket overflowed: that step scans every latent person
    unsigned due_scap;    // persons per shard
    int due_on;           // 0: every step scans every latent person
};

// ---- predicates (src/ModelMacros.h:6-23, src/Person.cpp) ----
__device__ __forceinline__ int age_group(const Dev& d, int p) { return (int)d.age[p] / CS_AGE_GROUP_WIDTH; }
__device__ __forceinline__ bool care(const Dev& d, int p) { return (d.pflags[p] & 1) != 0; }
__device__ __forceinline__ int keyw(const Dev& d, int p) { return (d.pflags[p] >> 1) & 1; }
__device__ __forceinline__ bool esoc(const Dev& d, int p) { return (d.pflags[p] & 4) != 0; }
// the same from the copies in Hot (gp = r * N + p): a caller that holds the
// person's Hot record then needs no PStat line
__device__ __forceinline__ int age_group_g(const Dev& d, size_t gp) { return (int)d.hage[gp] / CS_AGE_GROUP_WIDTH; }
__device__ __forceinline__ bool care_g(const Dev& d, size_t gp) { return (d.hpfl[gp] & 1) != 0; }
__device__ __forceinline__ bool esoc_g(const Dev& d, size_t gp) { return (d.hpfl[gp] & 4) != 0; }
__device__ __forceinline__ int keyw_g(const Dev& d, size_t gp) { return (d.hpfl[gp] >> 1) & 1; }
__device__ __forceinline__ bool is_dead_code(int8_t s) { return s == CS_DEAD_ASYMP || s == CS_DEAD_SYMP; }
__device__ __forceinline__ bool host_absent(const Dev& d, size_t gp, unsigned ts) { return d.t_as[gp] <= ts && d.t_ae[gp] > ts; }
// d.np_off[t] for a run-time t, from the device copy: a run-time index into
// the by-value Dev makes the compiler copy the whole struct to every lane's
// scratch first
__device__ __forceinline__ int np_off_of(const Dev& d, int t) { return d.np_off_d[t]; }
__device__ __forceinline__ bool place_closed(const Dev& d, int r, int t, int k, unsigned ts)
{
    size_t i = (size_t)r * d.np_off[CS_MAX_PLACE_TYPES] + np_off_of(d, t) + k;
    return d.pl_cs[i] <= ts && d.pl_ce[i] > ts;
}
__device__ __forceinline__ bool treated(const Dev& d, size_t gp, unsigned ts)
{
    return d.P->interventions_active && d.t_te[gp] > ts && d.t_ts[gp] <= ts;
}
__device__ __forceinline__ bool vacced(const Dev& d, size_t gp, unsigned ts)
{
    return d.P->interventions_active && (int)d.t_vs[gp] + (int)d.P->us_vacc_time_to_efficacy <= (int)ts;
}
__device__ __forceinline__ bool vacced_switch(const Dev& d, size_t gp) { return d.t_vs[gp] >= d.P->us_vacc_time_efficacy_switch; }
__device__ __forceinline__ bool quarantined(const Dev& d, size_t gp, unsigned ts)
{
    return d.P->interventions_active && d.qc[gp] == 1 &&
           (int)d.t_qs[gp] + (int)d.P->us_hquarantine_house_duration > (int)ts && d.t_qs[gp] <= ts;
}
__device__ __forceinline__ bool isolated(const Dev& d, size_t gp, unsigned ts)
{
    const cs_params* P = d.P;
    return P->interventions_active && (int)d.t_is[gp] + (int)P->us_case_isolation_delay <= (int)ts &&
           (int)d.t_is[gp] + (int)P->us_case_isolation_delay + (int)P->us_case_isolation_duration > (int)ts;
}
__device__ __forceinline__ bool dct_traced(const Dev& d, size_t gp) { return d.P->interventions_active && d.dct[gp] == 1; }
// P.ControlPropCasesId of realisation r: its own when interventions run (the
// alert switches it), else the run's
__device__ __forceinline__ double control_prop(const Dev& d, int r)
{
    return d.P->interventions_active ? d.ist[r].ctrl : d.P->control_prop_cases_id;
}
__device__ __forceinline__ bool socdist(const Dev& d, int r, int p)
{
    return d.P->interventions_active && d.m_socdist[(size_t)r * d.NM + d.mcell[p]] == 2;
}

// the kernel families of KernelStruct (src/Kernels.cpp:68-106), on the host:
// the lookup tables are lookup_[i] = k(i delta), hi_res_[i] = k(i delta / E)
// (src/Kernels.cpp:17-45)
inline double kernel_fn(const cs_params& P, double r2)
{
    const double sh = P.kernel_shape, sc = P.kernel_scale, p3 = P.kernel_p3, p4 = P.kernel_p4;
    double t, dd;
    switch (P.kernel_type) {
    case 1: return exp(-sqrt(r2) / sc);
    case 2: t = -sh * log(sqrt(r2) / sc + 1.0); return (t < -690.0) ? 0.0 : exp(t);
    case 3: return exp(-r2 / (sc * sc));
    case 4: return (r2 > sc * sc) ? 0.0 : 1.0;
    case 5: t = 0.5 * sh * log(r2 / (sc * sc)); return (t > 690.0) ? 0.0 : (1.0 / (exp(t) + 1.0));
    case 6: t = log(sqrt(r2) / sc + 1.0); return (t < -690.0) ? 0.0 : (exp(-sh * t) + p3 * exp(-p4 * t)) / (1.0 + p3);
    default: dd = sqrt(r2); t = -sh * log(dd / sc + 1.0); return (t < -690.0) ? 0.0 : exp(t - pow(dd / p3, p4));
    }
}

// a person's counters by administrative unit (StateT's *_adunit counters):
// CS_AD_COPIES copies, a block adding to the copy its index picks, so that
// regions of few units do not send every block's adds to one address;
// readers sum the copies (ad_total)
__device__ __forceinline__ unsigned long long* ad_counter(const Dev& d, int r, int p, int field)
{
    const size_t copy = blockIdx.x & (CS_AD_COPIES - 1);
    return d.ad_cnt + ((copy * d.R + r) * d.NA + d.padunit[p]) * CS_TSU_FIELDS + field;
}

__device__ __forceinline__ long long ad_total(const Dev& d, int r, int u, int field)
{
    long long s = 0;
    for (int k = 0; k < CS_AD_COPIES; k++)
        s += (long long)d.ad_cnt[(((size_t)k * d.R + r) * d.NA + u) * CS_TSU_FIELDS + field];
    return s;
}

// ---- kernel lookup (src/Kernels.cpp:108-129) ----
__device__ __forceinline__ double kernel_num(const Dev& d, double r2)
{
    const cs_params* P = d.P;
    double t = r2 / P->kernel_delta;
    const int sz = P->kernel_size;
    if (t > (double)sz) {
        atomicOr(d.status, (unsigned)FLAG_RANGE);
        return 0.0;
    }
    double s = t * (double)P->kernel_expansion;
    if (s < (double)sz) {
        t = s - __builtin_floor(s);
        int i0 = (int)s, i1 = (int)(s + 1.0);
        if (i1 > sz) i1 = sz;
        t = (1.0 - t) * d.kh[i0] + t * d.kh[i1];
    } else {
        s = t - __builtin_floor(t);
        int i0 = (int)t, i1 = (int)(t + 1.0);
        if (i1 > sz) i1 = sz;
        t = (1.0 - s) * d.kl[i0] + s * d.kl[i1];
    }
    return t;
}

// ---- distances (src/Dist.cpp:14-168) ----
__device__ __forceinline__ double periodic_xy(const cs_params* P, double x, double y)
{
    if (P->do_periodic_boundaries) {
        if (x > P->degrees_width * 0.5) x = P->degrees_width - x;
        if (y > P->degrees_height * 0.5) y = P->degrees_height - y;
    }
    return x * x + y * y;
}

__device__ __forceinline__ double dist2_utm_inl(const cs_params* P, double x1, double y1, double x2, double y2)
{
    double x, y, cy1, cy2, yt, xi, yi;
    x = __builtin_fabs(x1 - x2) / 2;
    y = __builtin_fabs(y1 - y2) / 2;
    xi = __builtin_floor(x);
    yi = __builtin_floor(y);
    x -= xi;
    y -= yi;
    x = (1 - x) * P->sinx[(int)xi] + x * P->sinx[((int)xi) + 1];
    y = (1 - y) * P->sinx[(int)yi] + y * P->sinx[((int)yi) + 1];
    yt = __builtin_fabs(y1 + P->bbox_bottom_left_y);
    yi = __builtin_floor(yt);
    cy1 = yt - yi;
    cy1 = (1 - cy1) * P->cosx[((int)yi)] + cy1 * P->cosx[((int)yi) + 1];
    yt = __builtin_fabs(y2 + P->bbox_bottom_left_y);
    yi = __builtin_floor(yt);
    cy2 = yt - yi;
    cy2 = (1 - cy2) * P->cosx[((int)yi)] + cy2 * P->cosx[((int)yi) + 1];
    x = __builtin_fabs(1000 * (y * y + x * x * cy1 * cy2));
    xi = __builtin_floor(x);
    x -= xi;
    y = (1 - x) * P->asin2sqx[((int)xi)] + x * P->asin2sqx[((int)xi) + 1];
    const double er = ((double)(1852 * 60 * 360) / 3.1415926535) / 2;
    return 4 * er * er * y;
}

// the same out of line, for kernels whose register budget is tight
__device__ __noinline__ double dist2_utm(const cs_params* P, double x1, double y1, double x2, double y2)
{
    return dist2_utm_inl(P, x1, y1, x2, y2);
}

// dist2 of two household locations (src/Dist.cpp dist2 over Households[].loc)
__device__ __forceinline__ double dist2_ll(const cs_params* P, float2 la, float2 lb)
{
    if (P->do_utm_coords) return dist2_utm(P, la.x, la.y, lb.x, lb.y);
    double x = __builtin_fabs((double)la.x - (double)lb.x);
    double y = __builtin_fabs((double)la.y - (double)lb.y);
    return periodic_xy(P, x, y);
}

__device__ __forceinline__ double dist2_pp(const Dev& d, int a, int b) { return dist2_ll(d.P, d.ploc[a], d.ploc[b]); }

// dist2_cc_min (src/Dist.cpp:79-144) on cell grid coordinates: cell l is
// (lx, ly) = (l / nch, l % nch).  The reference moves the corner cells by
// whole indices (i += nch: the next column; i++: the next row, wrapping
// into the next column at the top), then divides again; here the same
// moves are made on the coordinates, so every double is formed from the
// same integers as the reference's, with no integer division.
__device__ __forceinline__ void cc_next_row(int& x, int& y, int nch)
{
    if (++y == nch) {
        y = 0;
        x++;
    }
}
// INL: the UTM distance inlined (loops that evaluate many independent pairs)
template <bool INL = false>
__device__ __forceinline__ double dist2_cc_min_xy(const cs_params* P, int lx, int ly, int mx, int my)
{
    const int nch = P->nch;
    int ix = lx, iy = ly, jx = mx, jy = my;
    if (P->do_utm_coords) {
        const double PI_ = 3.1415926535;
        if (P->cell_width * ((double)abs(mx - lx)) > PI_) {
            if (mx > lx) jx++;
            else if (mx < lx) ix++;
        } else {
            if (mx > lx) ix++;
            else if (mx < lx) jx++;
        }
        if (my > ly) cc_next_row(ix, iy, nch);
        else if (my < ly) cc_next_row(jx, jy, nch);
        return (INL ? dist2_utm_inl : dist2_utm)(P, P->cell_width * __builtin_fabs((double)ix), P->cell_height * __builtin_fabs((double)iy),
                         P->cell_width * __builtin_fabs((double)jx), P->cell_height * __builtin_fabs((double)jy));
    }
    if (P->do_periodic_boundaries && (P->cell_width * ((double)abs(mx - lx)) > P->degrees_width * 0.5)) {
        if (mx > lx) jx++;
        else if (mx < lx) ix++;
    } else {
        if (mx > lx) ix++;
        else if (mx < lx) jx++;
    }
    if (P->do_periodic_boundaries && (P->degrees_height * ((double)abs(my - ly)) > P->degrees_height * 0.5)) {
        if (my > ly) cc_next_row(jx, jy, nch);
        else if (my < ly) cc_next_row(ix, iy, nch);
    } else {
        if (my > ly) cc_next_row(ix, iy, nch);
        else if (my < ly) cc_next_row(jx, jy, nch);
    }
    double x = P->cell_width * __builtin_fabs((double)(ix - jx));
    double y = P->cell_height * __builtin_fabs((double)(iy - jy));
    return periodic_xy(P, x, y);
}

__device__ __noinline__ double dist2_cc_min(const cs_params* P, int l, int m)
{
    return dist2_cc_min_xy(P, l / P->nch, l % P->nch, m / P->nch, m % P->nch);
}

// ---- InverseCdf::choose for a given uniform (src/InverseCdf.cpp:60-74) ----
__device__ __forceinline__ uint16_t icdf_choose(const double* c, double mean, double u, double tspd)
{
    double q;
    int i = (int)__builtin_floor(q = u * CS_CDF_RES);
    q -= (double)i;
    double ti = -mean * clog(q * c[i + 1] + (1.0 - q) * c[i]);
    return (uint16_t)__builtin_floor(0.5 + (ti * tspd));
}

// ---- hazards (src/CalcInfSusc.cpp:12-80); r = realisation, p = person ----
// The infector-side hazards read the person's Hot fields through an
// accessor: HotById reads them from d.hot by person (one field at a time),
// HotRec from a record already in registers (the slot copy, Dev.shot, of an
// infectious person).  Both give the same doubles.
struct HotById {
    const Dev& d;
    size_t gp;
    __device__ __forceinline__ int8_t inf() const { return d.inf[gp]; }
    __device__ __forceinline__ uint8_t flags() const { return d.hflags[gp]; }
    __device__ __forceinline__ uint8_t pflags() const { return d.hpfl[gp]; }
    __device__ __forceinline__ uint8_t age() const { return d.hage[gp]; }
    __device__ __forceinline__ float infness() const { return d.infness[gp]; }
    __device__ __forceinline__ uint16_t t_lat() const { return d.t_lat[gp]; }
};
struct HotRec {
    const Hot& h;
    __device__ __forceinline__ int8_t inf() const { return h.inf; }
    __device__ __forceinline__ uint8_t flags() const { return h.flags; }
    __device__ __forceinline__ uint8_t pflags() const { return h.pflags; }
    __device__ __forceinline__ uint8_t age() const { return h.age; }
    __device__ __forceinline__ float infness() const { return h.infness; }
    __device__ __forceinline__ uint16_t t_lat() const { return h.t_lat; }
};

template <class H>
__device__ __forceinline__ double inf_profile_t(const Dev& d, const H& h, unsigned ts)
{
    int idx = (int)ts - (int)h.t_lat() - 1;
    idx = idx < 0 ? 0 : (idx >= CS_MAX_INFECTIOUS_STEPS ? CS_MAX_INFECTIOUS_STEPS - 1 : idx);
    return d.P->infectiousness[idx];
}
__device__ __forceinline__ double inf_profile(const Dev& d, size_t gp, unsigned ts) { return inf_profile_t(d, HotById{d, gp}, ts); }

template <class H>
__device__ __forceinline__ double person_inf_t(const Dev& d, size_t gp, const H& h, unsigned ts)
{
    const cs_params* P = d.P;
    return (treated(d, gp, ts) ? P->treat_inf_drop : 1.0)
         * (vacced(d, gp, ts) ? P->vacc_inf_drop : 1.0)
         * __builtin_fabs((double)h.infness())
         * inf_profile_t(d, h, ts);
}
__device__ __forceinline__ double person_inf(const Dev& d, size_t gp, unsigned ts) { return person_inf_t(d, gp, HotById{d, gp}, ts); }

// the resident count of p's household: its size (in p's Hot record) unless
// the flags say otherwise
template <class H>
__device__ __forceinline__ int hh_nhr_t(const Dev& d, const H& h, int p)
{
    return (h.pflags() & PF_NHR_DIFFERS) ? (int)d.hnhr[d.hh[p]] : (int)(h.flags() >> HOT_NH_SHIFT);
}

template <class H>
__device__ __forceinline__ double house_inf_t(const Dev& d, int r, int p, const H& h, unsigned ts)
{
    const cs_params* P = d.P;
    size_t gp = (size_t)r * d.N + p;
    bool iso = isolated(d, gp, ts), dc = dct_traced(d, gp), qu = quarantined(d, gp, ts);
    return ((iso && !dc) ? P->efficacies[0][CS_HOUSE] : 1.0)
         * (dc ? P->efficacies[5][CS_HOUSE] : 1.0)
         * ((qu && !dc && !iso) ? P->efficacies[1][CS_HOUSE] : 1.0)
         * P->household_denom_lookup[hh_nhr_t(d, h, p) - 1]
         * ((h.pflags() & 1) ? P->care_home_resident_household_scaling : 1.0)
         * (treated(d, gp, ts) ? P->treat_inf_drop : 1.0)
         * (vacced(d, gp, ts) ? P->vacc_inf_drop : 1.0)
         * (P->no_infectiousness_sd_in_hh ? ((h.infness() < 0) ? P->sympt_infectiousness : P->asympt_infectiousness)
                                          : __builtin_fabs((double)h.infness()))
         * inf_profile_t(d, h, ts);
}
__device__ __forceinline__ double house_inf(const Dev& d, int r, int p, unsigned ts)
{
    return house_inf_t(d, r, p, HotById{d, (size_t)r * d.N + p}, ts);
}

template <class H>
__device__ __forceinline__ double place_inf_t(const Dev& d, int r, int p, const H& h, int pt, unsigned ts)
{
    const cs_params* P = d.P;
    size_t gp = (size_t)r * d.N + p;
    bool iso = isolated(d, gp, ts), dc = dct_traced(d, gp), qu = quarantined(d, gp, ts);
    const bool care = (h.pflags() & 1) != 0;
    return ((iso && !dc) ? P->efficacies[0][pt] : 1.0)
         * (dc ? P->efficacies[5][pt] : 1.0)
         * ((qu && !care && !dc && !iso) ? P->efficacies[1][pt] : 1.0)
         * ((h.inf() == CS_CASE && !care) ? P->sympt_place_type_contact_rate[pt] : 1.0)
         * P->place_type_trans[pt] / P->place_type_group_size_param1[pt] * person_inf_t(d, gp, h, ts);
}
__device__ __forceinline__ double place_inf(const Dev& d, int r, int p, int pt, unsigned ts)
{
    return place_inf_t(d, r, p, HotById{d, (size_t)r * d.N + p}, pt, ts);
}

template <class H>
__device__ __forceinline__ double spatial_inf_t(const Dev& d, int r, int p, const H& h, unsigned ts)
{
    const cs_params* P = d.P;
    size_t gp = (size_t)r * d.N + p;
    bool iso = isolated(d, gp, ts), dc = dct_traced(d, gp), qu = quarantined(d, gp, ts);
    return ((iso && !dc) ? P->efficacies[0][CS_SPATIAL] : 1.0)
         * (dc ? P->efficacies[5][CS_SPATIAL] : 1.0)
         * ((qu && !(h.pflags() & 1) && !dc && !iso) ? P->efficacies[1][CS_SPATIAL] : 1.0)
         * ((h.inf() == CS_CASE) ? P->sympt_spatial_contact_rate : 1.0)
         * P->relative_spatial_contact[(int)h.age() / CS_AGE_GROUP_WIDTH]
         * person_inf_t(d, gp, h, ts);
}
__device__ __forceinline__ double spatial_inf(const Dev& d, int r, int p, unsigned ts)
{
    return spatial_inf_t(d, r, p, HotById{d, (size_t)r * d.N + p}, ts);
}

// an infectious person's slot copy of its Hot record (Dev.shot) after a
// change of its Hot fields by a writer that reaches it by person: the copy
// takes the record, keeping its own HOT_HH_DONE bit (set on the copy only)
__device__ __forceinline__ bool in_infectious_list(int8_t s)
{
    return s == CS_INF_ALMOST_SYMPT || s == CS_INF_ASYMPT || s == CS_CASE;
}
// (in words, stored as two 8-byte halves: a Hot local, or a 16-byte vector
// store, here made the compiler give these kernels LDS they do not use,
// costing them occupancy)
static_assert(offsetof(Hot, flags) == 1, "Hot.flags is byte 1 of the record's first word");
__device__ __forceinline__ void shot_sync_at(const Dev& d, size_t gp, size_t slot)
{
    uint4 v = reinterpret_cast<const uint4*>(d.hot)[gp];
    v.x |= reinterpret_cast<const uint32_t*>(d.shot + slot)[0] & ((uint32_t)HOT_HH_DONE << 8);
    uint2* o = reinterpret_cast<uint2*>(d.shot + slot);
    o[0] = make_uint2(v.x, v.y);
    o[1] = make_uint2(v.z, v.w);
}
__device__ __forceinline__ void shot_sync(const Dev& d, size_t gp, int p)
{
    if (!in_infectious_list(d.inf[gp])) return;
    shot_sync_at(d, gp, (gp - (size_t)p) + d.cstart[d.pc[p]] + d.listpos[gp]);
}

// the same with the two age groups given (a caller that already holds them)
__device__ __forceinline__ double person_susc_ag(const Dev& d, size_t gp, int ag, int ag_infector, unsigned ts)
{
    const cs_params* P = d.P;
    return P->waifw[ag][ag_infector]
         * P->age_susceptibility[ag] * (double)d.susc[gp]
         * (treated(d, gp, ts) ? P->treat_susc_drop : 1.0)
         * (vacced(d, gp, ts) ? (vacced_switch(d, gp) ? P->vacc_susc_drop2 : P->vacc_susc_drop) : 1.0);
}

__device__ __forceinline__ double person_susc(const Dev& d, int r, int p, unsigned ts, int infector)
{
    const size_t gp = (size_t)r * d.N + p;
    return person_susc_ag(d, gp, age_group_g(d, gp), age_group_g(d, (size_t)r * d.N + infector), ts);
}

// the infector's age group given (a caller that already holds its record)
__device__ __forceinline__ double house_susc_ag(const Dev& d, int r, int p, unsigned ts, int ag_infector)
{
    const cs_params* P = d.P;
    size_t gp = (size_t)r * d.N + p;
    return person_susc_ag(d, gp, age_group_g(d, gp), ag_infector, ts)
         * (socdist(d, r, p) ? (esoc_g(d, gp) ? P->efficacies[4][CS_HOUSE] : P->efficacies[3][CS_HOUSE]) : 1.0)
         * (dct_traced(d, gp) ? P->efficacies[5][CS_HOUSE] : 1.0)
         * (care_g(d, gp) ? P->care_home_resident_household_scaling : 1.0);
}

__device__ __forceinline__ double house_susc(const Dev& d, int r, int p, unsigned ts, int infector)
{
    return house_susc_ag(d, r, p, ts, age_group_g(d, (size_t)r * d.N + infector));
}

__device__ __forceinline__ double place_susc(const Dev& d, int r, int p, int pt, unsigned ts)
{
    const cs_params* P = d.P;
    size_t gp = (size_t)r * d.N + p;
    bool dc = dct_traced(d, gp);
    return ((quarantined(d, gp, ts) && !care_g(d, gp) && !dc) ? P->efficacies[1][pt] : 1.0)
         * (socdist(d, r, p) ? (esoc_g(d, gp) ? P->efficacies[4][pt] : P->efficacies[3][pt]) : 1.0)
         * (dc ? P->efficacies[5][pt] : 1.0);
}

__device__ __forceinline__ double spatial_susc(const Dev& d, int r, int p, unsigned ts)
{
    const cs_params* P = d.P;
    size_t gp = (size_t)r * d.N + p;
    bool dc = dct_traced(d, gp);
    return ((quarantined(d, gp, ts) && !care_g(d, gp) && !dc) ? P->efficacies[1][CS_SPATIAL] : 1.0)
         * (socdist(d, r, p) ? (esoc_g(d, gp) ? P->efficacies[4][CS_SPATIAL] : P->efficacies[3][CS_SPATIAL]) : 1.0)
         * (dc ? P->efficacies[5][CS_SPATIAL] : 1.0)
         * P->relative_spatial_contact_susc[age_group_g(d, gp)];
}

}  // namespace cs
