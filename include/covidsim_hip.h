/*
 * covidsim_hip.h -- C ABI of the MI355X (gfx950) infection-step library.
 *
 * The library replaces CovidSim's per-timestep infection step:
 *   InfectSweep(double t, int run)   src/Sweep.cpp:283-881 (declared src/Sweep.h:4-9)
 *   IncubRecoverySweep(double t)     src/Sweep.cpp:883-979
 * together with the hazards of src/CalcInfSusc.cpp, the spatial-kernel lookup
 * of src/Kernels.cpp, the cell sampling tables of UpdateProbs
 * (src/CovidSim.cpp:2752-2802) and the state transitions those sweeps call in
 * src/Update.cpp.  CovidSim itself has no plugin API: its host code calls the
 * two sweeps as free functions over global arrays (src/CovidSim.cpp:1488,
 * :1490).  A shim translation unit defines those two symbols and forwards to
 * the functions below (the binding is shown in INTEGRATION.md).
 *
 * Conventions
 *  - Every function returns 0 on success and a negative CS_E* code on error;
 *    cs_last_error() gives the message.  Errors are never silent (the
 *    reference drops overflowing infection-queue entries, src/Sweep.cpp:2043;
 *    here an overflow is reported and the step is re-run with a larger queue).
 *  - Plain C types only; every pointer argument is host memory owned by the
 *    caller and is not retained after the call returns.
 *  - One caller thread per engine; calls are ordered on the engine's stream.
 *  - All realisations held by an engine advance together in one launch.
 */
#ifndef COVIDSIM_HIP_H_
#define COVIDSIM_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- sizes that mirror the reference's compile-time constants ---------- */
#define CS_NUM_AGE_GROUPS      17     /* src/Constants.h NUM_AGE_GROUPS */
#define CS_AGE_GROUP_WIDTH     5      /* src/Constants.h AGE_GROUP_WIDTH */
#define CS_MAX_PLACE_TYPES     5      /* src/Country.h MAX_NUM_PLACE_TYPES */
#define CS_HOUSE               5      /* src/Constants.h House   (setting index) */
#define CS_SPATIAL             6      /* src/Constants.h Spatial (setting index) */
#define CS_NUM_SETTINGS        7      /* MAX_NUM_PLACE_TYPES + 2 */
#define CS_NUM_NPI             6      /* CI, HQ, PC, SD, ESD, DCT (src/Constants.h:121-126) */
#define CS_MAX_HOUSEHOLD_SIZE  10     /* src/Country.h */
#define CS_CDF_RES             20     /* src/Constants.h CDF_RES */
#define CS_MAX_INFECTIOUS_STEPS 2550  /* src/Constants.h */
#define CS_DAYS_PER_YEAR       364
#define CS_INFECT_TYPE_MASK    16
#define CS_MAX_HOLIDAYS        64
#define CS_UNSET_TIME          65534  /* USHRT_MAX - 1, the reference's "never" */

/* InfStat codes, src/InfStat.h:7-38 */
#define CS_SUSCEPTIBLE          0
#define CS_LATENT               1
#define CS_INF_ALMOST_SYMPT    (-1)
#define CS_INF_ASYMPT           2
#define CS_CASE                (-2)
#define CS_RECOVERED_ASYMP      3
#define CS_RECOVERED_SYMP      (-3)
#define CS_IMMUNE_AT_START      4
#define CS_DEAD_ASYMP           5
#define CS_DEAD_SYMP           (-5)

/* Severity codes, src/InfStat.h:41-58 */
#define CS_SEV_ASYMPTOMATIC 0
#define CS_SEV_MILD         1
#define CS_SEV_ILI          2
#define CS_SEV_SARI         3
#define CS_SEV_CRITICAL     4
#define CS_SEV_STEPDOWN     5
#define CS_SEV_DEAD         6
#define CS_SEV_RECOVERED    7

/* error codes */
#define CS_OK              0
#define CS_EINVAL        (-1)   /* bad argument / inconsistent sizes          */
#define CS_EHIP          (-2)   /* HIP runtime error                           */
#define CS_ENOMEM        (-3)   /* device allocation failed                    */
#define CS_ESTATE        (-4)   /* call out of order (e.g. no population yet)  */
#define CS_ERANGE        (-5)   /* kernel r^2 beyond the lookup table, src/Kernels.cpp:110-115 */
#define CS_EUNSUPPORTED  (-6)   /* parameter combination outside this path     */
#define CS_EOVERFLOW     (-7)   /* event queue full (retried internally first) */

/* ICDF table slots in cs_params.icdf[] (src/Param.h:168-170) */
enum {
    CS_ICDF_LATENT = 0, CS_ICDF_INFECTIOUS, CS_ICDF_MILD_RECOV, CS_ICDF_ILI_RECOV,
    CS_ICDF_SARI_RECOV, CS_ICDF_CRIT_CRITRECOV, CS_ICDF_CRITRECOV_RECOV,
    CS_ICDF_ILI_SARI, CS_ICDF_SARI_CRIT, CS_ICDF_ILI_DEATH, CS_ICDF_SARI_DEATH,
    CS_ICDF_STEPDOWN_DEATH, CS_ICDF_CRIT_DEATH, CS_NUM_ICDF
};
/* per-age mean slots in cs_params.mean_by_age[] (src/Param.h:173-175) */
enum {
    CS_MEAN_MILD_RECOV = 0, CS_MEAN_ILI_RECOV, CS_MEAN_SARI_RECOV, CS_MEAN_CRIT_CRITRECOV,
    CS_MEAN_CRITRECOV_RECOV, CS_MEAN_STEPDOWN_DEATH, CS_MEAN_ILI_SARI, CS_MEAN_SARI_CRIT,
    CS_MEAN_CRIT_DEATH, CS_MEAN_SARI_DEATH, CS_MEAN_ILI_DEATH, CS_NUM_MEANS
};

/*
 * Derived model constants the path reads.  Field names follow Param
 * (src/Param.h); values are those after Params::ReadParams and the
 * run-start updates of InitModel (src/CovidSim.cpp:1140-1187).
 */
typedef struct cs_params {
    /* time (src/ReadParams.cpp:1754-1760) */
    double time_steps_per_day;         /* P.TimeStepsPerDay */
    double model_time_step;            /* P.ModelTimeStep   */
    double false_positive_rate;        /* P.FalsePositiveRate (only 0 is supported) */
    int    do_seasonality;
    int    pad0;
    double seasonality[CS_DAYS_PER_YEAR];
    /* transmission */
    double local_beta;                 /* P.LocalBeta (InitTransmissionCoeffs) */
    double household_trans;            /* P.HouseholdTrans */
    double household_denom_lookup[CS_MAX_HOUSEHOLD_SIZE];
    int    do_households, do_places, num_place_types, hotel_place_type;
    int    place_type_no_air_num, care_home_place_type;
    int    do_symptoms, do_severity, do_latent, do_infectiousness_profile;
    int    do_real_sympt_withdrawal, absenteeism_place_closure;
    int    no_infectiousness_sd_in_hh, do_death, do_ad_units, output_ad_unit_age;
    int    include_stepdown_to_death, do_digital_contact_tracing;
    int    do_blanket_move_restr;
    int    interventions_active;       /* 1: an intervention is configured (the subset below is served, any other refused) */
    double place_type_trans[CS_MAX_PLACE_TYPES];
    double place_type_group_size_param1[CS_MAX_PLACE_TYPES];
    double place_type_prop_between_group_links[CS_MAX_PLACE_TYPES];
    double sympt_place_type_contact_rate[CS_MAX_PLACE_TYPES];
    double sympt_place_type_withdrawal_prop[CS_MAX_PLACE_TYPES];
    double sympt_spatial_contact_rate;
    double relative_spatial_contact[CS_NUM_AGE_GROUPS];
    double relative_spatial_contact_susc[CS_NUM_AGE_GROUPS];
    double waifw[CS_NUM_AGE_GROUPS][CS_NUM_AGE_GROUPS];          /* [susceptible][infector] */
    double waifw_spatial[CS_NUM_AGE_GROUPS][CS_NUM_AGE_GROUPS];
    double age_susceptibility[CS_NUM_AGE_GROUPS];
    double age_infectiousness[CS_NUM_AGE_GROUPS];
    double efficacies[CS_NUM_NPI][CS_NUM_SETTINGS];              /* P.Efficacies */
    double care_home_resident_household_scaling, care_home_resident_spatial_scaling;
    double care_home_worker_group_scaling, care_home_resident_place_scaling, care_home_rel_prob_hosp;
    double treat_inf_drop, vacc_inf_drop, treat_susc_drop, vacc_susc_drop, vacc_susc_drop2;
    double treat_sympt_drop, vacc_sympt_drop, treat_death_drop;
    uint16_t us_vacc_time_to_efficacy, us_vacc_time_efficacy_switch;
    uint16_t us_hquarantine_house_duration, us_case_isolation_delay;
    uint16_t us_case_isolation_duration, us_case_absenteeism_delay;
    uint16_t us_case_absenteeism_duration, pad1;
    /* natural history */
    int    n_infectiousness;           /* valid entries of infectiousness[] */
    int    pad2;
    double infectiousness[CS_MAX_INFECTIOUS_STEPS]; /* P.infectiousness, per step */
    double latent_period, infectious_period, latent_to_sympt_delay;
    double sympt_infectiousness, asympt_infectiousness, infectiousness_sd, susceptibility_sd;
    double proportion_symptomatic[CS_NUM_AGE_GROUPS];
    double icdf[CS_NUM_ICDF][CS_CDF_RES + 1];                    /* exp(-x) transformed tables */
    double mean_by_age[CS_NUM_MEANS][CS_NUM_AGE_GROUPS];
    double prop_ili[CS_NUM_AGE_GROUPS], prop_sari[CS_NUM_AGE_GROUPS], prop_critical[CS_NUM_AGE_GROUPS];
    double cfr_critical[CS_NUM_AGE_GROUPS], cfr_sari[CS_NUM_AGE_GROUPS], cfr_ili[CS_NUM_AGE_GROUPS];
    double cfr_critical_scale, cfr_sari_scale, cfr_ili_scale;
    double control_prop_cases_id;      /* P.ControlPropCasesId during the run */
    int    case_absent_child_age_cutoff;
    int    place_close_round_household;
    double case_absent_child_prop_adult_carers;
    double location_initial_infection[2]; /* P.LocationInitialInfection[0] after seeding */
    /* spatial kernel (src/Kernels.h:75-115); tables are passed separately */
    int    kernel_size;                /* KernelLookup.size_ */
    int    kernel_expansion;           /* KernelLookup.expansion_factor_ */
    double kernel_delta;               /* KernelLookup.delta_ */
    int    kernel_type;                /* KernelStruct.type_ (1..7) */
    int    pad3;
    double kernel_shape, kernel_scale, kernel_p3, kernel_p4;
    /* geometry (src/Dist.cpp) */
    int    do_utm_coords, do_periodic_boundaries;
    int    nch;                        /* P.nch: cells per column */
    int    pad4;
    double cell_width, cell_height;    /* P.in_cells_ */
    double degrees_width, degrees_height; /* P.in_degrees_ */
    double bbox_bottom_left_y;         /* P.SpatialBoundingBox.bottom_left().y */
    double sinx[361], cosx[361], asin2sqx[1001];
    /* holidays (src/Sweep.cpp:889-918) */
    int    do_holidays, num_holidays;
    double holiday_effect[CS_MAX_PLACE_TYPES];
    double holiday_start[CS_MAX_HOLIDAYS], holiday_duration[CS_MAX_HOLIDAYS];
    double holidays_start_day_simtime;
    /* InitTransmissionCoeffs inputs (src/SetupModel.cpp:640-846) */
    double r0;                         /* P.R0, the target reproduction number */
    int    fix_local_beta;             /* P.FixLocalBeta: keep local_beta as given */
    int    do_si;                      /* P.DoSI */
    int    do_partial_immunity;        /* P.DoPartialImmunity */
    int    got_waifw_spatial;          /* P.Got_WAIFW_Matrix_Spatial */
    double initial_immunity[CS_NUM_AGE_GROUPS];
    double care_home_prop_residents;   /* P.CareHomePropResidents */
    /* ---- the intervention subset, served when interventions_active != 0 ----
     * Case isolation and household quarantine of detected cases
     * (DoDetectedCase, src/Update.cpp:525-792), place closure by microcell on
     * global triggers (TreatSweep, src/Sweep.cpp:1651-1740; DoPlaceClose and
     * DoPlaceOpen, src/Update.cpp:1095-1295), and the alert, trigger and
     * start-time logic of RecordSample and CalibrationThresholdCheck
     * (src/CovidSim.cpp:3150-3160, :3485-3689).  Field names follow Param;
     * times are in days, trigger windows in output samples.  cs_run_steps is
     * the only entry point that runs them (the single-sweep entry points
     * refuse an active intervention set: TreatSweep and the triggers run
     * between sweeps, in RunModel). */
    int    do_global_triggers, do_icu_triggers, do_admin_triggers, do_per_capita_triggers;
    int    triggers_sampling_interval;          /* P.TriggersSamplingInterval */
    int    trigger_alert_on_deaths;             /* P.TriggerAlertOnDeaths */
    int    case_or_death_threshold_before_alert;
    int    window_to_evaluate_trigger_alert;    /* P.WindowToEvaluateTriggerAlert */
    int    do_no_calibration, do_alert_trigger_after_interv;
    double date_trigger_reached_cal_time;       /* P.DateTriggerReached_CalTime (< 0: none) */
    double pre_alert_control_prop_cases_id, post_alert_control_prop_cases_id;
    double case_isolation_time_start_base, case_isolation_prop, case_isolation_policy_duration;
    double case_isolation_cell_inc_thresh;      /* P.CaseIsolation_CellIncThresh */
    double hquarantine_time_start_base, hquarantine_delay, hquarantine_policy_duration;
    double hquarantine_prop_house_compliant, hquarantine_prop_indiv_compliant;
    double hhquar_cell_inc_thresh;              /* P.HHQuar_CellIncThresh */
    int    do_hq_retrigger, do_place_close_once_only;
    double place_close_time_start_base, place_close_delay_mean, place_close_duration;  /* duration: P.PlaceCloseDurationBase */
    int    place_close_cell_inc_thresh;         /* P.PlaceCloseCellIncThresh1 */
    int    place_close_cell_inc_stop_thresh, place_close_inc_trig, place_close_by_admin_unit;
    double place_close_frac_inc_trig;
    double place_close_prop_attending[CS_MAX_PLACE_TYPES];
    /* interventions outside the subset: each must be off (start times of
     * 10000 days or more, switches 0); cs_run_steps refuses anything else */
    double treat_time_start_base, vacc_time_start_base, soc_dist_time_start_base, move_restr_time_start_base;
    double key_worker_proph_time_start_base, airport_close_time_start_base, place_close_time_start_base2;
    double place_close_indep_thresh, vacc_time_start_geo;
    int    do_intervention_delays_by_adunit, vary_efficacies_over_time, do_mass_vacc, pad5;
} cs_params;

/*
 * The population a realisation runs on, as CovidSim's SetupModel leaves it
 * (src/Model.h:295-303).  Person-indexed arrays have n_people entries;
 * place_links/place_group_links are [n_people][CS_MAX_PLACE_TYPES].
 * Places of type t: member list place_members[t][place_member_start[t][k] ..
 * +place_n[t][k]), groups place_group_start[t][place_groups_start[t][k] + g]
 * (offsets into the member list of that place) and place_group_size likewise.
 */
typedef struct cs_population {
    int32_t n_people, n_households;
    int32_t n_cells;                  /* NumPopulatedCells */
    int32_t n_cells_total;            /* NumCells (ncw*nch)  */
    int32_t n_mcells;                 /* total microcells    */
    int32_t n_adunits;
    /* persons */
    const int32_t*  hh;               /* Person.hh     */
    const int32_t*  mcell;            /* Person.mcell  */
    const int32_t*  pcell;            /* Person.pcell  (index into Cells[]) */
    const uint8_t*  age;              /* Person.age (years) */
    const uint8_t*  flags;            /* bit0 care_home_resident, bit1 keyworker, bit2 esocdist_comply, bit3 digitalContactTracingUser */
    const int32_t*  place_links;      /* Person.PlaceLinks        */
    const uint16_t* place_group_links;/* Person.PlaceGroupLinks   */
    /* households */
    const int32_t*  hh_first;         /* Household.FirstPerson */
    const uint16_t* hh_nh;            /* Household.nh  */
    const uint16_t* hh_nhr;           /* Household.nhr */
    const float*    hh_loc;           /* Household.loc, [n_households][2] */
    /* populated cells, in CellLookup order */
    const int32_t*  cell_index;       /* CellLookup[j] - Cells      */
    const int32_t*  cell_n;           /* Cell.n                     */
    const int32_t*  cell_member_start;/* offset of cell j in cell_members (n_cells+1 entries) */
    const int32_t*  cell_members;     /* Cell.members (concatenated) */
    /* microcells */
    const int32_t*  mcell_adunit;     /* Microcell.adunit */
    /* places */
    int32_t         n_places[CS_MAX_PLACE_TYPES];
    const int32_t*  place_n[CS_MAX_PLACE_TYPES];
    const int32_t*  place_mcell[CS_MAX_PLACE_TYPES];
    const int32_t*  place_member_start[CS_MAX_PLACE_TYPES];  /* n_places+1 */
    const int32_t*  place_members[CS_MAX_PLACE_TYPES];
    const int32_t*  place_groups_start[CS_MAX_PLACE_TYPES];  /* n_places+1, into group arrays */
    const int32_t*  place_group_start[CS_MAX_PLACE_TYPES];
    const int32_t*  place_group_size[CS_MAX_PLACE_TYPES];
} cs_population;

/*
 * Dynamic state of one realisation (the mutable fields of Person, Cell and
 * Place that the path reads or writes).  Used to load a realisation from a
 * host that initialised it (InitModel + SeedInfection) and to copy it back.
 * Person arrays: n_people entries.  cell_* : n_cells entries (populated-cell
 * order).  cell_perm: the permutation array of every populated cell,
 * concatenated with the offsets of cs_population.cell_member_start
 * (Cell.susceptible[0..n) = S | L | I | R+D segments).
 */
typedef struct cs_state {
    int8_t*   inf;                    /* InfStat code        */
    uint8_t*  severity_current;
    uint8_t*  severity_final;
    uint8_t*  bits;                   /* bit0 to_die, bit1 detected */
    int32_t*  listpos;
    int32_t*  infector;
    int16_t*  infect_type;
    float*    infectiousness;
    float*    susc;
    float*    prob_care;              /* Person.ProbCare */
    uint16_t* infection_time;
    uint16_t* latent_time;
    uint16_t* recovery_or_death_time;
    uint16_t* sari_time;
    uint16_t* critical_time;
    uint16_t* stepdown_time;
    uint16_t* detected_time;
    uint16_t* absent_start_time;
    uint16_t* absent_stop_time;
    /* intervention fields read by the hazards (only when interventions_active) */
    uint16_t* treat_start_time;
    uint16_t* treat_stop_time;
    uint16_t* vacc_start_time;
    uint16_t* isolation_start_time;
    uint16_t* quar_start_time;        /* HostsQuarantine[].start_time */
    uint8_t*  quar_comply;            /* HostsQuarantine[].comply     */
    uint8_t*  dct_traced;             /* Person.digitalContactTraced  */
    /* cells */
    int32_t*  cell_S; int32_t* cell_L; int32_t* cell_I; int32_t* cell_R; int32_t* cell_D;
    int32_t*  cell_cumTC;
    int32_t*  cell_perm;
    /* places (per type, n_places entries) */
    uint16_t* place_close_start[CS_MAX_PLACE_TYPES];
    uint16_t* place_close_end[CS_MAX_PLACE_TYPES];
    /* microcells (n_mcells entries; intervention state read by hazards) */
    uint8_t*  mcell_socdist;          /* TreatStat */
    uint8_t*  mcell_moverest;
} cs_state;

/*
 * Counters a step adds to (the StateT fields of src/Model.h:68-122 that the
 * path writes).  cs_get_counters returns the totals accumulated since the
 * last call with reset != 0.  Age arrays have CS_NUM_AGE_GROUPS entries.
 */
#define CS_NUM_SEV_KINDS 8
enum { CS_SK_MILD = 0, CS_SK_ILI, CS_SK_SARI, CS_SK_CRITICAL, CS_SK_CRITRECOV,
       CS_SK_DEATH_ILI, CS_SK_DEATH_SARI, CS_SK_DEATH_CRITICAL };
typedef struct cs_counters {
    int64_t cumI, cumC, cumDC, cumFC, cumTG, cumSI, nTG;
    int64_t cumAC, cumACS, cumAA;
    double  sumRad2, maxRad2;
    int64_t cumItype[CS_INFECT_TYPE_MASK];
    int64_t cumIa[CS_NUM_AGE_GROUPS], cumCa[CS_NUM_AGE_GROUPS], cumDa[CS_NUM_AGE_GROUPS];
    int64_t cumI_keyworker[2], cumC_keyworker[2];
    /* severity prevalence change (signed) and cumulative incidence, total and by age */
    int64_t sev_prev[CS_NUM_SEV_KINDS], sev_cum[CS_NUM_SEV_KINDS];
    int64_t sev_prev_age[CS_NUM_SEV_KINDS][CS_NUM_AGE_GROUPS];
    int64_t sev_cum_age[CS_NUM_SEV_KINDS][CS_NUM_AGE_GROUPS];
    /* population totals after the step */
    int64_t S, L, I, R, D;
    /* intervention counters (StateT.cumHQ, cumAH, cumAPC, cumAPA, cumAPCS):
     * households quarantined, adults absent for a quarantined household,
     * adults and children absent from closed places, adults absent to look
     * after children of closed places */
    int64_t cumHQ, cumAH, cumAPC, cumAPA, cumAPCS;
} cs_counters;

/* Event-count tallies used for the algorithmic-bytes roofline figure. */
typedef struct cs_work_counts {
    int64_t n_inf, n_hh_inf, n_hh_member, n_place_link, n_place_cand;
    int64_t n_sp_draw, n_sp_susc, n_new, n_active, n_transition, n_cells;
    int64_t n_sp_scan;   /* cum_trans entries read by the destination-cell scans */
    int64_t n_sp_probe;  /* cumulative-infectiousness entries read by the infector searches */
    int64_t n_hh_scan;   /* household members read by the household scans: a household whose scan found no
                            susceptible member is not scanned again for that infector (no one returns to
                            susceptible), so this is at most n_hh_member */
} cs_work_counts;

typedef struct cs_engine cs_engine;

const char* cs_last_error(const cs_engine* e);
const char* cs_version(void);

/* lifetime */
int  cs_create(cs_engine** out, int device);
void cs_destroy(cs_engine* e);

/* set-up: parameters, kernel tables (KernelLookup lookup_/hi_res_, each
 * kernel_size+1 doubles; NULL builds them here from the kernel fields of
 * cs_params, src/Kernels.cpp:17-45), population (copied to HBM).  The cell
 * max_trans table (src/Kernels.cpp:48-63) is built on the device.
 * cs_set_population checks every index the kernels follow and returns
 * CS_EINVAL naming the first bad one; households hold at most
 * CS_MAX_HOUSEHOLD_SIZE persons (hh_nh and hh_nhr), as the reference's
 * household table does (src/SetupModel.cpp:1287-1288). */
int  cs_set_params(cs_engine* e, const cs_params* p, const double* kernel_lookup, const double* kernel_hi_res);
int  cs_set_population(cs_engine* e, const cs_population* pop);

/* realisations: allocate n_real realisations, each keyed by its own 64-bit
 * seed.  cs_reset_realisations runs InitModel's per-person reset
 * (src/CovidSim.cpp:1006-1037) and SeedInfection's random-location branch
 * (src/CovidSim.cpp:1244-1381) on the device for every realisation, then
 * UpdateProbs (src/CovidSim.cpp:2752-2802). */
int  cs_alloc_realisations(cs_engine* e, int n_real, const uint64_t* seeds);
int  cs_reset_realisations(cs_engine* e, int n_seed_infections);
/* per-realisation parameters for sweeps and calibration (the reference runs
 * one parameter set per process; here every realisation may differ):
 * local_beta[r] replaces cs_params.local_beta (src/Sweep.cpp:300) and
 * n_seed[r] the seed count of the next cs_reset_realisations for
 * realisation r.  Either may be NULL (no override); each call replaces the
 * previous overrides; cs_alloc_realisations clears them. */
int  cs_set_realisation_params(cs_engine* e, const double* local_beta, const int32_t* n_seed);
/* per-realisation intervention trigger thresholds: threshold[r] replaces
 * the three thresholds the global trigger value is compared with
 * (cs_params.case_isolation_cell_inc_thresh, hhquar_cell_inc_thresh and
 * place_close_cell_inc_thresh; src/CovidSim.cpp:3640-3660 and
 * src/Sweep.cpp:1426-1437 read P.CaseIsolation_CellIncThresh,
 * P.HHQuar_CellIncThresh and P.PlaceCloseCellIncThresh) for realisation r,
 * as a threshold sweep of the reference would run one process per value.
 * NULL clears the overrides; cs_alloc_realisations clears them too. */
int  cs_set_realisation_triggers(cs_engine* e, const double* threshold);

/* host-initialised state in and out (drop-in use under CovidSim's RunModel) */
int  cs_load_state(cs_engine* e, int real, const cs_state* s);
int  cs_store_state(cs_engine* e, int real, cs_state* s);

/* the two sweeps the library replaces; t = CurrSimTime in days */
int  cs_infect_sweep(cs_engine* e, double t);
int  cs_incub_recovery_sweep(cs_engine* e, double t);
/* UpdateProbs(0) for every realisation (src/CovidSim.cpp:2752-2802) */
int  cs_update_probs(cs_engine* e);
/* RunModel's day loop without host round trips: n_steps timesteps of
 * InfectSweep + IncubRecoverySweep + the S-drop UpdateProbs check
 * (src/CovidSim.cpp:1426-1522).  steps_per_sample (> 0) is the number of
 * timesteps per output step: totals are recorded after each output step,
 * and, as in RunModel, the S-drop check reads S as of the last sample.
 * Successive calls continue one series (continuing calls should start at a
 * sample boundary); cs_alloc_realisations, cs_reset_realisations and
 * cs_load_state start a new one.  By default the launches of one output step
 * are captured once into a HIP graph and replayed on a device-side clock
 * (results are identical either way; see cs_set_step_graphs). */
int  cs_run_steps(cs_engine* e, double t0, int n_steps, int steps_per_sample);

/* A batch: engines of one device, each a region of its own (population,
 * parameters, realisations, per-realisation parameters), advanced together
 * by cs_batch_run_steps: every output step, each engine's captured step
 * graph is launched on the engine's own stream, all of them before any
 * wait, so that regions of uneven size run side by side on the device's
 * hardware queues.  The reference runs one region per
 * process and its realisations one after another (/NR, src/CovidSim.cpp:
 * 333-391).  Every engine's results equal those of cs_run_steps on it alone
 * (a region's streams are its own).  The batch does not own the engines;
 * each stays usable alone (its series, counters and state as after
 * cs_run_steps), and must outlive the batch.  An engine that overflows a
 * work list is grown and brought level directly, as cs_run_steps does.
 * Engines with step graphs off or profiling on run one after another. */
typedef struct cs_batch cs_batch;
int  cs_batch_create(cs_batch** out, cs_engine* const* engines, int n);
void cs_batch_destroy(cs_batch* b);
const char* cs_batch_last_error(const cs_batch* b);
int  cs_batch_run_steps(cs_batch* b, double t0, int n_steps, int steps_per_sample);
/* destination tables of the spatial draws, chosen at the next
 * cs_alloc_realisations.  stride 0: the reference's full tables, max_trans
 * [NC][NC] and per realisation cum_trans [NC][NC] with InvCDF [NC][1025]
 * (src/CovidSim.cpp:2752-2802, src/Kernels.cpp:48-63); stride > 0:
 * checkpointed tables, for regions whose NC x NC tables do not fit (NC =
 * 224,448 cells makes one table 201 GB): max_trans is evaluated where it is
 * read and each cum_trans row keeps its sequential float prefix at every
 * stride-th cell only (R x NC x ceil(NC / stride) floats), a destination
 * search redoing one block's prefixes in the same order.  Both give the
 * same destinations for the same uniforms, bit for bit.  -1 (default):
 * full tables while (realisations + 1) x NC^2 floats fit 16 GiB, else the
 * smallest power-of-two stride (at least 8) whose checkpoints and bucket
 * index fit 16 GiB.  cs_get_dest_tables gives the stride in use after an
 * allocation (the requested value before).  A search finds the first cell
 * (or block) whose normalised prefix reaches u by binary search within the
 * bracket of u's bucket of 1/1024: InvCDF with full tables; with
 * checkpoints, a per-row index of each bucket's first block (R x NC x 1025
 * int32), kept while it fits 4 GiB (cs_set_dest_index(e, 0) drops it, the
 * search then spanning the row's blocks).  Both choices, for the next
 * cs_alloc_realisations, leave every destination unchanged. */
int  cs_set_dest_tables(cs_engine* e, int stride);
int  cs_get_dest_tables(const cs_engine* e, int* stride);
int  cs_set_dest_index(cs_engine* e, int enable);
/* 1 (default): cs_run_steps replays captured step graphs; 0: it launches
 * every kernel of every step from the host */
int  cs_set_step_graphs(cs_engine* e, int enable);
/* capacity of the infection event queue and of the spatial draw list (in
 * chunks of up to 256 draws of one source cell); 0 keeps a value.  The
 * event queue is 8 equal shards (appends spread over 8 counts), so its
 * capacity is rounded up to a multiple of 8.  Both grow by doubling when a
 * step overflows them (the step is re-run, with the same result), so
 * capacities only trade memory against re-runs. */
int  cs_set_queue_capacity(cs_engine* e, uint32_t events, uint32_t draw_chunks);
int  cs_get_queue_capacity(cs_engine* e, uint32_t* events, uint32_t* draw_chunks);
/* capacity of each step's latent due list: the persons whose latency ends in
 * that step, listed when they are infected, so that the step visits only
 * them (after a reset; a loaded state is scanned in full until the next
 * reset).  The list is 16 shards of per_step / 16 persons (rounded up), a
 * block appending to one shard; a step one of whose shards overflows scans
 * every latent person instead, with the same result; 0 keeps the value.  No counterpart in the reference,
 * which scans every latent person every step (src/Sweep.cpp:922-940). */
int  cs_set_due_capacity(cs_engine* e, uint32_t per_step);
/* launch shapes: blocks of the per-person passes (4 waves each, one wave
 * per chunk of 64 persons or 256 spatial draws) and of the per-cell spatial
 * sums (4 waves each, one wave per cell); 0 keeps a value.  Results do not
 * depend on them. */
int  cs_set_launch_shape(cs_engine* e, uint32_t grid_waves, uint32_t grid_cells);
/* the batched list moves of a step (the reference's per-person list swaps
 * in DoInfect, DoIncub, DoRecover and DoDeath, src/Update.cpp:100-400): a
 * cell with at most per_cell moves (default and maximum 4096) is moved in
 * LDS, a larger one through a bitmap of its list segment in device memory
 * (sized at cs_alloc_realisations from the largest cell); both pair the
 * positions alike, so results do not depend on per_cell (0 sends every
 * cell through device memory).  No counterpart in the reference. */
int  cs_set_move_capacity(cs_engine* e, int per_cell);
/* autotuner: at the current state and time t, times the InfectSweep event
 * phase (it reads the state and writes only scratch, so it can be repeated)
 * for each candidate launch shape, best of reps, keeps the largest shape
 * within 3% of the fastest (timing noise on a small state must not pick a
 * grid too small for later, larger states), and sizes the event queue at 4
 * times this state's events */
typedef struct cs_tuning {
    uint32_t grid_waves, grid_cells;
    uint32_t event_capacity, draw_capacity;
    uint32_t events_at_state;
    float    event_phase_ms;
} cs_tuning;
int  cs_autotune(cs_engine* e, double t, int reps, cs_tuning* out);

/* results */
int  cs_get_counters(cs_engine* e, int real, cs_counters* out, int reset);
/* the intervention state of realisation real (the P and AdUnits fields that
 * CalibrationThresholdCheck sets, src/CovidSim.cpp:3485-3689, one set per
 * realisation here): whether the alert has been reached, the start times of
 * case isolation, household quarantine and place closure (days; 1e10: not
 * set) and the trigger value of the last sample (State.trigDetectedCases) */
typedef struct cs_intervention_state {
    int32_t alert;                     /* the alert condition has held at a sample */
    int32_t trigger;                   /* State.trigDetectedCases of the last sample */
    double  control_prop_cases_id;     /* P.ControlPropCasesId now */
    double  case_isolation_start, hquarantine_start, place_close_start;
    int32_t mcells_closed;             /* microcells whose places are closed now (TreatStat::Treated) */
    int32_t pad;
} cs_intervention_state;
int  cs_get_intervention_state(cs_engine* e, int real, cs_intervention_state* out);
/* series [n_real][n_samples][CS_TS_FIELDS] recorded by cs_run_steps since the
 * series last started; out holds max_samples per realisation.  Per sample
 * (RecordSample, src/CovidSim.cpp:3024-3483):
 *   0-4    S, L, I, R, D totals at the sample
 *   5-7    incidence since the previous sample: infections, cases, deaths
 *   8-15   incidence since the previous sample of Mild, ILI, SARI, Critical,
 *          CritRecov, and of deaths from ILI, SARI and Critical (the
 *          incMild .. incDeath_Critical fields)
 *   16-20  prevalence of Mild, ILI, SARI, Critical, CritRecov, counted from
 *          the last zeroing of the counters (cs_reset_realisations, or
 *          cs_get_counters with reset), as StateT's running totals are
 * The incidence fields are differences of the cumulative counters, so a
 * series must not be interleaved with cs_get_counters(..., reset != 0). */
#define CS_TS_FIELDS 21
enum { CS_TS_S = 0, CS_TS_L, CS_TS_I, CS_TS_R, CS_TS_D, CS_TS_INC_I, CS_TS_INC_C, CS_TS_INC_DEATH,
       CS_TS_INC_SEV = 8,    /* + CS_SK_* kind, 8 kinds */
       CS_TS_PREV_SEV = 16 }; /* + CS_SK_MILD .. CS_SK_CRITRECOV */
#define CS_TS_LAST 11   /* cumulative counters a sample differences: cumI, cumC, D, 8 severity kinds */
int  cs_get_timeseries(cs_engine* e, int32_t* out, int max_samples, int* n_samples);
/* the same samples by age group and by administrative unit (RecordSample's
 * by-age and by-unit fields, src/CovidSim.cpp:3024-3483, from StateT's
 * cumIa/cumCa/cumDa, the severity counters by age, and cumI_adunit,
 * cumC_adunit, cumDC_adunit and cumD_adunit, src/Update.cpp:159, :872,
 * :892, :976).  By age: out [n_real][max_samples][CS_TSA_FIELDS]
 * [CS_NUM_AGE_GROUPS]; by unit: out [n_real][max_samples][CS_TSU_FIELDS]
 * [n_adunits], kept when do_ad_units is set and every person's unit lies in
 * [0, n_adunits) (CS_ESTATE otherwise). */
#define CS_TSA_FIELDS 16
enum { CS_TSA_INC_I = 0, CS_TSA_INC_C, CS_TSA_INC_DEATH,
       CS_TSA_INC_SEV = 3,     /* + CS_SK_* kind, 8 kinds */
       CS_TSA_PREV_SEV = 11 }; /* + CS_SK_MILD .. CS_SK_CRITRECOV */
#define CS_TSA_CUM 11   /* by-age fields a sample differences: the incidences before CS_TSA_PREV_SEV */
#define CS_TSU_FIELDS 4
enum { CS_TSU_INC_I = 0, CS_TSU_INC_C, CS_TSU_INC_DC, CS_TSU_INC_DEATH };
int  cs_get_timeseries_age(cs_engine* e, int32_t* out, int max_samples, int* n_samples);
int  cs_get_timeseries_adunit(cs_engine* e, int32_t* out, int max_samples, int* n_samples);
/* realisation real's counters by administrative unit: out [n_adunits][4] =
 * cumI, cumC, cumDC, cumD (StateT's *_adunit counters); reset != 0 zeroes
 * them (as cs_get_counters, not to be interleaved with a series) */
int  cs_get_adunit_counters(cs_engine* e, int real, int64_t* out, int reset);
int  cs_get_work_counts(cs_engine* e, cs_work_counts* out, int reset);
int  cs_sync(cs_engine* e);

/* InitTransmissionCoeffs (src/SetupModel.cpp:640-846): the reproduction
 * numbers of one pass over the population with per-person draws of
 * infectiousness and infectious period, and the spatial beta that makes the
 * total equal cs_params.r0 (unless fix_local_beta).  The draws come from the
 * counter-based stream keyed by `seed`.  prop_age_group is P.PropAgeGroup
 * ([n_adunits][CS_NUM_AGE_GROUPS]), read only when got_waifw_spatial; may be
 * NULL otherwise.  On success, unless fix_local_beta, the engine's local_beta
 * is set to the result (the reference sets P.LocalBeta). */
typedef struct cs_transmission_coeffs {
    double r0_household;               /* P.R0household */
    double household_sar;              /* household secondary attack rate */
    double r0_places;                  /* P.R0places */
    double r0_places_by_type[CS_MAX_PLACE_TYPES];
    double spatial_r0;                 /* spatial R0 per unit local_beta */
    double local_beta;                 /* P.LocalBeta */
    double r0_spatial;                 /* P.R0spatial */
    double r0;                         /* overall R0 */
    double mean_infectious_days;       /* mean sampled infectious period */
    double mean_infectious_steps;
} cs_transmission_coeffs;
int  cs_init_transmission_coeffs(cs_engine* e, const double* prop_age_group, uint64_t seed,
                                 cs_transmission_coeffs* out);
/* the same pass from given per-person draws: infectiousness (signed: <0 is
 * symptomatic, as Person.infectiousness), infectious period in timesteps
 * (Person.recovery_or_death_time) and, optionally, the InfStat code that
 * decides which household members count as susceptible (NULL: all).  The
 * engine's parameters are not changed. */
int  cs_transmission_coeffs_from_draws(cs_engine* e, const double* prop_age_group, const float* infectiousness,
                                       const uint16_t* infectious_steps, const int8_t* infstat,
                                       cs_transmission_coeffs* out);

/* ---- population generator ----------------------------------------------
 * The set-up the reference runs before its day loop, on the device:
 * SetupPopulation (src/SetupModel.cpp:848-1642: microcell populations from
 * the density, households, household locations, AssignHouseholdAges
 * :1820-1988, the age correction, places and their locations),
 * AssignPeopleToPlaces (:1989-2466: eligibility, place sizes, the capacity
 * adjustment, the nearest-places choice and the cell-gravity assignment)
 * and StratifyPlaces (:2468-2601: groups).  Its random streams are the
 * library's counter-based ones, so a population equals the reference's in
 * law, not draw for draw.  The concurrent pieces keep the reference's law
 * where the reference is sequential: microcell and place counts are one
 * multinomial, split as a tree of binomials; capacity changes draw their
 * places independently; the cell-gravity assignment runs in rounds in which
 * every unplaced person proposes slots as the reference's loop does
 * (destination cell by the capacity-weighted cell tables, a uniform slot,
 * acceptance by kernel over max_trans) and a slot goes to its first
 * claimant, the tables being rebuilt from the slots left after each round.
 * Unsupported (CS_EUNSUPPORTED): school files (nsp > 0), regional
 * demography files, and InhibitInterAdunitPlaceAssignment. */
typedef struct cs_popgen_params {
    int32_t pop_size;                  /* P.PopSize */
    int32_t mcw, mch;                  /* P.total_microcells_wide_/high_ */
    int32_t nmcl;                      /* P.NMCL */
    int32_t ncw, nch;                  /* P.ncw, P.nch */
    double  mc_width, mc_height;       /* P.in_microcells_ (degrees) */
    /* reading the density (src/SetupModel.cpp:73-103, :882-940): the box's
     * bottom-left corner, the longitude cut and the administrative-unit codes */
    double  bbox_x0, bbox_y0;          /* P.SpatialBoundingBox.bottom_left() */
    double  longitude_cut_line;        /* P.LongitudeCutLine */
    int32_t do_ad_units, do_adunit_boundaries;      /* P.DoAdUnits, P.DoAdunitBoundaries */
    int32_t adunit_level1_mask, adunit_level1_divisor;  /* P.AdunitLevel1Mask, P.AdunitLevel1Divisor */
    int32_t do_households;             /* P.DoHouseholds */
    int32_t do_correct_age_dist;       /* P.DoCorrectAgeDist */
    double  household_size_cdf[CS_MAX_HOUSEHOLD_SIZE];  /* P.HouseholdSizeDistrib[0], cumulative */
    double  prop_age_group[CS_NUM_AGE_GROUPS];          /* P.PropAgeGroup[0] */
    /* AssignHouseholdAges */
    double  one_pers_house_prob_old, one_pers_house_prob_young;
    double  two_pers_house_prob_old, two_pers_house_prob_young;
    double  one_child_two_pers_prob, zero_child_three_pers_prob, two_child_three_pers_prob;
    double  one_child_four_pers_prob, three_child_five_pers_prob;
    double  one_child_prob_youngest_under_five, two_children_prob_youngest_under_five;
    double  prob_youngest_child_under_five, young_and_single_slope;
    int32_t no_child_pers_age, old_pers_age, young_and_single, min_adult_age;
    int32_t max_mf_partner_age_gap, max_fm_partner_age_gap, min_parent_age_gap, max_parent_age_gap;
    int32_t max_child_age, mean_child_age_gap, older_gen_gap;
    /* places */
    int32_t num_place_types;           /* P.NumPlaceTypes */
    int32_t hotel_place_type;          /* P.HotelPlaceType (>= num_place_types: none) */
    int32_t nsp;                       /* P.nsp, places read from a school file (must be 0) */
    int32_t place_age_min[3][CS_MAX_PLACE_TYPES];   /* PlaceTypeAgeMin, ...Min2, ...Min3 */
    int32_t place_age_max[3][CS_MAX_PLACE_TYPES];
    double  place_prop[3][CS_MAX_PLACE_TYPES];      /* PlaceTypePropAgeGroup, 2, 3 */
    double  place_mean_size[CS_MAX_PLACE_TYPES];
    double  place_size_power[CS_MAX_PLACE_TYPES], place_size_offset[CS_MAX_PLACE_TYPES];
    double  place_size_sd[CS_MAX_PLACE_TYPES];
    double  place_size_min[CS_MAX_PLACE_TYPES], place_size_max[CS_MAX_PLACE_TYPES];
    int32_t place_nearest[CS_MAX_PLACE_TYPES];      /* PlaceTypeNearestNeighb */
    int32_t place_kernel_type[CS_MAX_PLACE_TYPES];
    double  place_kernel_scale[CS_MAX_PLACE_TYPES], place_kernel_shape[CS_MAX_PLACE_TYPES];
    double  place_kernel_p3[CS_MAX_PLACE_TYPES], place_kernel_p4[CS_MAX_PLACE_TYPES];
    double  place_group_size[CS_MAX_PLACE_TYPES];   /* PlaceTypeGroupSizeParam1 */
    int32_t inhibit_inter_adunit[CS_MAX_PLACE_TYPES];  /* InhibitInterAdunitPlaceAssignment != 0 (must be 0) */
    int32_t setup_seed1, setup_seed2;  /* P.setupSeed1/2 (network file header only) */
} cs_popgen_params;

/* sizes of a generated population (the arrays of cs_population) */
typedef struct cs_popgen_sizes {
    int32_t n_people, n_households, n_cells, n_cells_total, n_mcells;
    int32_t n_places[CS_MAX_PLACE_TYPES];
    int64_t n_members[CS_MAX_PLACE_TYPES];   /* persons linked to each type */
    int64_t n_groups[CS_MAX_PLACE_TYPES];
    int64_t n_unplaced[CS_MAX_PLACE_TYPES];  /* eligible persons the assignment left without a place */
} cs_popgen_sizes;

/* host buffers of a generated population, sized by cs_popgen_sizes: the
 * cs_population arrays (hh, mcell, pcell, age, place_links [n][CS_MAX_PLACE_TYPES],
 * place_group_links, households, populated cells and their members,
 * mcell_adunit, places by type) plus the place locations ([n][2] floats,
 * relative to the bounding box as household locations are).  NULL
 * fields are not copied. */
typedef struct cs_popgen_out {
    int32_t *hh, *mcell, *pcell;
    uint8_t *age;
    int32_t *place_links;
    uint16_t *place_group_links;
    int32_t *hh_first;
    uint16_t *hh_nh;
    float   *hh_loc;
    int32_t *cell_index, *cell_n, *cell_member_start, *cell_members;
    int32_t *mcell_adunit;
    int32_t *place_n[CS_MAX_PLACE_TYPES], *place_mcell[CS_MAX_PLACE_TYPES];
    int32_t *place_member_start[CS_MAX_PLACE_TYPES], *place_members[CS_MAX_PLACE_TYPES];
    int32_t *place_groups_start[CS_MAX_PLACE_TYPES], *place_group_start[CS_MAX_PLACE_TYPES];
    int32_t *place_group_size[CS_MAX_PLACE_TYPES];
    float   *place_loc[CS_MAX_PLACE_TYPES];
} cs_popgen_out;

typedef struct cs_popgen cs_popgen;
int  cs_popgen_create(cs_popgen** out, int device);
void cs_popgen_destroy(cs_popgen* g);
const char* cs_popgen_last_error(const cs_popgen* g);
/* one population: kp supplies the distance geometry and the kernel table
 * size (cs_params: do_utm_coords, periodic boundaries, region size,
 * bounding-box latitude, kernel_size, kernel_delta, kernel_expansion; the
 * place kernels come from pg); per microcell (mcw x mch, index x * mch + y)
 * the mean density of its density points (0: none), its country and its
 * administrative unit index (SetupPopulation's mcell_dens / mcell_num,
 * mcell_country and mcell_adunits). */
int  cs_popgen_run(cs_popgen* g, const cs_popgen_params* pg, const cs_params* kp, const double* mcell_density,
                   const int32_t* mcell_country, const int32_t* mcell_adunit, uint64_t seed);
int  cs_popgen_get_sizes(cs_popgen* g, cs_popgen_sizes* out);
int  cs_popgen_fetch(cs_popgen* g, const cs_popgen_out* out);
/* AssignHouseholdAges for given household sizes (household k takes the
 * stream of household k of a population generated with `seed`); ages in
 * household order, sum(sizes) entries.  For parity tests. */
int  cs_popgen_household_ages(cs_popgen* g, const cs_popgen_params* pg, uint64_t seed, const int32_t* sizes,
                              int n, uint8_t* ages);

/* ---- calibration emulator ---------------------------------------------
 * No counterpart in the reference (its calibration re-runs the simulator,
 * src/CovidSim.cpp:3485-3689).  A multilayer perceptron
 * n_in -> n_hidden -> n_hidden -> n_out with GELU activations and a linear
 * output, evaluated and trained on the GPU: BF16 weights and activations,
 * FP32 accumulation on MFMA, FP32 pre-activations, gradients, master weights
 * and Adam moments.  Parameters are one FP32 vector: W1 [n_hidden][n_in],
 * b1, W2 [n_hidden][n_hidden], b2, W3 [n_out][n_hidden], b3.  Loss: mean
 * squared error over all outputs of a batch.  Host arrays are row-major
 * [batch][features]. */
typedef struct cs_mlp cs_mlp;
int  cs_mlp_create(cs_mlp** out, int device, int n_in, int n_hidden, int n_out, uint64_t seed);
void cs_mlp_destroy(cs_mlp* m);
const char* cs_mlp_last_error(const cs_mlp* m);
int64_t cs_mlp_n_params(const cs_mlp* m);
int  cs_mlp_get_params(cs_mlp* m, float* out);
/* sets the parameters and restarts the optimiser */
int  cs_mlp_set_params(cs_mlp* m, const float* in);
int  cs_mlp_forward(cs_mlp* m, const float* x, int batch, float* y);
/* forward with the intermediate layers: FP32 pre-activations z1, z2 and the
 * BF16 activations h1, h2 (raw bits), [batch][n_hidden] each; any may be NULL */
int  cs_mlp_forward_trace(cs_mlp* m, const float* x, int batch, float* z1, float* z2,
                          uint16_t* h1, uint16_t* h2, float* y);
/* Allocation-free forward.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory (tight serving loops such as a
 * calibration that scores candidates one batch at a time, harnesses that
 * forbid hipMalloc).  Same result as cs_mlp_forward, bit for bit, but x
 * ([batch][n_in] FP32) and y ([batch][n_out] FP32) are DEVICE pointers owned
 * by the caller, and the batch's intermediate activations live in a fixed
 * device array of CS_MLP_NOALLOC_SCRATCH_BYTES that each GPU holds for the
 * life of the process (set up by cs_mlp_create).  The call allocates and
 * frees nothing; it runs on the emulator's stream and returns when y is
 * written.  A batch needing more than the array holds (batch x (2 n_in +
 * 4 n_hidden) bytes, each part rounded up to 256) is refused with CS_EINVAL.
 * All calls on one GPU, from any emulator, share that GPU's array: calls that
 * overlap on one GPU race. */
#define CS_MLP_NOALLOC_SCRATCH_BYTES 16777216
int  cs_mlp_forward_noalloc(cs_mlp* m, const float* x, int batch, float* y);
/* loss and its gradient with respect to every parameter, for one batch */
int  cs_mlp_loss_grad(cs_mlp* m, const float* x, const float* t, int batch, double* loss, float* grad);
/* training set kept in HBM (inputs stored as BF16) */
int  cs_mlp_set_data(cs_mlp* m, const float* x, const float* t, int n);
/* epochs of minibatch Adam (beta 0.9/0.999, eps 1e-8) over the training set
 * in an order drawn from `seed`; loss_per_epoch (may be NULL) gets each
 * epoch's mean batch loss */
int  cs_mlp_train(cs_mlp* m, int epochs, int batch, float lr, uint64_t seed, double* loss_per_epoch);

/* per-kernel device time, measured with HIP events recorded around each
 * launch on the engine's stream while profiling is enabled (cs_run_steps
 * then launches every kernel directly instead of replaying step graphs) */
enum { CS_K_INFECT_HOUSEHOLDS = 0, CS_K_SPATIAL_DRAWS, CS_K_RESOLVE, CS_K_LIST_MOVE,
       CS_K_LATENT_SCAN, CS_K_INFECTIOUS_SCAN, CS_K_UPDATE_PROBS, CS_K_CELL_SPATIAL_SUM,
       CS_K_INFECT_PLACES, CS_K_OTHER, CS_NUM_KERNEL_KINDS };
int  cs_set_profiling(cs_engine* e, int enable);
/* ms[k] total milliseconds and launches[k] launch count of kernel kind k
 * since profiling was last enabled or reset */
int  cs_get_profile(cs_engine* e, double* ms, int64_t* launches, int reset);

/* ---- direct evaluation of path components (parity tests) -------------- */
/* KernelLookup::num for n squared distances (src/Kernels.cpp:108-129) */
int  cs_kernel_num(cs_engine* e, const double* r2, double* out, int n);
/* dist2 between persons (src/Dist.cpp:49-60) */
int  cs_dist2_persons(cs_engine* e, const int32_t* a, const int32_t* b, double* out, int n);
/* cell tables of realisation real: max_trans [n_cells][n_cells],
 * cum_trans [n_cells][n_cells], invcdf [n_cells][1025], tot_prob, S0 */
int  cs_get_cell_tables(cs_engine* e, int real, float* max_trans, float* cum_trans,
                        int32_t* invcdf, float* tot_prob, int32_t* S0);
/* hazards (src/CalcInfSusc.cpp:12-80) of realisation real at TimeStepNow ts.
 * kind: 0 HouseInf(person) 1 PlaceInf(person,pt) 2 SpatialInf(person)
 *       3 PersonInf(person) 4 HouseSusc(person,infector) 5 PlaceSusc(person,pt)
 *       6 SpatialSusc(person) 7 PersonSusc(person,infector) */
int  cs_eval_hazards(cs_engine* e, int real, int kind, uint16_t ts, const int32_t* person,
                     const int32_t* other, double* out, int n);
/* InverseCdf::choose for given uniforms u (src/InverseCdf.cpp:60-74) */
int  cs_icdf_choose(cs_engine* e, int table, const double* u, const double* mean,
                    uint16_t* out, int n);
/* spatial indexing for given uniforms (src/Sweep.cpp:716-768): infector
 * index within a normalised cumulative list, destination cell from the
 * InvCDF + cum_trans scan of source cell `cell` in realisation real. */
int  cs_select_infector(cs_engine* e, const float* cell_inf, int n_inf, const double* u,
                        int32_t* out, int n);
int  cs_select_dest_cell(cs_engine* e, int real, int cell, const double* u, int32_t* out, int n);
/* raw counter-based RNG output (Philox4x32-10) for known-answer tests:
 * ctr_key holds n records {c0, c1, c2, c3, k0, k1}, out n * 4 words */
int  cs_philox(cs_engine* e, const uint32_t* ctr_key, uint32_t* out, int n);
/* sampler draws for distribution tests: out[i] is one draw from the stream
 * {key, c0, c1, i}; which: 0 uniform, 1 Binomial(a, b), 2 Poisson(a),
 * 3 standard normal, 4 Gamma(rate a, shape b) */
int  cs_sample(cs_engine* e, int which, uint64_t key, uint32_t c0, uint32_t c1,
               double a, double b, double* out, int n);

#ifdef __cplusplus
}
#endif
#endif /* COVIDSIM_HIP_H_ */
