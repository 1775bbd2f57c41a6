/*
 * cs_oracle.h -- CPU oracle for the infection step.  TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the CPU baseline, never as the thing
 * measured or shipped.
 *
 * A sequential C restatement of CovidSim's InfectSweep / IncubRecoverySweep
 * path (src/Sweep.cpp:283-979, src/Update.cpp, src/CalcInfSusc.cpp,
 * src/Kernels.cpp, src/Dist.cpp, src/InverseCdf.cpp, UpdateProbs in
 * src/CovidSim.cpp:2752-2802) over the cs_params / cs_population /
 * cs_state views of include/covidsim_hip.h.  Random numbers come from the
 * same counter-based streams as the GPU library (Philox4x32-10 keyed by the
 * realisation seed, counter = timestep/pass, entity, sub, block), so for the
 * same inputs the oracle and the GPU must produce identical trajectories.
 */
#ifndef CS_ORACLE_H_
#define CS_ORACLE_H_
#include "../include/covidsim_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_engine or_engine;

int  or_create(or_engine** out, const cs_params* p, const double* kernel_lookup,
               const double* kernel_hi_res, const cs_population* pop);
void or_destroy(or_engine* e);
const char* or_last_error(const or_engine* e);
int  or_alloc_realisations(or_engine* e, int n_real, const uint64_t* seeds);
int  or_set_realisation_params(or_engine* e, const double* local_beta, const int32_t* n_seed);
int  or_set_realisation_triggers(or_engine* e, const double* threshold);
int  or_reset_realisations(or_engine* e, int n_seed_infections);
int  or_load_state(or_engine* e, int real, const cs_state* s);
int  or_store_state(or_engine* e, int real, cs_state* s);
int  or_infect_sweep(or_engine* e, double t);
int  or_incub_recovery_sweep(or_engine* e, double t);
int  or_update_probs(or_engine* e);
int  or_run_steps(or_engine* e, double t0, int n_steps, int steps_per_sample);
int  or_get_counters(or_engine* e, int real, cs_counters* out, int reset);
int  or_get_timeseries(or_engine* e, int32_t* out, int max_samples, int* n_samples);
int  or_get_timeseries_age(or_engine* e, int32_t* out, int max_samples, int* n_samples);
int  or_get_timeseries_adunit(or_engine* e, int32_t* out, int max_samples, int* n_samples);
int  or_get_adunit_counters(or_engine* e, int real, int64_t* out, int reset);
int  or_get_work_counts(or_engine* e, cs_work_counts* out, int reset);
/* the intervention subset: NULL when p lies inside it, else the reason */
const char* or_interventions_unsupported(const cs_params* p);
int  or_get_intervention_state(or_engine* e, int real, cs_intervention_state* out);

/* InitTransmissionCoeffs: draws + sums, and the sums from given draws */
int  or_init_transmission_coeffs(or_engine* e, const double* prop_age_group, uint64_t seed,
                                 cs_transmission_coeffs* out, float* susc_out, float* infness_out, uint16_t* rec_out);
int  or_transmission_coeffs_from_draws(or_engine* e, const double* prop_age_group, const float* infness,
                                       const uint16_t* rec, const int8_t* infstat, cs_transmission_coeffs* out);

int  or_kernel_num(or_engine* e, const double* r2, double* out, int n);
int  or_dist2_persons(or_engine* e, const int32_t* a, const int32_t* b, double* out, int n);
int  or_get_cell_tables(or_engine* e, int real, float* max_trans, float* cum_trans,
                        int32_t* invcdf, float* tot_prob, int32_t* S0);
int  or_eval_hazards(or_engine* e, int real, int kind, uint16_t ts, const int32_t* person,
                     const int32_t* other, double* out, int n);
int  or_icdf_choose(or_engine* e, int table, const double* u, const double* mean, uint16_t* out, int n);
int  or_select_infector(const float* cell_inf, int n_inf, const double* u, int32_t* out, int n);
int  or_select_dest_cell(or_engine* e, int real, int cell, const double* u, int32_t* out, int n);
void or_philox(const uint32_t* ctr_key, uint32_t* out, int n);
/* samplers exposed for distribution tests: draws from stream (key, c0..c2) */
int  or_sample(int which, uint64_t key, uint32_t c0, uint32_t c1, uint32_t c2,
               double a, double b, int n, double* out);
/* AssignHouseholdAges (src/SetupModel.cpp:1820-1988) on the generator's
 * streams, for given household sizes */
int  or_household_ages(const cs_popgen_params* p, uint64_t seed, const int32_t* sizes, int n, uint8_t* ages);
/* the nearest-places choice of AssignPeopleToPlaces (src/SetupModel.cpp:2250-2350)
 * for given persons (household location [n][2], microcell), places of type tp
 * by microcell (pl_off over pg->mcw * pg->mch microcells, locations [][2]) */
int  or_nearest_places(const cs_params* kp, const cs_popgen_params* pg, int tp, uint64_t seed, const int32_t* persons,
                       const float* hloc, const int32_t* pmcell, int n, const int64_t* pl_off, const float* pl_loc,
                       const int32_t* mc_country, int32_t* out);
double or_log(double x);
double or_exp(double x);

#ifdef __cplusplus
}
#endif
#endif
