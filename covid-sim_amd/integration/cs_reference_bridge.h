// cs_reference_bridge.h -- copies between CovidSim's global model state and
// the C ABI's flat views (include/covidsim_hip.h).
//
// For translation units compiled inside a CovidSim build: include it after
// the reference's own headers (Param.h, Model.h, Kernels.h with
// KernelLookup's tables readable).  It is used by the drop-in shim
// (cs_shim.cpp) and by the test harness that writes golden vectors, so both
// see CovidSim's state through exactly the same field mapping.
//
//   fill_params(c)        Param P (src/Param.h) -> cs_params
//   PopulationView        Hosts/Households/CellLookup/Mcells/Places
//                         (src/Model.h:295-303) -> cs_population
//   StateView             the mutable fields of Person, PersonQuarantine,
//                         Cell, Place and Microcell -> cs_state, and back
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "covidsim_hip.h"

namespace cs_bridge {

inline void copy_icdf(double* dst, InverseCdf& src)
{
    for (int i = 0; i <= CDF_RES; i++) dst[i] = src[i];
}

// cs_params from the reference's Param P as the run sees it (after ReadParams
// and InitModel's run-start updates).
inline void fill_params(cs_params* c)
{
    memset(c, 0, sizeof(*c));
    c->time_steps_per_day = P.TimeStepsPerDay;
    c->model_time_step = P.ModelTimeStep;
    c->false_positive_rate = P.FalsePositiveRate;
    c->do_seasonality = P.DoSeasonality;
    for (int i = 0; i < DAYS_PER_YEAR; i++) c->seasonality[i] = P.DoSeasonality ? P.Seasonality[i] : 1.0;
    c->local_beta = P.LocalBeta;
    c->household_trans = P.HouseholdTrans;
    for (int i = 0; i < MAX_HOUSEHOLD_SIZE; i++) c->household_denom_lookup[i] = P.HouseholdDenomLookup[i];
    c->do_households = P.DoHouseholds;
    c->do_places = P.DoPlaces;
    c->num_place_types = P.NumPlaceTypes;
    c->hotel_place_type = P.HotelPlaceType;
    c->place_type_no_air_num = P.PlaceTypeNoAirNum;
    c->care_home_place_type = P.CareHomePlaceType;
    c->do_symptoms = P.DoSymptoms;
    c->do_severity = P.DoSeverity;
    c->do_latent = P.DoLatent;
    c->do_infectiousness_profile = P.DoInfectiousnessProfile;
    c->do_real_sympt_withdrawal = P.DoRealSymptWithdrawal;
    c->absenteeism_place_closure = P.AbsenteeismPlaceClosure;
    c->no_infectiousness_sd_in_hh = P.NoInfectiousnessSDinHH;
    c->do_death = P.DoDeath;
    c->do_ad_units = P.DoAdUnits;
    c->output_ad_unit_age = P.OutputAdUnitAge;
    c->include_stepdown_to_death = P.IncludeStepDownToDeath;
    c->do_digital_contact_tracing = P.DoDigitalContactTracing;
    c->do_blanket_move_restr = P.DoBlanketMoveRestr;
    c->interventions_active = 0;
    for (int i = 0; i < MAX_NUM_PLACE_TYPES; i++) {
        c->place_type_trans[i] = P.PlaceTypeTrans[i];
        c->place_type_group_size_param1[i] = P.PlaceTypeGroupSizeParam1[i];
        c->place_type_prop_between_group_links[i] = P.PlaceTypePropBetweenGroupLinks[i];
        c->sympt_place_type_contact_rate[i] = P.SymptPlaceTypeContactRate[i];
        c->sympt_place_type_withdrawal_prop[i] = P.SymptPlaceTypeWithdrawalProp[i];
        c->holiday_effect[i] = P.HolidayEffect[i];
    }
    c->sympt_spatial_contact_rate = P.SymptSpatialContactRate;
    for (int a = 0; a < NUM_AGE_GROUPS; a++) {
        c->relative_spatial_contact[a] = P.RelativeSpatialContact[a];
        c->relative_spatial_contact_susc[a] = P.RelativeSpatialContactSusc[a];
        c->age_susceptibility[a] = P.AgeSusceptibility[a];
        c->age_infectiousness[a] = P.AgeInfectiousness[a];
        c->proportion_symptomatic[a] = P.ProportionSymptomatic[a];
        for (int b = 0; b < NUM_AGE_GROUPS; b++) {
            c->waifw[a][b] = P.WAIFW_Matrix[a][b];
            c->waifw_spatial[a][b] = P.WAIFW_Matrix_SpatialOnly[a][b];
        }
        c->mean_by_age[CS_MEAN_MILD_RECOV][a] = P.Mean_MildToRecovery[a];
        c->mean_by_age[CS_MEAN_ILI_RECOV][a] = P.Mean_ILIToRecovery[a];
        c->mean_by_age[CS_MEAN_SARI_RECOV][a] = P.Mean_SARIToRecovery[a];
        c->mean_by_age[CS_MEAN_CRIT_CRITRECOV][a] = P.Mean_CriticalToCritRecov[a];
        c->mean_by_age[CS_MEAN_CRITRECOV_RECOV][a] = P.Mean_CritRecovToRecov[a];
        c->mean_by_age[CS_MEAN_STEPDOWN_DEATH][a] = P.Mean_StepdownToDeath[a];
        c->mean_by_age[CS_MEAN_ILI_SARI][a] = P.Mean_ILIToSARI[a];
        c->mean_by_age[CS_MEAN_SARI_CRIT][a] = P.Mean_SARIToCritical[a];
        c->mean_by_age[CS_MEAN_CRIT_DEATH][a] = P.Mean_CriticalToDeath[a];
        c->mean_by_age[CS_MEAN_SARI_DEATH][a] = P.Mean_SARIToDeath[a];
        c->mean_by_age[CS_MEAN_ILI_DEATH][a] = P.Mean_ILIToDeath[a];
        c->prop_ili[a] = P.Prop_ILI_ByAge[a];
        c->prop_sari[a] = P.Prop_SARI_ByAge[a];
        c->prop_critical[a] = P.Prop_Critical_ByAge[a];
        c->cfr_critical[a] = P.CFR_Critical_ByAge[a];
        c->cfr_sari[a] = P.CFR_SARI_ByAge[a];
        c->cfr_ili[a] = P.CFR_ILI_ByAge[a];
    }
    for (int i = 0; i < P.NumInterventionClasses && i < CS_NUM_NPI; i++)
        for (int j = 0; j < P.NumInfectionSettings && j < CS_NUM_SETTINGS; j++)
            c->efficacies[i][j] = P.Efficacies[i][j];
    c->care_home_resident_household_scaling = P.CareHomeResidentHouseholdScaling;
    c->care_home_resident_spatial_scaling = P.CareHomeResidentSpatialScaling;
    c->care_home_worker_group_scaling = P.CareHomeWorkerGroupScaling;
    c->care_home_resident_place_scaling = P.CareHomeResidentPlaceScaling;
    c->care_home_rel_prob_hosp = P.CareHomeRelProbHosp;
    c->treat_inf_drop = P.TreatInfDrop;
    c->vacc_inf_drop = P.VaccInfDrop;
    c->treat_susc_drop = P.TreatSuscDrop;
    c->vacc_susc_drop = P.VaccSuscDrop;
    c->vacc_susc_drop2 = P.VaccSuscDrop2;
    c->treat_sympt_drop = P.TreatSympDrop;
    c->vacc_sympt_drop = P.VaccSympDrop;
    c->treat_death_drop = P.TreatDeathDrop;
    c->us_vacc_time_to_efficacy = P.usVaccTimeToEfficacy;
    c->us_vacc_time_efficacy_switch = P.usVaccTimeEfficacySwitch;
    c->us_hquarantine_house_duration = P.usHQuarantineHouseDuration;
    c->us_case_isolation_delay = P.usCaseIsolationDelay;
    c->us_case_isolation_duration = P.usCaseIsolationDuration;
    c->us_case_absenteeism_delay = P.usCaseAbsenteeismDelay;
    c->us_case_absenteeism_duration = P.usCaseAbsenteeismDuration;
    c->n_infectiousness = MAX_INFECTIOUS_STEPS;
    for (int i = 0; i < MAX_INFECTIOUS_STEPS; i++) c->infectiousness[i] = P.infectiousness[i];
    c->latent_period = P.LatentPeriod;
    c->infectious_period = P.InfectiousPeriod;
    c->latent_to_sympt_delay = P.LatentToSymptDelay;
    c->sympt_infectiousness = P.SymptInfectiousness;
    c->asympt_infectiousness = P.AsymptInfectiousness;
    c->infectiousness_sd = P.InfectiousnessSD;
    c->susceptibility_sd = P.SusceptibilitySD;
    copy_icdf(c->icdf[CS_ICDF_LATENT], P.latent_icdf);
    copy_icdf(c->icdf[CS_ICDF_INFECTIOUS], P.infectious_icdf);
    copy_icdf(c->icdf[CS_ICDF_MILD_RECOV], P.MildToRecovery_icdf);
    copy_icdf(c->icdf[CS_ICDF_ILI_RECOV], P.ILIToRecovery_icdf);
    copy_icdf(c->icdf[CS_ICDF_SARI_RECOV], P.SARIToRecovery_icdf);
    copy_icdf(c->icdf[CS_ICDF_CRIT_CRITRECOV], P.CriticalToCritRecov_icdf);
    copy_icdf(c->icdf[CS_ICDF_CRITRECOV_RECOV], P.CritRecovToRecov_icdf);
    copy_icdf(c->icdf[CS_ICDF_ILI_SARI], P.ILIToSARI_icdf);
    copy_icdf(c->icdf[CS_ICDF_SARI_CRIT], P.SARIToCritical_icdf);
    copy_icdf(c->icdf[CS_ICDF_ILI_DEATH], P.ILIToDeath_icdf);
    copy_icdf(c->icdf[CS_ICDF_SARI_DEATH], P.SARIToDeath_icdf);
    copy_icdf(c->icdf[CS_ICDF_STEPDOWN_DEATH], P.StepdownToDeath_icdf);
    copy_icdf(c->icdf[CS_ICDF_CRIT_DEATH], P.CriticalToDeath_icdf);
    c->cfr_critical_scale = P.CFR_Critical_Scale_Current;
    c->cfr_sari_scale = P.CFR_SARI_Scale_Current;
    c->cfr_ili_scale = P.CFR_ILI_Scale_Current;
    c->control_prop_cases_id = P.ControlPropCasesId;
    c->case_absent_child_age_cutoff = P.CaseAbsentChildAgeCutoff;
    c->place_close_round_household = P.PlaceCloseRoundHousehold;
    c->case_absent_child_prop_adult_carers = P.CaseAbsentChildPropAdultCarers;
    c->location_initial_infection[0] = P.LocationInitialInfection[0][0];
    c->location_initial_infection[1] = P.LocationInitialInfection[0][1];
    c->kernel_size = P.KernelLookup.size_;
    c->kernel_expansion = P.KernelLookup.expansion_factor_;
    c->kernel_delta = P.KernelLookup.delta_;
    c->kernel_type = P.MoveKernel.type_;
    c->kernel_shape = P.MoveKernel.shape_;
    c->kernel_scale = P.MoveKernel.scale_;
    c->kernel_p3 = P.MoveKernel.p3_;
    c->kernel_p4 = P.MoveKernel.p4_;
    c->do_utm_coords = P.DoUTM_coords;
    c->do_periodic_boundaries = P.DoPeriodicBoundaries;
    c->nch = P.nch;
    c->cell_width = P.in_cells_.width;
    c->cell_height = P.in_cells_.height;
    c->degrees_width = P.in_degrees_.width;
    c->degrees_height = P.in_degrees_.height;
    c->bbox_bottom_left_y = P.SpatialBoundingBox.bottom_left().y;
    for (int i = 0; i <= 360; i++) { c->sinx[i] = P.sinx[i]; c->cosx[i] = P.cosx[i]; }
    for (int i = 0; i <= 1000; i++) c->asin2sqx[i] = P.asin2sqx[i];
    c->do_holidays = P.DoHolidays;
    c->num_holidays = std::min(P.NumHolidays, CS_MAX_HOLIDAYS);
    for (int i = 0; i < c->num_holidays; i++) {
        c->holiday_start[i] = P.HolidayStartTime[i];
        c->holiday_duration[i] = P.HolidayDuration[i];
    }
    c->holidays_start_day_simtime = P.HolidaysStartDay_SimTime;
    c->r0 = P.R0;
    c->fix_local_beta = P.FixLocalBeta;
    c->do_si = P.DoSI;
    c->do_partial_immunity = P.DoPartialImmunity;
    c->got_waifw_spatial = P.Got_WAIFW_Matrix_Spatial;
    for (int a = 0; a < NUM_AGE_GROUPS; a++) c->initial_immunity[a] = P.InitialImmunity[a];
    c->care_home_prop_residents = P.CareHomePropResidents;
    // the intervention subset and the switches of everything outside it
    c->do_global_triggers = P.DoGlobalTriggers;
    c->do_icu_triggers = P.DoICUTriggers;
    c->do_admin_triggers = P.DoAdminTriggers;
    c->do_per_capita_triggers = P.DoPerCapitaTriggers;
    c->triggers_sampling_interval = P.TriggersSamplingInterval;
    c->trigger_alert_on_deaths = P.TriggerAlertOnDeaths;
    c->case_or_death_threshold_before_alert = P.CaseOrDeathThresholdBeforeAlert;
    c->window_to_evaluate_trigger_alert = P.WindowToEvaluateTriggerAlert;
    c->do_no_calibration = P.DoNoCalibration;
    c->do_alert_trigger_after_interv = P.DoAlertTriggerAfterInterv;
    c->date_trigger_reached_cal_time = P.DateTriggerReached_CalTime;
    c->pre_alert_control_prop_cases_id = P.PreAlertControlPropCasesId;
    c->post_alert_control_prop_cases_id = P.PostAlertControlPropCasesId;
    c->case_isolation_time_start_base = P.CaseIsolationTimeStartBase;
    c->case_isolation_prop = P.CaseIsolationProp;
    c->case_isolation_policy_duration = P.CaseIsolationPolicyDuration;
    c->case_isolation_cell_inc_thresh = P.CaseIsolation_CellIncThresh;
    c->hquarantine_time_start_base = P.HQuarantineTimeStartBase;
    c->hquarantine_delay = P.HQuarantineDelay;
    c->hquarantine_policy_duration = P.HQuarantinePolicyDuration;
    c->hquarantine_prop_house_compliant = P.HQuarantinePropHouseCompliant;
    c->hquarantine_prop_indiv_compliant = P.HQuarantinePropIndivCompliant;
    c->hhquar_cell_inc_thresh = P.HHQuar_CellIncThresh;
    c->do_hq_retrigger = P.DoHQretrigger;
    c->do_place_close_once_only = P.DoPlaceCloseOnceOnly;
    c->place_close_time_start_base = P.PlaceCloseTimeStartBase;
    c->place_close_delay_mean = P.PlaceCloseDelayMean;
    c->place_close_duration = P.PlaceCloseDurationBase;
    c->place_close_cell_inc_thresh = P.PlaceCloseCellIncThresh1;
    c->place_close_cell_inc_stop_thresh = P.PlaceCloseCellIncStopThresh;
    c->place_close_inc_trig = P.PlaceCloseIncTrig1;
    c->place_close_by_admin_unit = P.PlaceCloseByAdminUnit;
    c->place_close_frac_inc_trig = P.PlaceCloseFracIncTrig;
    for (int i = 0; i < MAX_NUM_PLACE_TYPES; i++) c->place_close_prop_attending[i] = P.PlaceClosePropAttending[i];
    c->treat_time_start_base = P.TreatTimeStartBase;
    c->vacc_time_start_base = P.VaccTimeStartBase;
    c->soc_dist_time_start_base = P.SocDistTimeStartBase;
    c->move_restr_time_start_base = P.MoveRestrTimeStartBase;
    c->key_worker_proph_time_start_base = P.KeyWorkerProphTimeStartBase;
    c->airport_close_time_start_base = P.AirportCloseTimeStartBase;
    c->place_close_time_start_base2 = P.PlaceCloseTimeStartBase2;
    c->place_close_indep_thresh = P.PlaceCloseIndepThresh;
    c->vacc_time_start_geo = P.VaccTimeStartGeo;
    c->do_intervention_delays_by_adunit = P.DoInterventionDelaysByAdUnit;
    c->vary_efficacies_over_time = P.VaryEfficaciesOverTime;
    c->do_mass_vacc = P.DoMassVacc;
    // any intervention configured to start at all: the library then serves
    // the subset above or refuses the run
    const double never = 10000.0;
    c->interventions_active = (P.CaseIsolationTimeStartBase < never || P.HQuarantineTimeStartBase < never ||
                               P.PlaceCloseTimeStartBase < never || P.TreatTimeStartBase < never ||
                               P.VaccTimeStartBase < never || P.SocDistTimeStartBase < never ||
                               P.MoveRestrTimeStartBase < never || P.KeyWorkerProphTimeStartBase < never ||
                               P.AirportCloseTimeStartBase < never || P.VaccTimeStartGeo < never || P.DoMassVacc) ? 1 : 0;
}

// InfStat code of a person, read through the public predicates
// (src/Person.cpp; codes of src/InfStat.h:7-38).
inline int infstat_code(const Person& a)
{
    if (a.is_susceptible()) return CS_SUSCEPTIBLE;
    if (a.is_latent()) return CS_LATENT;
    if (a.is_infectious_almost_symptomatic()) return CS_INF_ALMOST_SYMPT;
    if (a.is_infectious_asymptomatic_not_case()) return CS_INF_ASYMPT;
    if (a.is_case()) return CS_CASE;
    if (a.is_immune_at_start()) return CS_IMMUNE_AT_START;
    if (a.is_dead_was_asymp()) return CS_DEAD_ASYMP;
    if (a.is_dead_was_symp()) return CS_DEAD_SYMP;
    if (a.is_recovered_symp()) return CS_RECOVERED_SYMP;
    return CS_RECOVERED_ASYMP;
}

// Set a person's InfStat through the public setters; recovered and dead
// states derive their sign from the state before (src/Person.cpp:111-154).
inline void set_infstat(Person& a, int code)
{
    switch (code) {
    case CS_SUSCEPTIBLE: a.set_susceptible(); break;
    case CS_LATENT: a.set_latent(); break;
    case CS_INF_ALMOST_SYMPT: a.set_infectious_almost_symptomatic(); break;
    case CS_INF_ASYMPT: a.set_infectious_asymptomatic_not_case(); break;
    case CS_CASE: a.set_case(); break;
    case CS_IMMUNE_AT_START: a.set_immune_at_start(); break;
    case CS_RECOVERED_SYMP: a.set_case(); a.set_recovered(); break;
    case CS_RECOVERED_ASYMP: a.set_infectious_asymptomatic_not_case(); a.set_recovered(); break;
    case CS_DEAD_SYMP: a.set_case(); a.set_dead(); break;
    case CS_DEAD_ASYMP: a.set_infectious_asymptomatic_not_case(); a.set_dead(); break;
    default: break;
    }
}

// cs_popgen_params from Param P as ReadParams leaves it (the generator's
// inputs: src/SetupModel.cpp:848-2601 read these fields).  ReadParams stores
// "Prob Youngest Child Under Five" into TwoChildrenProbYoungestUnderFive
// and never sets ProbYoungestChildUnderFive (src/ReadParams.cpp:2289-2290);
// the fields are copied as the run has them.
inline void fill_popgen_params(cs_popgen_params* g)
{
    memset(g, 0, sizeof(*g));
    g->pop_size = P.PopSize;
    g->mcw = P.total_microcells_wide_;
    g->mch = P.total_microcells_high_;
    g->nmcl = P.NMCL;
    g->ncw = P.ncw;
    g->nch = P.nch;
    g->mc_width = P.in_microcells_.width;
    g->mc_height = P.in_microcells_.height;
    g->bbox_x0 = P.SpatialBoundingBox.bottom_left().x;
    g->bbox_y0 = P.SpatialBoundingBox.bottom_left().y;
    g->longitude_cut_line = P.LongitudeCutLine;
    g->do_ad_units = P.DoAdUnits;
    g->do_adunit_boundaries = P.DoAdunitBoundaries;
    g->adunit_level1_mask = P.AdunitLevel1Mask;
    g->adunit_level1_divisor = P.AdunitLevel1Divisor;
    g->do_households = P.DoHouseholds;
    g->do_correct_age_dist = P.DoCorrectAgeDist;
    for (int i = 0; i < MAX_HOUSEHOLD_SIZE; i++) g->household_size_cdf[i] = P.HouseholdSizeDistrib[0][i];
    for (int i = 0; i < NUM_AGE_GROUPS; i++) g->prop_age_group[i] = P.PropAgeGroup[0][i];
    g->one_pers_house_prob_old = P.OnePersHouseProbOld;
    g->one_pers_house_prob_young = P.OnePersHouseProbYoung;
    g->two_pers_house_prob_old = P.TwoPersHouseProbOld;
    g->two_pers_house_prob_young = P.TwoPersHouseProbYoung;
    g->one_child_two_pers_prob = P.OneChildTwoPersProb;
    g->zero_child_three_pers_prob = P.ZeroChildThreePersProb;
    g->two_child_three_pers_prob = P.TwoChildThreePersProb;
    g->one_child_four_pers_prob = P.OneChildFourPersProb;
    g->three_child_five_pers_prob = P.ThreeChildFivePersProb;
    g->one_child_prob_youngest_under_five = P.OneChildProbYoungestChildUnderFive;
    g->two_children_prob_youngest_under_five = P.TwoChildrenProbYoungestUnderFive;
    g->prob_youngest_child_under_five = P.ProbYoungestChildUnderFive;
    g->young_and_single_slope = P.YoungAndSingleSlope;
    g->no_child_pers_age = P.NoChildPersAge;
    g->old_pers_age = P.OldPersAge;
    g->young_and_single = P.YoungAndSingle;
    g->min_adult_age = P.MinAdultAge;
    g->max_mf_partner_age_gap = P.MaxMFPartnerAgeGap;
    g->max_fm_partner_age_gap = P.MaxFMPartnerAgeGap;
    g->min_parent_age_gap = P.MinParentAgeGap;
    g->max_parent_age_gap = P.MaxParentAgeGap;
    g->max_child_age = P.MaxChildAge;
    g->mean_child_age_gap = P.MeanChildAgeGap;
    g->older_gen_gap = P.OlderGenGap;
    g->num_place_types = P.NumPlaceTypes;
    g->hotel_place_type = P.HotelPlaceType;
    g->nsp = P.nsp;
    for (int t = 0; t < MAX_NUM_PLACE_TYPES; t++) {
        g->place_age_min[0][t] = P.PlaceTypeAgeMin[t];
        g->place_age_max[0][t] = P.PlaceTypeAgeMax[t];
        g->place_prop[0][t] = P.PlaceTypePropAgeGroup[t];
        g->place_age_min[1][t] = P.PlaceTypeAgeMin2[t];
        g->place_age_max[1][t] = P.PlaceTypeAgeMax2[t];
        g->place_prop[1][t] = P.PlaceTypePropAgeGroup2[t];
        g->place_age_min[2][t] = P.PlaceTypeAgeMin3[t];
        g->place_age_max[2][t] = P.PlaceTypeAgeMax3[t];
        g->place_prop[2][t] = P.PlaceTypePropAgeGroup3[t];
        g->place_mean_size[t] = P.PlaceTypeMeanSize[t];
        g->place_size_power[t] = P.PlaceTypeSizePower[t];
        g->place_size_offset[t] = P.PlaceTypeSizeOffset[t];
        g->place_size_sd[t] = P.PlaceTypeSizeSD[t];
        g->place_size_min[t] = P.PlaceTypeSizeMin[t];
        g->place_size_max[t] = P.PlaceTypeSizeMax[t];
        g->place_nearest[t] = P.PlaceTypeNearestNeighb[t];
        g->place_kernel_type[t] = P.PlaceTypeKernelType[t];
        g->place_kernel_scale[t] = P.PlaceTypeKernelScale[t];
        g->place_kernel_shape[t] = P.PlaceTypeKernelShape[t];
        g->place_kernel_p3[t] = P.PlaceTypeKernelP3[t];
        g->place_kernel_p4[t] = P.PlaceTypeKernelP4[t];
        g->place_group_size[t] = P.PlaceTypeGroupSizeParam1[t];
        g->inhibit_inter_adunit[t] = (P.DoAdUnits && P.InhibitInterAdunitPlaceAssignment[t] > 0) ? 1 : 0;
    }
    g->setup_seed1 = P.setupSeed1;
    g->setup_seed2 = P.setupSeed2;
}

inline int n_mcells_total() { return P.NMCL * P.NMCL * P.NumCells; }

// The static population as flat arrays (the layout cs_population describes).
struct PopulationView {
    std::vector<int32_t> hh, mcell, pcell, plinks, hfirst, cidx, cn, cstart, cmem, madunit;
    std::vector<uint8_t> age, flags;
    std::vector<uint16_t> pglinks, hnh, hnhr;
    std::vector<float> hloc;
    std::vector<int32_t> pl_n[CS_MAX_PLACE_TYPES], pl_mc[CS_MAX_PLACE_TYPES], pl_ms[CS_MAX_PLACE_TYPES],
        pl_mem[CS_MAX_PLACE_TYPES], pl_gs0[CS_MAX_PLACE_TYPES], pl_gst[CS_MAX_PLACE_TYPES], pl_gsz[CS_MAX_PLACE_TYPES];

    void gather()
    {
        const int N = P.PopSize, NH = P.NumHouseholds, NC = P.NumPopulatedCells, NM = n_mcells_total();
        hh.resize(N); mcell.resize(N); pcell.resize(N); age.resize(N); flags.resize(N);
        plinks.resize((size_t)N * CS_MAX_PLACE_TYPES);
        pglinks.resize((size_t)N * CS_MAX_PLACE_TYPES);
        for (int i = 0; i < N; i++) {
            const Person& a = Hosts[i];
            hh[i] = a.hh; mcell[i] = a.mcell; pcell[i] = a.pcell; age[i] = a.age;
            flags[i] = (uint8_t)((a.care_home_resident ? 1 : 0) | (a.keyworker ? 2 : 0) |
                                 (a.esocdist_comply ? 4 : 0) | (a.digitalContactTracingUser ? 8 : 0));
            for (int t = 0; t < CS_MAX_PLACE_TYPES; t++) {
                plinks[(size_t)i * CS_MAX_PLACE_TYPES + t] = (t < P.NumPlaceTypes) ? a.PlaceLinks[t] : -1;
                pglinks[(size_t)i * CS_MAX_PLACE_TYPES + t] = (t < P.NumPlaceTypes) ? a.PlaceGroupLinks[t] : 0;
            }
        }
        hfirst.resize(NH); hnh.resize(NH); hnhr.resize(NH); hloc.resize((size_t)NH * 2);
        for (int h = 0; h < NH; h++) {
            hfirst[h] = Households[h].FirstPerson; hnh[h] = Households[h].nh; hnhr[h] = Households[h].nhr;
            hloc[2 * h] = Households[h].loc.x; hloc[2 * h + 1] = Households[h].loc.y;
        }
        cidx.resize(NC); cn.resize(NC); cstart.assign(NC + 1, 0); cmem.clear();
        for (int j = 0; j < NC; j++) {
            Cell* c = CellLookup[j];
            cidx[j] = (int32_t)(c - Cells); cn[j] = c->n;
            for (int k = 0; k < c->n; k++) cmem.push_back(c->members[k]);
            cstart[j + 1] = cstart[j] + c->n;
        }
        madunit.resize(NM);
        for (int m = 0; m < NM; m++) madunit[m] = Mcells[m].adunit;
        for (int t = 0; t < P.NumPlaceTypes && t < CS_MAX_PLACE_TYPES; t++) {
            const int np = P.Nplace[t];
            pl_n[t].assign(np, 0); pl_mc[t].assign(np, 0);
            pl_ms[t].assign(np + 1, 0); pl_gs0[t].assign(np + 1, 0);
            pl_mem[t].clear(); pl_gst[t].clear(); pl_gsz[t].clear();
            for (int k = 0; k < np; k++) {
                const Place& pl = Places[t][k];
                pl_n[t][k] = pl.n; pl_mc[t][k] = pl.mcell;
                for (int m = 0; m < pl.n; m++) pl_mem[t].push_back(pl.members[m]);
                pl_ms[t][k + 1] = pl_ms[t][k] + pl.n;
                for (int g = 0; g < pl.ng; g++) { pl_gst[t].push_back(pl.group_start[g]); pl_gsz[t].push_back(pl.group_size[g]); }
                pl_gs0[t][k + 1] = pl_gs0[t][k] + pl.ng;
            }
        }
    }

    cs_population view() const
    {
        cs_population pop;
        memset(&pop, 0, sizeof(pop));
        pop.n_people = P.PopSize; pop.n_households = P.NumHouseholds; pop.n_cells = P.NumPopulatedCells;
        pop.n_cells_total = P.NumCells; pop.n_mcells = n_mcells_total(); pop.n_adunits = P.NumAdunits;
        pop.hh = hh.data(); pop.mcell = mcell.data(); pop.pcell = pcell.data(); pop.age = age.data();
        pop.flags = flags.data(); pop.place_links = plinks.data(); pop.place_group_links = pglinks.data();
        pop.hh_first = hfirst.data(); pop.hh_nh = hnh.data(); pop.hh_nhr = hnhr.data(); pop.hh_loc = hloc.data();
        pop.cell_index = cidx.data(); pop.cell_n = cn.data(); pop.cell_member_start = cstart.data();
        pop.cell_members = cmem.data(); pop.mcell_adunit = madunit.data();
        for (int t = 0; t < P.NumPlaceTypes && t < CS_MAX_PLACE_TYPES; t++) {
            pop.n_places[t] = P.Nplace[t];
            pop.place_n[t] = pl_n[t].data(); pop.place_mcell[t] = pl_mc[t].data();
            pop.place_member_start[t] = pl_ms[t].data(); pop.place_members[t] = pl_mem[t].data();
            pop.place_groups_start[t] = pl_gs0[t].data(); pop.place_group_start[t] = pl_gst[t].data();
            pop.place_group_size[t] = pl_gsz[t].data();
        }
        return pop;
    }
};

// The dynamic state of one realisation as flat arrays (the layout cs_state
// describes), with copies from and back to the reference's globals.
struct StateView {
    std::vector<int8_t> inf;
    std::vector<uint8_t> sevc, sevf, bits, qc, dct, msd, mmr;
    std::vector<int32_t> listpos, infector, cS, cL, cI, cR, cD, cTC, perm;
    std::vector<int16_t> itype;
    std::vector<float> infn, susc, pcare;
    std::vector<uint16_t> t_inf, t_lat, t_rod, t_sari, t_crit, t_step, t_det, t_as, t_ae, t_ts, t_te, t_vs, t_is, t_qs;
    std::vector<uint16_t> pcs[CS_MAX_PLACE_TYPES], pce[CS_MAX_PLACE_TYPES];
    std::vector<int32_t> cstart;  // cell offsets into perm

    void resize()
    {
        const size_t N = P.PopSize, NC = P.NumPopulatedCells, NM = n_mcells_total();
        inf.resize(N); sevc.resize(N); sevf.resize(N); bits.resize(N); qc.resize(N); dct.resize(N);
        listpos.resize(N); infector.resize(N); itype.resize(N); infn.resize(N); susc.resize(N); pcare.resize(N);
        for (auto* v : {&t_inf, &t_lat, &t_rod, &t_sari, &t_crit, &t_step, &t_det, &t_as, &t_ae,
                        &t_ts, &t_te, &t_vs, &t_is, &t_qs}) v->resize(N);
        for (auto* v : {&cS, &cL, &cI, &cR, &cD, &cTC}) v->resize(NC);
        perm.resize(N); msd.resize(NM); mmr.resize(NM);
        cstart.assign(NC + 1, 0);
        for (size_t j = 0; j < NC; j++) cstart[j + 1] = cstart[j] + CellLookup[j]->n;
        for (int t = 0; t < P.NumPlaceTypes && t < CS_MAX_PLACE_TYPES; t++) {
            pcs[t].resize(std::max(1, P.Nplace[t])); pce[t].resize(std::max(1, P.Nplace[t]));
        }
    }

    // reference globals -> arrays
    void gather()
    {
        resize();
        const int N = P.PopSize, NC = P.NumPopulatedCells, NM = n_mcells_total();
        for (int i = 0; i < N; i++) {
            const Person& a = Hosts[i];
            inf[i] = (int8_t)infstat_code(a);
            sevc[i] = (uint8_t)a.Severity_Current; sevf[i] = (uint8_t)a.Severity_Final;
            bits[i] = (uint8_t)((a.to_die ? 1 : 0) | (a.detected ? 2 : 0));
            listpos[i] = a.listpos; infector[i] = a.infector; itype[i] = a.infect_type;
            infn[i] = a.infectiousness; susc[i] = a.susc; pcare[i] = a.ProbCare;
            t_inf[i] = a.infection_time; t_lat[i] = a.latent_time; t_rod[i] = a.recovery_or_death_time;
            t_sari[i] = a.SARI_time; t_crit[i] = a.Critical_time; t_step[i] = a.Stepdown_time;
            t_det[i] = a.detected_time; t_as[i] = a.absent_start_time; t_ae[i] = a.absent_stop_time;
            t_ts[i] = a.treat_start_time; t_te[i] = a.treat_stop_time; t_vs[i] = a.vacc_start_time;
            t_is[i] = a.isolation_start_time;
            t_qs[i] = HostsQuarantine[i].start_time; qc[i] = HostsQuarantine[i].comply;
            dct[i] = a.digitalContactTraced;
        }
        for (int j = 0; j < NC; j++) {
            const Cell* c = CellLookup[j];
            cS[j] = c->S; cL[j] = c->L; cI[j] = c->I; cR[j] = c->R; cD[j] = c->D; cTC[j] = c->cumTC;
            memcpy(&perm[cstart[j]], c->susceptible, sizeof(int) * (size_t)c->n);
        }
        for (int t = 0; t < P.NumPlaceTypes && t < CS_MAX_PLACE_TYPES; t++)
            for (int k = 0; k < P.Nplace[t]; k++) {
                pcs[t][k] = Places[t][k].close_start_time;
                pce[t][k] = Places[t][k].close_end_time;
            }
        for (int m = 0; m < NM; m++) { msd[m] = (uint8_t)Mcells[m].socdist; mmr[m] = (uint8_t)Mcells[m].moverest; }
    }

    // arrays -> reference globals: the fields the sweeps write
    void scatter() const
    {
        const int N = P.PopSize, NC = P.NumPopulatedCells;
        for (int i = 0; i < N; i++) {
            Person& a = Hosts[i];
            if (infstat_code(a) != inf[i]) set_infstat(a, inf[i]);
            a.Severity_Current = (Severity)sevc[i]; a.Severity_Final = (Severity)sevf[i];
            a.to_die = bits[i] & 1; a.detected = (bits[i] >> 1) & 1;
            a.listpos = listpos[i]; a.infector = infector[i]; a.infect_type = itype[i];
            a.infectiousness = infn[i];
            a.infection_time = t_inf[i]; a.latent_time = t_lat[i]; a.recovery_or_death_time = t_rod[i];
            a.SARI_time = t_sari[i]; a.Critical_time = t_crit[i]; a.Stepdown_time = t_step[i];
            a.detected_time = t_det[i]; a.absent_start_time = t_as[i]; a.absent_stop_time = t_ae[i];
        }
        for (int j = 0; j < NC; j++) {
            Cell* c = CellLookup[j];
            c->S = cS[j]; c->L = cL[j]; c->I = cI[j]; c->R = cR[j]; c->D = cD[j]; c->cumTC = cTC[j];
            memcpy(c->susceptible, &perm[cstart[j]], sizeof(int) * (size_t)c->n);
            c->latent = c->susceptible + c->S;
            c->infected = c->latent + c->L;
        }
        for (int t = 0; t < P.NumPlaceTypes && t < CS_MAX_PLACE_TYPES; t++)
            for (int k = 0; k < P.Nplace[t]; k++) {
                Places[t][k].close_start_time = pcs[t][k];
                Places[t][k].close_end_time = pce[t][k];
            }
    }

    cs_state view()
    {
        cs_state st;
        memset(&st, 0, sizeof(st));
        st.inf = inf.data(); st.severity_current = sevc.data(); st.severity_final = sevf.data(); st.bits = bits.data();
        st.listpos = listpos.data(); st.infector = infector.data(); st.infect_type = itype.data();
        st.infectiousness = infn.data(); st.susc = susc.data(); st.prob_care = pcare.data();
        st.infection_time = t_inf.data(); st.latent_time = t_lat.data(); st.recovery_or_death_time = t_rod.data();
        st.sari_time = t_sari.data(); st.critical_time = t_crit.data(); st.stepdown_time = t_step.data();
        st.detected_time = t_det.data(); st.absent_start_time = t_as.data(); st.absent_stop_time = t_ae.data();
        st.treat_start_time = t_ts.data(); st.treat_stop_time = t_te.data(); st.vacc_start_time = t_vs.data();
        st.isolation_start_time = t_is.data(); st.quar_start_time = t_qs.data(); st.quar_comply = qc.data();
        st.dct_traced = dct.data();
        st.cell_S = cS.data(); st.cell_L = cL.data(); st.cell_I = cI.data(); st.cell_R = cR.data();
        st.cell_D = cD.data(); st.cell_cumTC = cTC.data(); st.cell_perm = perm.data();
        for (int t = 0; t < P.NumPlaceTypes && t < CS_MAX_PLACE_TYPES; t++) {
            st.place_close_start[t] = pcs[t].data();
            st.place_close_end[t] = pce[t].data();
        }
        st.mcell_socdist = msd.data(); st.mcell_moverest = mmr.data();
        return st;
    }
};

}  // namespace cs_bridge
