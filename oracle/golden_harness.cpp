// golden_harness.cpp -- test infrastructure, not product code.
//
// Links the unmodified CovidSim objects built by oracle/Makefile.ref with
// Sweep.cpp compiled as InfectSweep_cpu/IncubRecoverySweep_cpu, and supplies
// the two sweep symbols itself (src/Sweep.h:4-9).  Each call forwards to the
// reference sweep unchanged, so a harness run reproduces covidsim_ref's run
// exactly; around the calls it writes what the parity tests need:
//
//   CS_DUMP_DIR=<dir>   required to dump anything
//   CS_DUMP_STEPS=a,b   TimeStepNow values at which the state of realisation
//                       0 is written before InfectSweep (default "0")
//   CS_DUMP_GOLDEN=1    at the first call also write golden vectors: the
//                       kernel lookup num() (src/Kernels.cpp:108-129), person
//                       dist2 (src/Dist.cpp:49-60), InverseCdf::choose
//                       (src/InverseCdf.cpp:60-74) with its uniform, and the
//                       eight hazards of src/CalcInfSusc.cpp over a grid of
//                       synthetic person states
//   CS_DUMP_ITC=1       at the first call, after the dumps above, re-run the
//                       reference's InitTransmissionCoeffs
//                       (src/SetupModel.cpp:640-846) from a fixed RNG seed,
//                       replay its per-person draws with the same generator
//                       to recover them (one thread: the draws come in person
//                       order), write them with the resulting coefficients
//                       (itc_*.npy) and exit; the run does not continue
//   CS_SERIES=<file>    append one record per step and realisation:
//                       run, TimeStepNow, S, L, I, R, D, cumI, cumC, cumD,
//                       cumItype[16], then the start times (in timesteps,
//                       -1 while unset) of place closure (P.PlaceCloseTimeStart),
//                       case isolation and household quarantine (of admin
//                       unit 0: with global triggers every unit has the same)
//                       and State.trigDetectedCases (int32 each, 30 in all),
//                       read after IncubRecoverySweep
//
// Output arrays are .npy files; cs_params is written as its raw bytes
// (params.bin) with the layout of include/covidsim_hip.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include <functional>
#include <limits>
#include <climits>
#include <iostream>
#include <fstream>
#include <sstream>

// KernelLookup keeps its tables private (src/Kernels.h:79-87); the harness
// reads them in place, so the class is opened up for this one TU.
#define private public
#include "Kernels.h"
#undef private

#include "CalcInfSusc.h"
#include "Constants.h"
#include "Dist.h"
#include "InverseCdf.h"
#include "Model.h"
#include "ModelMacros.h"
#include "Param.h"
#include "Rand.h"
#include "SetupModel.h"
#include "Sweep.h"

#include "covidsim_hip.h"
#include "cs_reference_bridge.h"

void InfectSweep_cpu(double t, int run);
void IncubRecoverySweep_cpu(double t);

namespace {

std::string g_dir;
std::vector<int> g_steps;
bool g_golden = false;
bool g_started = false;
FILE* g_series = nullptr;
int g_last_run = -1;

void save_npy(const std::string& name, const char* descr, const std::vector<size_t>& shape,
              const void* data, size_t elem)
{
    std::string path = g_dir + "/" + name + ".npy";
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { perror(path.c_str()); exit(2); }
    std::string sh = "(";
    size_t n = 1;
    for (size_t i = 0; i < shape.size(); i++) {
        sh += std::to_string(shape[i]);
        sh += (shape.size() == 1) ? "," : (i + 1 < shape.size() ? ", " : "");
        n *= shape[i];
    }
    sh += ")";
    std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
    size_t total = 10 + hdr.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    hdr += std::string(pad, ' ') + "\n";
    unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    fwrite(magic, 1, 8, f);
    uint16_t hl = (uint16_t)hdr.size();
    fwrite(&hl, 2, 1, f);
    fwrite(hdr.data(), 1, hdr.size(), f);
    if (n) fwrite(data, elem, n, f);
    fclose(f);
}
template <class T> const char* npdt();
template <> const char* npdt<int8_t>() { return "|i1"; }
template <> const char* npdt<uint8_t>() { return "|u1"; }
template <> const char* npdt<int16_t>() { return "<i2"; }
template <> const char* npdt<uint16_t>() { return "<u2"; }
template <> const char* npdt<int32_t>() { return "<i4"; }
template <> const char* npdt<uint32_t>() { return "<u4"; }
template <> const char* npdt<float>() { return "<f4"; }
template <> const char* npdt<double>() { return "<f8"; }
template <class T> void save(const std::string& name, const std::vector<T>& v)
{
    save_npy(name, npdt<T>(), {v.size()}, v.data(), sizeof(T));
}
template <class T> void save2(const std::string& name, const std::vector<T>& v, size_t rows, size_t cols)
{
    save_npy(name, npdt<T>(), {rows, cols}, v.data(), sizeof(T));
}

using cs_bridge::fill_params;

void dump_params(const std::string& name)
{
    cs_params c;
    fill_params(&c);
    std::string path = g_dir + "/" + name;
    FILE* f = fopen(path.c_str(), "wb");
    fwrite(&c, sizeof(c), 1, f);
    fclose(f);
}

void dump_population()
{
    cs_bridge::PopulationView v;
    v.gather();
    const int N = P.PopSize;
    save("p_hh", v.hh); save("p_mcell", v.mcell); save("p_pcell", v.pcell);
    save("p_age", v.age); save("p_flags", v.flags);
    save2("p_place_links", v.plinks, N, MAX_NUM_PLACE_TYPES);
    save2("p_place_group_links", v.pglinks, N, MAX_NUM_PLACE_TYPES);
    save("h_first", v.hfirst); save("h_nh", v.hnh); save("h_nhr", v.hnhr); save2("h_loc", v.hloc, P.NumHouseholds, 2);
    save("c_index", v.cidx); save("c_n", v.cn); save("c_member_start", v.cstart); save("c_members", v.cmem);
    save("m_adunit", v.madunit);
    for (int t = 0; t < P.NumPlaceTypes; t++) {
        std::string s = std::to_string(t);
        save("pl" + s + "_n", v.pl_n[t]); save("pl" + s + "_mcell", v.pl_mc[t]); save("pl" + s + "_member_start", v.pl_ms[t]);
        save("pl" + s + "_members", v.pl_mem[t]); save("pl" + s + "_groups_start", v.pl_gs0[t]);
        save("pl" + s + "_group_start", v.pl_gst[t]); save("pl" + s + "_group_size", v.pl_gsz[t]);
    }
    const int NC = P.NumPopulatedCells, nm = cs_bridge::n_mcells_total();
    std::vector<int32_t> meta = {P.PopSize, P.NumHouseholds, NC, P.NumCells, nm, P.NumAdunits, P.NumPlaceTypes, P.NMCL};
    save("meta", meta);
    // the set-up's inputs and outputs the population generator is compared on:
    // its parameters, place locations, microcell populations and countries
    {
        cs_popgen_params g;
        cs_bridge::fill_popgen_params(&g);
        std::string path = g_dir + "/popgen_params.bin";
        FILE* f = fopen(path.c_str(), "wb");
        fwrite(&g, sizeof(g), 1, f);
        fclose(f);
        for (int t = 0; t < P.NumPlaceTypes; t++) {
            std::vector<float> loc((size_t)P.Nplace[t] * 2);
            for (int i = 0; i < P.Nplace[t]; i++) { loc[2 * i] = Places[t][i].loc.x; loc[2 * i + 1] = Places[t][i].loc.y; }
            save2("pl" + std::to_string(t) + "_loc", loc, P.Nplace[t], 2);
        }
        std::vector<int32_t> mn(nm), mcn(nm);
        for (int i = 0; i < nm; i++) { mn[i] = Mcells[i].n; mcn[i] = mcell_country[i]; }
        save("m_n", mn); save("m_country", mcn);
    }

    // kernel tables
    std::vector<double> lk(P.KernelLookup.lookup_), hr(P.KernelLookup.hi_res_);
    save("kernel_lookup", lk); save("kernel_hi_res", hr);
}

void dump_state(const std::string& prefix, unsigned short ts)
{
    cs_bridge::StateView v;
    v.gather();
    const std::string p = prefix;
    save(p + "inf", v.inf); save(p + "sev_cur", v.sevc); save(p + "sev_final", v.sevf); save(p + "bits", v.bits);
    save(p + "listpos", v.listpos); save(p + "infector", v.infector); save(p + "infect_type", v.itype);
    save(p + "infectiousness", v.infn); save(p + "susc", v.susc); save(p + "prob_care", v.pcare);
    save(p + "t_infection", v.t_inf); save(p + "t_latent", v.t_lat); save(p + "t_recovery", v.t_rod);
    save(p + "t_sari", v.t_sari); save(p + "t_critical", v.t_crit); save(p + "t_stepdown", v.t_step);
    save(p + "t_detected", v.t_det); save(p + "t_absent_start", v.t_as); save(p + "t_absent_stop", v.t_ae);
    save(p + "t_treat_start", v.t_ts); save(p + "t_treat_stop", v.t_te); save(p + "t_vacc_start", v.t_vs);
    save(p + "t_isolation_start", v.t_is); save(p + "t_quar_start", v.t_qs); save(p + "quar_comply", v.qc);
    save(p + "dct_traced", v.dct);

    const int NC = P.NumPopulatedCells;
    save(p + "cell_S", v.cS); save(p + "cell_L", v.cL); save(p + "cell_I", v.cI); save(p + "cell_R", v.cR);
    save(p + "cell_D", v.cD); save(p + "cell_cumTC", v.cTC); save(p + "cell_perm", v.perm);
    // the cell tables UpdateProbs keeps (not part of the state the library loads)
    std::vector<int32_t> S0(NC);
    std::vector<float> totp(NC), maxt((size_t)NC * NC), cumt((size_t)NC * NC);
    std::vector<int32_t> invcdf((size_t)NC * 1025);
    for (int j = 0; j < NC; j++) {
        Cell* c = CellLookup[j];
        S0[j] = c->S0;
        totp[j] = c->tot_prob;
        for (int m = 0; m < NC; m++) { maxt[(size_t)j * NC + m] = c->max_trans[m]; cumt[(size_t)j * NC + m] = c->cum_trans[m]; }
        for (int k = 0; k <= 1024; k++) invcdf[(size_t)j * 1025 + k] = c->InvCDF[k];
    }
    save(p + "cell_S0", S0);
    save(p + "tot_prob", totp); save2(p + "max_trans", maxt, NC, NC); save2(p + "cum_trans", cumt, NC, NC);
    save2(p + "invcdf", invcdf, NC, 1025);
    for (int t = 0; t < P.NumPlaceTypes; t++) {
        save(p + "pl" + std::to_string(t) + "_close_start", v.pcs[t]);
        save(p + "pl" + std::to_string(t) + "_close_end", v.pce[t]);
    }
    save(p + "m_socdist", v.msd); save(p + "m_moverest", v.mmr);
    std::vector<int32_t> tsv = {(int32_t)ts};
    save(p + "timestep", tsv);
    dump_params(p + "params.bin");
}

// 64-bit LCG so the golden sample choices do not touch the reference RNG
uint64_t g_lcg = 0x853c49e6748fea9bULL;
double lcg_u() { g_lcg = g_lcg * 6364136223846793005ULL + 1442695040888963407ULL; return (double)(g_lcg >> 11) * (1.0 / 9007199254740992.0); }
int lcg_i(int n) { return (int)(lcg_u() * n); }

void dump_golden()
{
    // kernel lookup num(r2): uniform over the table, the hi-res region, and the boundaries
    const CovidSim::TBD1::KernelLookup& K = P.KernelLookup;
    double rmax = K.delta_ * K.size_;
    std::vector<double> r2;
    for (int i = 0; i < 6000; i++) r2.push_back(lcg_u() * rmax);
    for (int i = 0; i < 3000; i++) r2.push_back(lcg_u() * rmax / K.expansion_factor_);
    r2.push_back(0.0); r2.push_back(rmax);
    double hb = K.delta_ * K.size_ / K.expansion_factor_;
    for (int d = -20; d <= 20; d++) r2.push_back(hb * (1.0 + d * 1e-13));
    for (int i = 0; i < 500; i++) r2.push_back(K.delta_ * (double)lcg_i(K.size_));
    std::vector<double> kv(r2.size());
    for (size_t i = 0; i < r2.size(); i++) kv[i] = K.num(r2[i]);
    save("g_kernel_r2", r2); save("g_kernel_num", kv);

    // dist2 between random person pairs (src/Dist.cpp:49-60)
    std::vector<int32_t> da, db;
    std::vector<double> dv;
    for (int i = 0; i < 4000; i++) {
        int a = lcg_i(P.PopSize), b = (i % 7 == 0) ? a : lcg_i(P.PopSize);
        da.push_back(a); db.push_back(b); dv.push_back(dist2(Hosts + a, Hosts + b));
    }
    save("g_dist_a", da); save("g_dist_b", db); save("g_dist2", dv);

    // InverseCdf::choose with its uniform captured from thread 0's stream
    InverseCdf* tabs[CS_NUM_ICDF] = {&P.latent_icdf, &P.infectious_icdf, &P.MildToRecovery_icdf, &P.ILIToRecovery_icdf,
        &P.SARIToRecovery_icdf, &P.CriticalToCritRecov_icdf, &P.CritRecovToRecov_icdf, &P.ILIToSARI_icdf,
        &P.SARIToCritical_icdf, &P.ILIToDeath_icdf, &P.SARIToDeath_icdf, &P.StepdownToDeath_icdf, &P.CriticalToDeath_icdf};
    std::vector<int32_t> ct;
    std::vector<double> cu, cm;
    std::vector<uint16_t> cvv;
    int32_t s1 = Xcg1[0], s2 = Xcg2[0];
    for (int tb = 0; tb < CS_NUM_ICDF; tb++)
        for (int i = 0; i < 400; i++) {
            double mean = 0.5 + 20.0 * lcg_u();
            int32_t a1 = Xcg1[0], a2 = Xcg2[0];
            double u = ranf_mt(0);
            Xcg1[0] = a1; Xcg2[0] = a2;
            unsigned short v = tabs[tb]->choose(mean, 0, P.TimeStepsPerDay);
            ct.push_back(tb); cu.push_back(u); cm.push_back(mean); cvv.push_back(v);
        }
    Xcg1[0] = s1; Xcg2[0] = s2;
    save("g_choose_table", ct); save("g_choose_u", cu); save("g_choose_mean", cm); save("g_choose_value", cvv);

    // hazards over synthetic states.  The harness perturbs the intervention
    // parameters so that every factor is exercised, gives distinct persons
    // synthetic states, evaluates the eight reference hazards, writes the
    // perturbed state (prefix hz_) and restores everything.
    double** effsave = P.Efficacies;
    static double effbuf[CS_NUM_NPI][CS_NUM_SETTINGS];
    static double* effrows[CS_NUM_NPI];
    for (int i = 0; i < CS_NUM_NPI; i++) {
        effrows[i] = effbuf[i];
        for (int j = 0; j < CS_NUM_SETTINGS; j++) effbuf[i][j] = 0.2 + 0.75 * lcg_u();
    }
    const double sv_tid = P.TreatInfDrop, sv_vid = P.VaccInfDrop, sv_tsd = P.TreatSuscDrop, sv_vsd = P.VaccSuscDrop,
        sv_vsd2 = P.VaccSuscDrop2, sv_chh = P.CareHomeResidentHouseholdScaling, sv_chs = P.CareHomeResidentSpatialScaling;
    const unsigned short sv_u1 = P.usVaccTimeToEfficacy, sv_u2 = P.usVaccTimeEfficacySwitch,
        sv_u3 = P.usHQuarantineHouseDuration, sv_u4 = P.usCaseIsolationDelay, sv_u5 = P.usCaseIsolationDuration;
    double rsc[NUM_AGE_GROUPS];
    for (int a = 0; a < NUM_AGE_GROUPS; a++) rsc[a] = P.RelativeSpatialContactSusc[a];
    P.Efficacies = effrows;
    P.TreatInfDrop = 0.61; P.VaccInfDrop = 0.83; P.TreatSuscDrop = 0.47; P.VaccSuscDrop = 0.39; P.VaccSuscDrop2 = 0.71;
    P.CareHomeResidentHouseholdScaling = 0.9; P.CareHomeResidentSpatialScaling = 0.55;
    P.usVaccTimeToEfficacy = 8; P.usVaccTimeEfficacySwitch = 120; P.usHQuarantineHouseDuration = 40;
    P.usCaseIsolationDelay = 2; P.usCaseIsolationDuration = 28;
    for (int a = 0; a < NUM_AGE_GROUPS; a++) P.RelativeSpatialContactSusc[a] = 0.5 + lcg_u();

    const int NG = std::min(6000, P.PopSize / 3);
    std::vector<int> order(P.PopSize);
    for (int i = 0; i < P.PopSize; i++) order[i] = i;
    for (int i = P.PopSize - 1; i > 0; i--) std::swap(order[i], order[lcg_i(i + 1)]);
    std::vector<int32_t> hp(NG), ho(NG), hpt(NG);
    std::vector<uint16_t> hts(NG);
    std::vector<double> out((size_t)NG * 8);
    std::vector<Person> saveH(P.PopSize);
    std::vector<PersonQuarantine> saveQ(HostsQuarantine);
    for (int i = 0; i < P.PopSize; i++) saveH[i] = Hosts[i];
    int nm = P.NMCL * P.NMCL * P.NumCells;
    std::vector<TreatStat> saveSD(nm);
    for (int i = 0; i < nm; i++) { saveSD[i] = Mcells[i].socdist; Mcells[i].socdist = (lcg_u() < 0.4) ? TreatStat::Treated : TreatStat::Untreated; }
    for (int g = 0; g < NG; g++) {
        int a = order[2 * g], b = order[2 * g + 1];
        int pt = lcg_i(P.NumPlaceTypes);
        unsigned short ts = (unsigned short)(60 + lcg_i(200));
        for (int who : {a, b}) {
            Person& h = Hosts[who];
            h.latent_time = (unsigned short)(ts - 1 - lcg_i(50));
            h.infectiousness = (float)((lcg_u() < 0.5 ? -1 : 1) * (0.1 + 2 * lcg_u()));
            h.susc = (float)(0.2 + lcg_u());
            h.treat_start_time = (unsigned short)(lcg_u() < 0.3 ? ts - lcg_i(10) : CS_UNSET_TIME);
            h.treat_stop_time = (unsigned short)(lcg_u() < 0.5 ? ts + 1 + lcg_i(10) : 0);
            h.vacc_start_time = (unsigned short)(lcg_u() < 0.3 ? lcg_i(ts) : CS_UNSET_TIME);
            h.isolation_start_time = (unsigned short)(lcg_u() < 0.3 ? ts - lcg_i(30) : CS_UNSET_TIME);
            h.digitalContactTraced = lcg_u() < 0.2 ? 1 : 0;
            h.care_home_resident = lcg_u() < 0.2 ? 1 : 0;
            h.esocdist_comply = lcg_u() < 0.5 ? 1 : 0;
            if (lcg_u() < 0.3) h.set_case(); else h.set_infectious_asymptomatic_not_case();
            HostsQuarantine[who].comply = (uint8_t)lcg_i(3);
            HostsQuarantine[who].start_time = (uint16_t)(lcg_u() < 0.5 ? ts - lcg_i(45) : CS_UNSET_TIME);
        }
        hp[g] = a; ho[g] = b; hpt[g] = pt; hts[g] = ts;
        double* o = &out[(size_t)g * 8];
        o[0] = CalcHouseInf(a, ts);
        o[1] = CalcPlaceInf(a, pt, ts);
        o[2] = CalcSpatialInf(a, ts);
        o[3] = CalcPersonInf(a, ts);
        o[4] = CalcHouseSusc(a, ts, b);
        o[5] = CalcPlaceSusc(a, pt, ts);
        o[6] = CalcSpatialSusc(a, ts);
        o[7] = CalcPersonSusc(a, ts, b);
    }
    save("g_hz_person", hp); save("g_hz_other", ho); save("g_hz_pt", hpt); save("g_hz_ts", hts);
    save2("g_hz_out", out, NG, 8);
    {
        std::vector<uint8_t> flags(P.PopSize);
        for (int i = 0; i < P.PopSize; i++) {
            const Person& a = Hosts[i];
            flags[i] = (uint8_t)((a.care_home_resident ? 1 : 0) | (a.keyworker ? 2 : 0) |
                                 (a.esocdist_comply ? 4 : 0) | (a.digitalContactTracingUser ? 8 : 0));
        }
        save("hz_p_flags", flags);
    }
    dump_state("hz_", 0);
    // restore
    for (int i = 0; i < P.PopSize; i++) Hosts[i] = saveH[i];
    HostsQuarantine = saveQ;
    for (int i = 0; i < nm; i++) Mcells[i].socdist = saveSD[i];
    for (int a = 0; a < NUM_AGE_GROUPS; a++) P.RelativeSpatialContactSusc[a] = rsc[a];
    P.Efficacies = effsave;
    P.TreatInfDrop = sv_tid; P.VaccInfDrop = sv_vid; P.TreatSuscDrop = sv_tsd; P.VaccSuscDrop = sv_vsd;
    P.VaccSuscDrop2 = sv_vsd2; P.CareHomeResidentHouseholdScaling = sv_chh; P.CareHomeResidentSpatialScaling = sv_chs;
    P.usVaccTimeToEfficacy = sv_u1; P.usVaccTimeEfficacySwitch = sv_u2; P.usHQuarantineHouseDuration = sv_u3;
    P.usCaseIsolationDelay = sv_u4; P.usCaseIsolationDuration = sv_u5;
}

void init_once()
{
    if (g_started) return;
    g_started = true;
    const char* d = getenv("CS_DUMP_DIR");
    if (d) g_dir = d;
    const char* s = getenv("CS_DUMP_STEPS");
    std::string steps = s ? s : "0";
    std::stringstream ss(steps);
    std::string tok;
    while (std::getline(ss, tok, ',')) if (!tok.empty()) g_steps.push_back(atoi(tok.c_str()));
    g_golden = getenv("CS_DUMP_GOLDEN") && atoi(getenv("CS_DUMP_GOLDEN"));
    const char* se = getenv("CS_SERIES");
    if (se) g_series = fopen(se, "ab");
}

// InitTransmissionCoeffs' per-person draws and results, for the oracle's
// restatement of its deterministic part
void dump_itc()
{
    const int N = P.PopSize;
    if (P.NumThreads != 1) { fprintf(stderr, "CS_DUMP_ITC needs /c:1\n"); exit(2); }
    std::vector<int8_t> infstat(N);
    for (int i = 0; i < N; i++) infstat[i] = (int8_t)cs_bridge::infstat_code(Hosts[i]);
    const double r0_in = P.R0, beta_before = P.LocalBeta;
    int32_t s1 = 1234567, s2 = 7654321;
    setall(&s1, &s2);
    InitTransmissionCoeffs();
    std::vector<double> out = {r0_in, P.R0household, P.R0places, P.LocalBeta, P.R0spatial, P.R0, beta_before};
    std::vector<float> infness(N), susc(N);
    for (int i = 0; i < N; i++) { infness[i] = Hosts[i].infectiousness; susc[i] = Hosts[i].susc; }
    // replay the draws of src/SetupModel.cpp:667-690 from the same seeds
    s1 = 1234567; s2 = 7654321;
    setall(&s1, &s2);
    std::vector<uint16_t> rec(N);
    for (int i = 0; i < N; i++) {
        const int a = HOST_AGE_GROUP(i);
        float su = 1.0f, in;
        if (P.SusceptibilitySD != 0)
            su = (float)(((P.DoPartialImmunity) ? (1.0 - P.InitialImmunity[a]) : 1.0) *
                         gen_gamma_mt(1 / (P.SusceptibilitySD * P.SusceptibilitySD), 1 / (P.SusceptibilitySD * P.SusceptibilitySD), 0));
        else
            su = (float)((P.DoPartialImmunity) ? (1.0 - P.InitialImmunity[a]) : 1.0);
        if (P.InfectiousnessSD == 0) in = (float)P.AgeInfectiousness[a];
        else in = (float)(P.AgeInfectiousness[a] * gen_gamma_mt(1 / (P.InfectiousnessSD * P.InfectiousnessSD), 1 / (P.InfectiousnessSD * P.InfectiousnessSD), 0));
        if (ranf_mt(0) < P.ProportionSymptomatic[a]) in *= (float)(-P.SymptInfectiousness);
        else in *= (float)P.AsymptInfectiousness;
        double q;
        int j = (int)floor((q = ranf_mt(0) * CDF_RES));
        q -= (double)j;
        rec[i] = (unsigned short)floor(0.5 - (P.InfectiousPeriod * log(q * P.infectious_icdf[j + 1] + (1.0 - q) * P.infectious_icdf[j]) / P.ModelTimeStep));
        if (in != infness[i] || su != susc[i]) { fprintf(stderr, "ITC replay diverged at person %d\n", i); exit(3); }
    }
    save("itc_infectiousness", infness); save("itc_susc", susc); save("itc_recovery", rec);
    save("itc_infstat", infstat); save("itc_out", out);
    dump_params("itc_params.bin");
}

int g_run_of_call = -1;

} // namespace

void InfectSweep(double t, int run)
{
    init_once();
    g_run_of_call = run;
    unsigned short ts = (unsigned short)(P.TimeStepsPerDay * t);
    if (!g_dir.empty() && run == 0) {
        if (ts == 0 && g_last_run != run) {
            dump_population();
            dump_params("params.bin");
            if (g_golden) dump_golden();
        }
        if (ts == 0 && getenv("CS_DUMP_ITC") && atoi(getenv("CS_DUMP_ITC"))) {
            dump_itc();
            exit(0);
        }
        for (int s : g_steps)
            if (s == ts) dump_state("s" + std::to_string(ts) + "_", ts);
    }
    g_last_run = run;
    InfectSweep_cpu(t, run);
}

void IncubRecoverySweep(double t)
{
    IncubRecoverySweep_cpu(t);
    if (g_series) {
        int32_t rec[10 + INFECT_TYPE_MASK + 4];
        int S = 0, L = 0, I = 0, R = 0, D = 0;
        for (int j = 0; j < P.NumPopulatedCells; j++) {
            S += CellLookup[j]->S; L += CellLookup[j]->L; I += CellLookup[j]->I;
            R += CellLookup[j]->R; D += CellLookup[j]->D;
        }
        int cumI = 0, cumC = 0;
        for (int k = 0; k < P.NumThreads; k++) { cumI += StateT[k].cumI; cumC += StateT[k].cumC; }
        rec[0] = g_run_of_call; rec[1] = (int32_t)(unsigned short)(P.TimeStepsPerDay * t);
        rec[2] = S; rec[3] = L; rec[4] = I; rec[5] = R; rec[6] = D; rec[7] = cumI; rec[8] = cumC; rec[9] = D;
        // StateT.cumItype is cleared by RecordSample before the first
        // timestep of each output step (src/CovidSim.cpp:3420-3421)
        static int itype_prev[INFECT_TYPE_MASK], itype_acc[INFECT_TYPE_MASK], acc_run = -1;
        if (acc_run != g_run_of_call) {
            acc_run = g_run_of_call;
            for (int q = 0; q < INFECT_TYPE_MASK; q++) itype_prev[q] = itype_acc[q] = 0;
        }
        const bool cleared = (rec[1] % P.NumModelTimeStepsPerOutputTimeStep) == 0;
        for (int q = 0; q < INFECT_TYPE_MASK; q++) {
            int cur = 0;
            for (int k = 0; k < P.NumThreads; k++) cur += StateT[k].cumItype[q];
            itype_acc[q] += cleared ? cur : cur - itype_prev[q];
            itype_prev[q] = cur;
            rec[10 + q] = itype_acc[q];
        }
        auto steps = [](double x) { return (int32_t)(x < 1e9 ? x * P.TimeStepsPerDay : -1); };
        rec[10 + INFECT_TYPE_MASK] = steps(P.PlaceCloseTimeStart);
        rec[11 + INFECT_TYPE_MASK] = P.NumAdunits > 0 ? steps(AdUnits[0].CaseIsolationTimeStart) : -1;
        rec[12 + INFECT_TYPE_MASK] = P.NumAdunits > 0 ? steps(AdUnits[0].HQuarantineTimeStart) : -1;
        rec[13 + INFECT_TYPE_MASK] = State.trigDetectedCases;
        fwrite(rec, sizeof(rec), 1, g_series);
        fflush(g_series);
    }
}
