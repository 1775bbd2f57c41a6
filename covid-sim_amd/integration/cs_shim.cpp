// cs_shim.cpp -- the binding a CovidSim maintainer adds to run the
// infection step on the GPU.
//
// CovidSim calls its two sweeps as free functions (src/Sweep.h:4-9, called
// from RunModel, src/CovidSim.cpp:1488 and :1490).  CovidSim's Sweep.cpp is
// compiled with the two renamed
//     -DInfectSweep=InfectSweep_cpu -DIncubRecoverySweep=IncubRecoverySweep_cpu
// and this translation unit supplies the original names:
//
//   InfectSweep(t, run)    -> cs_infect_sweep          (libcovidsim_hip)
//   IncubRecoverySweep(t)  -> cs_incub_recovery_sweep, then the state the
//                             host reads between timesteps is copied back
//
// At the first timestep of every realisation (t == 0, after the host's
// InitModel + SeedInfection) the host state is loaded into the library, with
// a stream seed taken from the realisation's chained run seeds
// (src/CovidSim.cpp:352-369).  When the host's UpdateProbs has rebuilt its
// cell tables (a cell's S0 changed) the library rebuilds its own from the
// same counts.  The counters the sweeps add are added to StateT[0], so
// RecordSample and the output files see the GPU run.
//
// CS_SHIM=cpu in the environment runs the reference's own sweeps instead.
// CS_SHIM_TIMING=1 prints, at exit, where the GPU path's time went: the two
// sweeps on the device, the state copies back to the host arrays after each
// timestep, the state loads at each realisation's start and the table
// rebuilds.
// Any library error ends the run through ERR_CRITICAL, the reference's
// convention for fatal errors.
//
// Scope: the no-intervention path; the library refuses parameter sets that
// switch interventions on.  The counters by administrative unit (cumI, cumC,
// cumDC, cumD per unit) are added to StateT[0] beside the others when the
// parameters keep them (P.DoAdUnits with units known).
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <string>
#include <vector>

// KernelLookup keeps its tables private (src/Kernels.h:79-87); the shim
// passes them to the library as they are, so the class is opened up for
// this one translation unit.
#define private public
#include "Kernels.h"
#undef private

#include "Constants.h"
#include "Error.h"
#include "InverseCdf.h"
#include "Model.h"
#include "Param.h"
#include "Sweep.h"

#include "cs_reference_bridge.h"

void InfectSweep_cpu(double t, int run);
void IncubRecoverySweep_cpu(double t);

namespace {

struct Shim {
    bool init = false, gpu = true;
    cs_engine* eng = nullptr;
    cs_params prm;
    cs_bridge::PopulationView pop;
    cs_bridge::StateView st;
    std::vector<int32_t> S0;  // cell S0 the library's tables were built from
    std::vector<int64_t> adc; // counters by administrative unit, [NumAdunits][CS_TSU_FIELDS]
    bool ad = false;          // the library keeps them
    bool timing = false;      // CS_SHIM_TIMING
    double t_sweep = 0, t_copy = 0, t_load = 0, t_probs = 0;
    long n_steps = 0, n_loads = 0;
};
Shim g;

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void report_timing()
{
    fprintf(stderr,
            "covidsim_hip shim: %ld timesteps: sweeps on the device %.3f s, state copies to the host %.3f s "
            "(%.3f ms a step), %ld state loads %.3f s, table rebuilds %.3f s\n",
            g.n_steps, g.t_sweep, g.t_copy, g.n_steps ? 1e3 * g.t_copy / (double)g.n_steps : 0.0, g.n_loads, g.t_load,
            g.t_probs);
}

void check(int rc, const char* what)
{
    if (rc) ERR_CRITICAL_FMT("covidsim_hip: %s failed (%d): %s\n", what, rc, cs_last_error(g.eng));
}

uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

void setup_engine()
{
    g.init = true;
    const char* mode = getenv("CS_SHIM");
    g.gpu = !(mode && strcmp(mode, "cpu") == 0);
    if (!g.gpu) return;
    g.timing = getenv("CS_SHIM_TIMING") != nullptr;
    if (g.timing) atexit(report_timing);
    check(cs_create(&g.eng, 0), "cs_create");
    cs_bridge::fill_params(&g.prm);
    check(cs_set_params(g.eng, &g.prm, P.KernelLookup.lookup_.data(), P.KernelLookup.hi_res_.data()), "cs_set_params");
    g.pop.gather();
    cs_population v = g.pop.view();
    check(cs_set_population(g.eng, &v), "cs_set_population");
}

// host -> library, after InitModel + SeedInfection of realisation `run`
void push_state(int run)
{
    cs_bridge::fill_params(&g.prm);  // run-start values (seed location, CFR scales)
    check(cs_set_params(g.eng, &g.prm, nullptr, nullptr), "cs_set_params");
    const uint64_t seed = splitmix64(((uint64_t)(uint32_t)P.nextRunSeed1 << 32 | (uint32_t)P.nextRunSeed2) ^ (uint64_t)run);
    check(cs_alloc_realisations(g.eng, 1, &seed), "cs_alloc_realisations");
    g.st.gather();
    cs_state v = g.st.view();
    check(cs_load_state(g.eng, 0, &v), "cs_load_state");
    cs_counters c;
    check(cs_get_counters(g.eng, 0, &c, 1), "cs_get_counters");
    g.adc.assign((size_t)P.NumAdunits * CS_TSU_FIELDS, 0);
    g.ad = P.DoAdUnits && P.NumAdunits > 0 && cs_get_adunit_counters(g.eng, 0, g.adc.data(), 1) == 0;
    g.S0.resize(P.NumPopulatedCells);
    for (int j = 0; j < P.NumPopulatedCells; j++) g.S0[j] = CellLookup[j]->S0;
}

void add_counters(const cs_counters& d)
{
    PopVar& z = StateT[0];
    z.cumI += (int)d.cumI; z.cumC += (int)d.cumC; z.cumDC += (int)d.cumDC; z.cumFC += (int)d.cumFC;
    z.cumTG += (int)d.cumTG; z.cumSI += (int)d.cumSI; z.nTG += (int)d.nTG;
    z.cumAC += (int)d.cumAC; z.cumACS += (int)d.cumACS; z.cumAA += (int)d.cumAA;
    z.sumRad2 += d.sumRad2;
    if (d.maxRad2 > z.maxRad2) z.maxRad2 = d.maxRad2;
    for (int q = 0; q < INFECT_TYPE_MASK; q++) z.cumItype[q] += (int)d.cumItype[q];
    for (int a = 0; a < NUM_AGE_GROUPS; a++) {
        z.cumIa[a] += (int)d.cumIa[a]; z.cumCa[a] += (int)d.cumCa[a]; z.cumDa[a] += (int)d.cumDa[a];
    }
    for (int k = 0; k < 2; k++) { z.cumI_keyworker[k] += (int)d.cumI_keyworker[k]; z.cumC_keyworker[k] += (int)d.cumC_keyworker[k]; }
    int* prev[5] = {&z.Mild, &z.ILI, &z.SARI, &z.Critical, &z.CritRecov};
    int* prev_age[5] = {z.Mild_age, z.ILI_age, z.SARI_age, z.Critical_age, z.CritRecov_age};
    int* cum[8] = {&z.cumMild, &z.cumILI, &z.cumSARI, &z.cumCritical, &z.cumCritRecov,
                   &z.cumDeath_ILI, &z.cumDeath_SARI, &z.cumDeath_Critical};
    int* cum_age[8] = {z.cumMild_age, z.cumILI_age, z.cumSARI_age, z.cumCritical_age, z.cumCritRecov_age,
                       z.cumDeath_ILI_age, z.cumDeath_SARI_age, z.cumDeath_Critical_age};
    for (int k = 0; k < CS_NUM_SEV_KINDS; k++) {
        *cum[k] += (int)d.sev_cum[k];
        for (int a = 0; a < NUM_AGE_GROUPS; a++) cum_age[k][a] += (int)d.sev_cum_age[k][a];
        if (k < 5) {
            *prev[k] += (int)d.sev_prev[k];
            for (int a = 0; a < NUM_AGE_GROUPS; a++) prev_age[k][a] += (int)d.sev_prev_age[k][a];
        }
    }
}

// library -> host: the state RecordSample, UpdateProbs and the output code read
void pull_state()
{
    cs_state v = g.st.view();
    check(cs_store_state(g.eng, 0, &v), "cs_store_state");
    g.st.scatter();
    cs_counters d;
    check(cs_get_counters(g.eng, 0, &d, 1), "cs_get_counters");
    add_counters(d);
    if (g.ad) {
        check(cs_get_adunit_counters(g.eng, 0, g.adc.data(), 1), "cs_get_adunit_counters");
        PopVar& z = StateT[0];
        for (int i = 0; i < P.NumAdunits; i++) {
            const int64_t* a = &g.adc[(size_t)i * CS_TSU_FIELDS];
            z.cumI_adunit[i] += (int)a[CS_TSU_INC_I]; z.cumC_adunit[i] += (int)a[CS_TSU_INC_C];
            z.cumDC_adunit[i] += (int)a[CS_TSU_INC_DC]; z.cumD_adunit[i] += (int)a[CS_TSU_INC_DEATH];
        }
    }
}

}  // namespace

void InfectSweep(double t, int run)
{
    if (!g.init) setup_engine();
    if (!g.gpu) {
        InfectSweep_cpu(t, run);
        return;
    }
    double t0 = g.timing ? now_s() : 0.0;
    if (t == 0) {
        push_state(run);
        if (g.timing) { g.t_load += now_s() - t0; g.n_loads++; }
    } else {
        // the host's UpdateProbs rebuilt its cell tables: rebuild ours from the same counts
        bool changed = false;
        for (int j = 0; j < P.NumPopulatedCells && !changed; j++) changed = CellLookup[j]->S0 != g.S0[j];
        if (changed) {
            check(cs_update_probs(g.eng), "cs_update_probs");
            for (int j = 0; j < P.NumPopulatedCells; j++) g.S0[j] = CellLookup[j]->S0;
            if (g.timing) { check(cs_sync(g.eng), "cs_sync"); g.t_probs += now_s() - t0; }
        }
    }
    t0 = g.timing ? now_s() : 0.0;
    check(cs_infect_sweep(g.eng, t), "cs_infect_sweep");
    if (g.timing) { check(cs_sync(g.eng), "cs_sync"); g.t_sweep += now_s() - t0; }
}

void IncubRecoverySweep(double t)
{
    if (!g.gpu) {
        IncubRecoverySweep_cpu(t);
        return;
    }
    double t0 = g.timing ? now_s() : 0.0;
    check(cs_incub_recovery_sweep(g.eng, t), "cs_incub_recovery_sweep");
    if (g.timing) {
        check(cs_sync(g.eng), "cs_sync");
        const double t1 = now_s();
        g.t_sweep += t1 - t0;
        t0 = t1;
    }
    pull_state();
    if (g.timing) { g.t_copy += now_s() - t0; g.n_steps++; }
}
