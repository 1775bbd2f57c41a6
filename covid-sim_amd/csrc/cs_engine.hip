// cs_engine.hip -- host side of libcovidsim_hip.so: device memory, uploads,
// step orchestration and the C ABI of include/covidsim_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/covidsim_hip.h"
#include "cs_kernels_itc.h"
#include "cs_kernels_itv.h"

using cs::Dev;

// destination tables chosen at allocation: the full NC x NC max_trans and
// per-realisation cum_trans rows while (realisations + 1) x NC^2 floats stay
// within CS_FULL_TABLE_BYTES, else checkpoints at the smallest power-of-two
// stride (at least 8) whose R x NC x NC / stride floats stay within
// CS_CK_TABLE_BYTES (a destination search redoes at most a stride of terms)
#define CS_FULL_TABLE_BYTES (16ull << 30)
#define CS_CK_TABLE_BYTES (16ull << 30)
// the rows' bucket index of checkpointed tables (R x NC x 1025 block
// numbers) is kept while it fits this
#define CS_CK_INDEX_BYTES (4ull << 30)

struct cs_engine {
    char err[512];
    int device = 0;
    hipStream_t stream = nullptr;
    cs_params P;
    bool have_params = false, have_pop = false;
    bool due_ok = false;             // the parameters allow latent due lists
    cs_params* dP = nullptr;
    double *dkl = nullptr, *dkh = nullptr;
    int N = 0, NH = 0, NC = 0, NCT = 0, NM = 0, NA = 0, NPT = 0, R = 0;
    int max_cn = 0;                  // the largest cell's persons
    int move_cap = CS_MOVE_CAP;      // Dev.move_cap (cs_set_move_capacity)
    int np[CS_MAX_PLACE_TYPES] = {0};
    std::vector<void*> pop_allocs, real_allocs;
    Dev d;
    int* d_sel = nullptr;
    bool adunits_valid = false;      // every person's administrative unit lies in [0, n_adunits)
    double* rbeta_buf = nullptr;     // [R] storage behind Dev.rbeta
    int32_t* rseed_buf = nullptr;    // [R] storage behind Dev.rseed_n
    double* rthr_buf = nullptr;      // [R] storage behind Dev.rthr
    int dest_stride = -1;            // destination tables asked for: -1 chosen at allocation, 0 full, > 0 checkpoint stride
    bool dest_index = true;          // checkpointed tables keep the rows' bucket index (Dev.ck_inv) when it fits
    int cxy_nch = 0;                 // the nch Dev.cxy was built with
    float* mt_buf = nullptr;         // [NC][NC] max_trans (full tables only; built when first needed)
    uint8_t* pin = nullptr;          // pinned host staging of cs_store_state (grown as needed)
    size_t pin_cap = 0;
    int ts_max = 0, ts_n = 0;
    std::vector<uint8_t> flags0;     // [N] each person's Hot flags after a reset (no links, household size), for loaded states
    std::vector<uint8_t> h_age, h_flags;  // [N] the copies a loaded state's Hot records take
    bool ts_fresh = true;            // no sample since the last alloc/reset/load: the series restarts
    unsigned grid_waves = 4096;      // blocks of 4 waves for the per-person passes (a wave per chunk)
    unsigned grid_cells = 2048;      // blocks for the per-cell spatial sums (4 waves, a wave per cell)
    unsigned grid_moves = 4096;      // blocks for the batched list moves (each a range of cells; not tuned)
    // replayed step graphs (cs_run_steps): one graph per output step
    bool use_graphs = true;
    hipGraphExec_t gexec = nullptr;
    int graph_sps = 0;

    // profiling: event pairs per kernel kind, resolved on cs_get_profile
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_open;
    double prof_ms[CS_NUM_KERNEL_KINDS] = {0};
    long long prof_n[CS_NUM_KERNEL_KINDS] = {0};
};

namespace {

int set_err(cs_engine* e, int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(e->err, sizeof(e->err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(e, x)                                                                    \
    do {                                                                                \
        hipError_t _st = (x);                                                           \
        if (_st != hipSuccess)                                                          \
            return set_err(e, CS_EHIP, "%s: %s (%s:%d)", #x, hipGetErrorString(_st), __FILE__, __LINE__); \
    } while (0)

template <class T>
int dalloc(cs_engine* e, std::vector<void*>& list, T** out, size_t n)
{
    void* p = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t st = hipMalloc(&p, bytes);
    if (st != hipSuccess) return set_err(e, CS_ENOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(st));
    list.push_back(p);
    *out = (T*)p;
    return 0;
}

template <class T>
int dupload(cs_engine* e, std::vector<void*>& list, const T** out, const T* src, size_t n)
{
    T* p;
    int rc = dalloc(e, list, &p, n);
    if (rc) return rc;
    if (n && src) HIPCHK(e, hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    *out = p;
    return 0;
}

void free_list(std::vector<void*>& list)
{
    for (void* p : list) (void)hipFree(p);
    list.clear();
}

// KernelStruct families (src/Kernels.cpp:68-106) with the host's libm, so
// the tables equal KernelLookup::init's (src/Kernels.cpp:17-45, norm 1.0)

hipEvent_t pool_event(cs_engine* e)
{
    if (!e->ev_pool.empty()) {
        hipEvent_t ev = e->ev_pool.back();
        e->ev_pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    (void)hipEventCreate(&ev);
    return ev;
}

struct ProfScope {  // records a start/stop event pair around one launch
    cs_engine* e;
    int kind;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(cs_engine* e_, int k) : e(e_), kind(k)
    {
        if (e->prof) {  // never inside a captured graph: run_steps launches directly while profiling
            a = pool_event(e);
            b = pool_event(e);
            (void)hipEventRecord(a, e->stream);
        }
    }
    ~ProfScope()
    {
        if (e->prof) {
            (void)hipEventRecord(b, e->stream);
            e->ev_open.push_back({kind, {a, b}});
        }
    }
};

void resolve_profile(cs_engine* e)
{
    for (auto& x : e->ev_open) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, x.second.first, x.second.second) == hipSuccess) {
            e->prof_ms[x.first] += ms;
            e->prof_n[x.first]++;
        }
        e->ev_pool.push_back(x.second.first);
        e->ev_pool.push_back(x.second.second);
    }
    e->ev_open.clear();
}

unsigned grid_for(long long n, int block, unsigned cap = 65535u * 16)
{
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (long long)cap) g = cap;
    return (unsigned)g;
}

// the parameter sets this library serves: the path without interventions,
// or with the intervention subset (case isolation and household quarantine
// of detected cases, place closure by microcell, on global detected-case or
// ICU triggers, with the outbreak alert of CalibrationThresholdCheck when no
// calibration is asked for); anything else is refused, never approximated
const char* unsupported(const cs_params& p)
{
    const double never = 10000.0;
    if (p.false_positive_rate != 0.0) return "false-positive infections are outside this path";
    if (p.do_digital_contact_tracing) return "digital contact tracing is outside this path";
    if (p.do_blanket_move_restr) return "movement restrictions are outside this path";
    if (p.absenteeism_place_closure) return "absenteeism place closure is outside this path";
    if (!p.interventions_active) return nullptr;
    if (!p.do_global_triggers || p.do_admin_triggers) return "interventions need global triggers (admin-unit and microcell triggers are outside this path)";
    if (p.do_per_capita_triggers) return "per-capita triggers are outside this path";
    if (p.do_no_calibration || p.do_alert_trigger_after_interv || p.date_trigger_reached_cal_time >= 0)
        return "calibration to a trigger date is outside this path";
    if (p.place_close_by_admin_unit || p.do_intervention_delays_by_adunit) return "admin-unit place closure and delays are outside this path";
    if (p.vary_efficacies_over_time) return "efficacies varying over time are outside this path";
    if (p.do_mass_vacc || p.treat_time_start_base < never || p.vacc_time_start_base < never || p.vacc_time_start_geo < never ||
        p.soc_dist_time_start_base < never || p.move_restr_time_start_base < never || p.key_worker_proph_time_start_base < never ||
        p.airport_close_time_start_base < never)
        return "treatment, vaccination, social distancing, movement restrictions, key-worker prophylaxis and airport closure are outside this path";
    if (p.place_close_time_start_base2 < never) return "a second place closure is outside this path";
    if (p.place_close_indep_thresh > 0) return "the independent place-closure threshold is outside this path";
    if (p.place_close_cell_inc_stop_thresh != 0) return "reopening places on a stop threshold is outside this path";
    if (p.triggers_sampling_interval < 1 || p.window_to_evaluate_trigger_alert < 1) return "trigger windows must be at least one sample";
    return nullptr;
}

// a whole run (reset, cs_run_steps): the served parameter sets
int check_path(cs_engine* e)
{
    const char* why = unsupported(e->P);
    return why ? set_err(e, CS_EUNSUPPORTED, "%s", why) : 0;
}

// one sweep called on its own (the drop-in's InfectSweep and
// IncubRecoverySweep): interventions need the triggers and TreatSweep that
// RunModel runs between sweeps, so only cs_run_steps serves them
int check_sweep(cs_engine* e)
{
    int rc = check_path(e);
    if (rc) return rc;
    if (e->P.interventions_active)
        return set_err(e, CS_EUNSUPPORTED, "interventions run only in cs_run_steps (the triggers and TreatSweep run between sweeps)");
    return 0;
}

int check_flags(cs_engine* e)
{
    unsigned f = 0;
    HIPCHK(e, hipMemcpyAsync(&f, e->d.status, sizeof(f), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    if (f & cs::FLAG_RANGE) return set_err(e, CS_ERANGE, "r too large in NumKernel (distance beyond the kernel table)");
    if (f & cs::FLAG_OVERFLOW) return set_err(e, CS_EOVERFLOW, "infection event queue overflow (capacity %u)", e->d.ev_cap);
    if (f & cs::FLAG_BATCH) return set_err(e, CS_ESTATE, "a batched list move found a list position outside its cell's segment");
    if (f & cs::FLAG_STUCK) return set_err(e, CS_ERANGE, "spatial draw found no live candidate in %d attempts", CS_MAX_ATTEMPTS);
    if (f & cs::FLAG_TRIG) return set_err(e, CS_EUNSUPPORTED, "a trigger window reaches back more than %d samples", CS_TRIG_RING - 1);
    if (f & cs::FLAG_LIST) return set_err(e, CS_EOVERFLOW, "intervention lists full (detected cases %u, closing places %u)", e->d.det_cap, e->d.close_cap);
    return 0;
}

// the batched list moves of one mode: cells within the LDS arrays' reach in
// k_list_move, the cells it queues in k_list_move_big
void launch_list_move(cs_engine* e, hipStream_t st, int mode)
{
    cs::k_list_move<<<e->grid_moves, 256, 0, st>>>(e->d, mode);
    cs::k_list_move_big<<<CS_MOVE_BIG_BLOCKS, 256, 0, st>>>(e->d, mode);
}

// sweep kernels that append events; state is untouched until resolution

int launch_infect_events(cs_engine* e)
{
    Dev& d = e->d;
    hipStream_t st = e->stream;
    const unsigned nrc = (unsigned)((long long)d.R * d.NC);
    {
        ProfScope ps(e, CS_K_OTHER);
        cs::k_zero_counts<<<1, CS_EV_SHARDS * CS_EV_STRIDE, 0, st>>>(d.ev_count, CS_EV_SHARDS * CS_EV_STRIDE);
        cs::k_zero_counts<<<1, CS_EV_SHARDS * CS_EV_STRIDE, 0, st>>>(d.pc_count, CS_EV_SHARDS * CS_EV_STRIDE);
        cs::k_zero_counts<<<1, 64, 0, st>>>(d.wl, cs::WL_N);
        cs::k_compact<<<grid_for(nrc, 256), 256, 0, st>>>(d, 1, 1);
    }
    { ProfScope ps(e, CS_K_INFECT_HOUSEHOLDS); cs::k_infect_households<<<e->grid_waves, 256, 0, st>>>(d); }
    {
        ProfScope ps(e, CS_K_INFECT_PLACES);
        cs::k_infect_places<<<e->grid_waves, 256, 0, st>>>(d);
        cs::k_place_contacts<<<e->grid_waves, 256, 0, st>>>(d);
    }
    { ProfScope ps(e, CS_K_CELL_SPATIAL_SUM); cs::k_cell_spatial_sum<<<e->grid_cells, CS_SUM_BLOCK, 0, st>>>(d); }
    { ProfScope ps(e, CS_K_OTHER); cs::k_compact<<<grid_for(nrc, 256), 256, 0, st>>>(d, 2, 0); }
    {
        ProfScope ps(e, CS_K_SPATIAL_DRAWS);
        if (d.ck) cs::k_spatial_draws<true><<<e->grid_waves, 256, 0, st>>>(d);
        else cs::k_spatial_draws<false><<<e->grid_waves, 256, 0, st>>>(d);
    }
    HIPCHK(e, hipGetLastError());
    return 0;
}

int launch_resolve(cs_engine* e)
{
    Dev& d = e->d;
    hipStream_t st = e->stream;
    {
        ProfScope ps(e, CS_K_RESOLVE);
        cs::k_resolve_claim<<<e->grid_waves, 256, 0, st>>>(d);
        cs::k_resolve_apply<<<e->grid_waves, 256, 0, st>>>(d);
        cs::k_resolve_reset<<<e->grid_waves, 256, 0, st>>>(d);
        cs::k_cnt_fold<<<dim3(d.R, (cs::CS_CNT_WORDS + 63) / 64), 64, 0, st>>>(d);
    }
    { ProfScope ps(e, CS_K_LIST_MOVE); launch_list_move(e, st, 0); }
    HIPCHK(e, hipGetLastError());
    return 0;
}

int launch_incub(cs_engine* e)
{
    Dev& d = e->d;
    hipStream_t st = e->stream;
    const unsigned nrc = (unsigned)((long long)d.R * d.NC);
    if (e->P.do_places && e->P.num_holidays > 0) {
        long long n = (long long)d.R * d.np_off[CS_MAX_PLACE_TYPES];
        // a holiday starts in few steps: a modest grid (the places loop is
        // grid-strided) keeps the other steps' launch to a uniform test
        cs::k_holidays<<<grid_for(n, 256, 2048), 256, 0, st>>>(d);
    }
    {
        ProfScope ps(e, CS_K_OTHER);
        cs::k_zero_counts<<<1, 64, 0, st>>>(d.wl, cs::WL_N);
        cs::k_compact<<<grid_for(nrc, 256), 256, 0, st>>>(d, 0, 0);
    }
    {
        ProfScope ps(e, CS_K_LATENT_SCAN);
        // the due list when it holds the step's persons, else the scan (each
        // kernel does nothing in the other case)
        cs::k_latent_due<<<e->grid_waves, 256, 0, st>>>(d);
        cs::k_latent_scan<<<e->grid_waves, 256, 0, st>>>(d);
        cs::k_due_clear<<<1, 64, 0, st>>>(d);
    }
    { ProfScope ps(e, CS_K_LIST_MOVE); launch_list_move(e, st, 1); }
    {
        ProfScope ps(e, CS_K_OTHER);
        cs::k_zero_counts<<<1, 64, 0, st>>>(d.care_count, 1);
        cs::k_zero_counts<<<1, 64, 0, st>>>(d.wl, cs::WL_N);
        if (e->P.interventions_active) cs::k_zero_counts<<<1, 64, 0, st>>>(d.itv_n, 3);
        cs::k_compact<<<grid_for(nrc, 256), 256, 0, st>>>(d, 1, 0);
    }
    { ProfScope ps(e, CS_K_INFECTIOUS_SCAN); cs::k_infectious_scan<<<e->grid_waves, 256, 0, st>>>(d); }
    if (e->P.interventions_active) {  // DoDetectedCase: quarantine, then isolation
        ProfScope ps(e, CS_K_OTHER);
        cs::k_detected_hq<<<256, 256, 0, st>>>(d);
        cs::k_detected_ci<<<256, 256, 0, st>>>(d);
    }
    {
        ProfScope ps(e, CS_K_OTHER);
        cs::k_care_eval<<<256, 256, 0, st>>>(d);
        cs::k_care_apply<<<256, 256, 0, st>>>(d);
    }
    { ProfScope ps(e, CS_K_LIST_MOVE); launch_list_move(e, st, 2); }
    { ProfScope ps(e, CS_K_OTHER); cs::k_cnt_fold<<<dim3(d.R, (cs::CS_CNT_WORDS + 63) / 64), 64, 0, st>>>(d); }
    HIPCHK(e, hipGetLastError());
    return 0;
}

// the full max_trans table (src/Kernels.cpp:48-63), built once per
// population and kernel
int ensure_max_trans(cs_engine* e)
{
    if (e->mt_buf) return 0;
    const size_t NC = (size_t)e->NC;
    if (hipMalloc(&e->mt_buf, NC * NC * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        e->mt_buf = nullptr;
        return set_err(e, CS_ENOMEM, "max_trans: %zu x %zu cells (use checkpointed destination tables)", NC, NC);
    }
    e->d.max_trans = e->mt_buf;
    cs::k_build_max_trans<<<grid_for((long long)NC * NC, 256, 1u << 30), 256, 0, e->stream>>>(e->d, e->mt_buf);
    HIPCHK(e, hipGetLastError());
    return check_flags(e);
}

int launch_update_probs(cs_engine* e, const int* sel)
{
    Dev& d = e->d;
    long long n = (long long)d.R * d.NC;
    ProfScope ps(e, CS_K_UPDATE_PROBS);
    cs::k_update_s0<<<grid_for(n, 256), 256, 0, e->stream>>>(d, sel);
    if (d.ck) {
        cs::k_update_rows_ck<<<grid_for(n, 64), 64, 0, e->stream>>>(d, sel);
        if (d.ck_inv) cs::k_ck_index<<<grid_for(n * 64, 256), 256, 0, e->stream>>>(d, sel);
    }
    else cs::k_update_rows_wave<<<grid_for(n * 64, 256, 4096), 256, 0, e->stream>>>(d, sel);
    HIPCHK(e, hipGetLastError());
    return 0;
}

void drop_graph(cs_engine* e)
{
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    e->gexec = nullptr;
}

// the place contact list: a quarter of the event queue's capacity (at least
// 8192 entries), whole shards; sized afresh whenever the queue is
int size_contacts(cs_engine* e)
{
    Dev& d = e->d;
    unsigned cap = std::max(d.ev_cap / 4u, 8192u);
    cap -= cap % CS_EV_SHARDS;
    if (d.pcl && cap == d.pc_cap) return 0;
    if (d.pcl) {
        (void)hipFree(d.pcl);
        e->real_allocs.erase(std::remove(e->real_allocs.begin(), e->real_allocs.end(), (void*)d.pcl), e->real_allocs.end());
        d.pcl = nullptr;
    }
    int rc = dalloc(e, e->real_allocs, &d.pcl, cap);
    if (rc) return rc;
    d.pc_cap = cap;
    return 0;
}

int grow_events(cs_engine* e)
{
    Dev& d = e->d;
    unsigned cap = d.ev_cap * 2u;  // stays a multiple of CS_EV_SHARDS
    if (cap < d.ev_cap) return set_err(e, CS_EOVERFLOW, "event queue cannot grow further");
    HIPCHK(e, hipStreamSynchronize(e->stream));
    drop_graph(e);  // the graph holds the old buffers
    (void)hipFree(d.ev_key); (void)hipFree(d.ev_gp); (void)hipFree(d.ev_infector); (void)hipFree(d.ev_type);
    auto drop = [&](void* p) { e->real_allocs.erase(std::remove(e->real_allocs.begin(), e->real_allocs.end(), p), e->real_allocs.end()); };
    drop(d.ev_key); drop(d.ev_gp); drop(d.ev_infector); drop(d.ev_type);
    int rc;
    if ((rc = dalloc(e, e->real_allocs, &d.ev_key, cap))) return rc;
    if ((rc = dalloc(e, e->real_allocs, &d.ev_gp, cap))) return rc;
    if ((rc = dalloc(e, e->real_allocs, &d.ev_infector, cap))) return rc;
    if ((rc = dalloc(e, e->real_allocs, &d.ev_type, cap))) return rc;
    d.ev_cap = cap;
    if ((rc = size_contacts(e))) return rc;  // the place contact list shares the flag
    // the spatial draw list shares the overflow flag: grow it too
    const unsigned dcap = d.draw_cap * 2u > d.draw_cap ? d.draw_cap * 2u : 0xffffffffu;
    (void)hipFree(d.drawl);
    drop(d.drawl);
    if ((rc = dalloc(e, e->real_allocs, &d.drawl, dcap))) return rc;
    d.draw_cap = dcap;
    unsigned zero = 0;
    HIPCHK(e, hipMemcpy(d.status, &zero, sizeof(zero), hipMemcpyHostToDevice));
    return 0;
}

// one InfectSweep at time t; an overflow of the event or draw lists is
// detected before any state changes and the sweep is re-run with larger
// lists (the counter-based streams make the re-run identical)
int infect_sweep(cs_engine* e, double t)
{
    for (int attempt = 0; attempt < 8; attempt++) {
        cs::k_step_begin<<<1, 64, 0, e->stream>>>(e->d, t, 0);
        int rc = launch_infect_events(e);
        if (rc) return rc;
        rc = check_flags(e);
        if (rc == CS_EOVERFLOW) {
            if ((rc = grow_events(e))) return rc;
            continue;
        }
        if (rc) return rc;
        if ((rc = launch_resolve(e))) return rc;
        return check_flags(e);
    }
    return set_err(e, CS_EOVERFLOW, "event queue overflow persists");
}

// the tail of a timestep after InfectSweep: IncubRecoverySweep, the S-drop
// UpdateProbs, the sample at the end of an output step, the device clock
int launch_step_tail(cs_engine* e, bool sample_done)
{
    Dev& d = e->d;
    int rc;
    if ((rc = launch_incub(e))) return rc;
    if (e->P.interventions_active) {  // TreatSweep and UpdateHostClosure (src/CovidSim.cpp:1498-1503)
        ProfScope ps(e, CS_K_OTHER);
        cs::k_zero_counts<<<1, 64, 0, e->stream>>>(d.itv_n + 1, 2);
        cs::k_treat_mcells<<<grid_for((long long)d.R * d.NM, 256), 256, 0, e->stream>>>(d);
        cs::k_close_members<<<e->grid_waves, 256, 0, e->stream>>>(d, 0);
        cs::k_close_members<<<e->grid_waves, 256, 0, e->stream>>>(d, 1);
        cs::k_host_closure<<<256, 256, 0, e->stream>>>(d);
    }
    cs::k_sdrop<<<grid_for(d.R, 64), 64, 0, e->stream>>>(d, e->d_sel);
    if ((rc = launch_update_probs(e, e->d_sel))) return rc;
    if (sample_done) cs::k_record<<<d.R, 256, 0, e->stream>>>(d, -1);
    cs::k_step_end<<<1, 64, 0, e->stream>>>(d, sample_done ? 1 : 0);
    HIPCHK(e, hipGetLastError());
    return 0;
}

// every launch of one output step (sps timesteps) on the device clock,
// captured once into a graph for replay; an overflow halts every later
// kernel of the replay, so the state stays that of the overflowing step's
// start and the host can re-run it directly
int capture_graph(cs_engine* e, int sps)
{
    drop_graph(e);
    Dev& d = e->d;
    HIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < sps && !rc; k++) {
        cs::k_step_begin<<<1, 64, 0, e->stream>>>(d, 0.0, 1);
        if (!(rc = launch_infect_events(e)) && !(rc = launch_resolve(e))) rc = launch_step_tail(e, k == sps - 1);
    }
    hipGraph_t g = nullptr;
    hipError_t st = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (st != hipSuccess) return set_err(e, CS_EHIP, "step graph capture: %s", hipGetErrorString(st));
    st = hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (st != hipSuccess) {
        e->gexec = nullptr;
        return set_err(e, CS_EHIP, "step graph instantiation: %s", hipGetErrorString(st));
    }
    e->graph_sps = sps;
    return 0;
}

int read_clock(cs_engine* e, cs::Clock* c)
{
    HIPCHK(e, hipMemcpyAsync(c, e->d.clock, sizeof(*c), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return 0;
}

// timesteps [from, n_steps) of a run, one at a time, from time t (the time
// of step `from`, accumulated step by step as the device clock and RunModel
// do); a sample follows every sps-th step counted from the run's start
int run_direct(cs_engine* e, double t, int from, int n_steps, int sps)
{
    int rc;
    for (int step = from; step < n_steps; step++) {
        if ((rc = infect_sweep(e, t))) return rc;
        if ((rc = launch_step_tail(e, (step + 1) % sps == 0))) return rc;
        t += e->P.model_time_step;
    }
    return 0;
}

// the lists of the intervention kernels, sized on the first run with an
// intervention set: the step's detected cases (at most one case per person;
// an eighth of the ensemble's persons, at least 2^16) and its closing places
int ensure_itv(cs_engine* e)
{
    Dev& d = e->d;
    if (!e->P.interventions_active || d.det) return 0;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    drop_graph(e);
    const size_t RN = (size_t)d.R * d.N;
    d.det_cap = (unsigned)std::min<size_t>(std::max<size_t>(RN / 8, (size_t)1 << 16), std::min<size_t>(RN, 0x7fffffff));
    d.close_cap = (unsigned)std::min<size_t>(std::max<size_t>((size_t)d.R * d.np_off[CS_MAX_PLACE_TYPES], 1), 0x7fffffff);
    int rc;
    if ((rc = dalloc(e, e->real_allocs, &d.det, d.det_cap))) return rc;
    if ((rc = dalloc(e, e->real_allocs, &d.close_l, d.close_cap))) return rc;
    return 0;
}

template <class T>
struct DBuf {
    T* p = nullptr;
    ~DBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

extern "C" {

const char* cs_last_error(const cs_engine* e) { return e ? e->err : "null engine"; }
const char* cs_version(void) { return "covidsim_hip 0.1 (gfx950)"; }

int cs_create(cs_engine** out, int device)
{
    if (!out) return CS_EINVAL;
    cs_engine* e = new cs_engine();
    *out = e;
    e->err[0] = 0;
    e->device = device;
    int n = 0;
    HIPCHK(e, hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_err(e, CS_EINVAL, "device %d not present (%d devices)", device, n);
    HIPCHK(e, hipSetDevice(device));
    HIPCHK(e, hipStreamCreate(&e->stream));
    memset(&e->d, 0, sizeof(e->d));
    int rc;
    if ((rc = dalloc(e, e->pop_allocs, &e->d.status, 1))) return rc;
    if ((rc = dalloc(e, e->pop_allocs, &e->d.wc, 1))) return rc;
    if ((rc = dalloc(e, e->pop_allocs, &e->d.sc, 1))) return rc;
    if ((rc = dalloc(e, e->pop_allocs, &e->d.clock, 1))) return rc;
    HIPCHK(e, hipMemset(e->d.sc, 0, sizeof(cs::StepScalars)));
    HIPCHK(e, hipMemset(e->d.clock, 0, sizeof(cs::Clock)));
    HIPCHK(e, hipMemset(e->d.status, 0, sizeof(unsigned)));
    HIPCHK(e, hipMemset(e->d.wc, 0, sizeof(cs::Work)));
    return 0;
}

void cs_destroy(cs_engine* e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    free_list(e->real_allocs);
    free_list(e->pop_allocs);
    if (e->mt_buf) (void)hipFree(e->mt_buf);
    if (e->pin) (void)hipHostFree(e->pin);
    if (e->dP) (void)hipFree(e->dP);
    if (e->dkl) (void)hipFree(e->dkl);
    if (e->dkh) (void)hipFree(e->dkh);
    resolve_profile(e);
    for (hipEvent_t ev : e->ev_pool) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int cs_set_params(cs_engine* e, const cs_params* p, const double* kl, const double* kh)
{
    if (!e || !p) return CS_EINVAL;
    HIPCHK(e, hipSetDevice(e->device));
    if (p->kernel_type < 1 || p->kernel_type > 7) return set_err(e, CS_EINVAL, "unknown kernel type %d", p->kernel_type);
    if (p->kernel_size <= 0 || !(p->kernel_delta > 0)) return set_err(e, CS_EINVAL, "kernel table size/delta not set");
    if (p->num_place_types < 0 || p->num_place_types > CS_MAX_PLACE_TYPES) return set_err(e, CS_EINVAL, "bad number of place types");
    if (p->num_holidays < 0 || p->num_holidays > CS_MAX_HOLIDAYS) return set_err(e, CS_EINVAL, "bad number of holidays");
    if (e->have_pop && p->nch != e->cxy_nch)
        return set_err(e, CS_ESTATE, "nch changed to %d with a population of %d cells high loaded (set the population again)", p->nch, e->cxy_nch);
    drop_graph(e);  // a captured step follows the host's parameters (holiday launches)
    bool new_tables = !e->have_params || e->P.kernel_size != p->kernel_size || e->P.kernel_type != p->kernel_type ||
                      e->P.kernel_delta != p->kernel_delta || e->P.kernel_shape != p->kernel_shape ||
                      e->P.kernel_scale != p->kernel_scale || e->P.kernel_p3 != p->kernel_p3 ||
                      e->P.kernel_p4 != p->kernel_p4 || e->P.kernel_expansion != p->kernel_expansion || kl || kh;
    e->P = *p;
    if (!e->dP) HIPCHK(e, hipMalloc(&e->dP, sizeof(cs_params)));
    HIPCHK(e, hipMemcpy(e->dP, p, sizeof(cs_params), hipMemcpyHostToDevice));
    if (new_tables) {
        size_t n = (size_t)p->kernel_size + 1;
        std::vector<double> hl, hh;
        if (!kl || !kh) {
            hl.resize(n);
            hh.resize(n);
            for (size_t i = 0; i < n; i++) {
                hl[i] = cs::kernel_fn(*p, (int)i * p->kernel_delta) / 1.0;
                hh[i] = cs::kernel_fn(*p, (int)i * p->kernel_delta / p->kernel_expansion) / 1.0;
            }
        }
        if (e->dkl) (void)hipFree(e->dkl);
        if (e->dkh) (void)hipFree(e->dkh);
        HIPCHK(e, hipMalloc(&e->dkl, n * sizeof(double)));
        HIPCHK(e, hipMalloc(&e->dkh, n * sizeof(double)));
        HIPCHK(e, hipMemcpy(e->dkl, kl ? kl : hl.data(), n * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(e, hipMemcpy(e->dkh, kh ? kh : hh.data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    e->d.P = e->dP;
    e->d.kl = e->dkl;
    e->d.kh = e->dkh;
    e->have_params = true;
    {   // latent due lists need every latent time within CS_DUE_RING steps: the
        // longest latent period is latent_period * -ln(smallest table entry)
        double cmin = 1.0;
        for (int i = 0; i <= CS_CDF_RES; i++) cmin = std::min(cmin, p->icdf[CS_ICDF_LATENT][i]);
        const double steps = p->latent_period * -std::log(cmin) * p->time_steps_per_day;
        e->due_ok = !p->do_latent || (cmin > 0 && steps + 3.0 < (double)CS_DUE_RING);
        if (!e->due_ok) e->d.due_on = 0;
    }
    if (e->have_pop && new_tables && e->mt_buf) {
        cs::k_build_max_trans<<<grid_for((long long)e->NC * e->NC, 256, 1u << 30), 256, 0, e->stream>>>(e->d, e->mt_buf);
        HIPCHK(e, hipGetLastError());
        int rc = check_flags(e);
        if (rc) return rc;
    }
    return 0;
}

static bool no_place_links_host(const int32_t* lk, int npt) { return cs::no_place_links(lk, npt); }

int cs_set_population(cs_engine* e, const cs_population* pop)
{
    if (!e || !pop) return CS_EINVAL;
    if (!e->have_params) return set_err(e, CS_ESTATE, "cs_set_params must come first");
    HIPCHK(e, hipSetDevice(e->device));
    if (pop->n_people <= 0 || pop->n_cells <= 0) return set_err(e, CS_EINVAL, "empty population");
    if ((uint64_t)pop->n_people * CS_MAX_PLACE_TYPES > 0xffffffffull)
        return set_err(e, CS_EINVAL, "%d persons: the place-link index (persons x place types) must fit 32 bits",
                       pop->n_people);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    drop_graph(e);
    free_list(e->real_allocs);
    e->R = 0;
    // keep the flags/work words, drop the rest of the old population
    std::vector<void*> keep = {e->d.status, e->d.wc, e->d.sc, e->d.clock};
    for (void* q : e->pop_allocs)
        if (std::find(keep.begin(), keep.end(), q) == keep.end()) (void)hipFree(q);
    e->pop_allocs = keep;
    Dev& d = e->d;
    const int N = pop->n_people, NC = pop->n_cells;
    e->N = d.N = N; e->NH = d.NH = pop->n_households; e->NC = d.NC = NC;
    e->NCT = pop->n_cells_total; e->NM = d.NM = pop->n_mcells; e->NA = pop->n_adunits;
    e->NPT = d.NPT = e->P.num_place_types;
    // host-side checks of every index the kernels follow
    for (int j = 0; j <= NC; j++)
        if (pop->cell_member_start[j] < 0 || pop->cell_member_start[j] > N || (j && pop->cell_member_start[j] < pop->cell_member_start[j - 1]))
            return set_err(e, CS_EINVAL, "cell_member_start not monotone within [0, n_people]");
    if (pop->cell_member_start[NC] != N) return set_err(e, CS_EINVAL, "cells do not hold every person");
    std::vector<int32_t> map(std::max(1, e->NCT), -1), pc(N);
    for (int j = 0; j < NC; j++) {
        if (pop->cell_index[j] < 0 || pop->cell_index[j] >= e->NCT) return set_err(e, CS_EINVAL, "cell_index out of range");
        if (pop->cell_n[j] != pop->cell_member_start[j + 1] - pop->cell_member_start[j])
            return set_err(e, CS_EINVAL, "cell_n disagrees with cell_member_start");
        map[pop->cell_index[j]] = j;
    }
    for (int i = 0; i < N; i++) {
        int c = (pop->pcell[i] >= 0 && pop->pcell[i] < e->NCT) ? map[pop->pcell[i]] : -1;
        if (c < 0) return set_err(e, CS_EINVAL, "person %d in an unpopulated or unknown cell", i);
        pc[i] = c;
        if (pop->hh[i] < 0 || pop->hh[i] >= e->NH) return set_err(e, CS_EINVAL, "person %d household out of range", i);
        if (pop->mcell[i] < 0 || pop->mcell[i] >= e->NM) return set_err(e, CS_EINVAL, "person %d microcell out of range", i);
    }
    for (int j = 0; j < N; j++) {
        int k = pop->cell_members[j];
        if (k < 0 || k >= N) return set_err(e, CS_EINVAL, "cell member out of range");
    }
    for (int h = 0; h < e->NH; h++)
        if (pop->hh_first[h] < 0 || pop->hh_first[h] + pop->hh_nh[h] > N || pop->hh_nh[h] > CS_MAX_HOUSEHOLD_SIZE || pop->hh_nhr[h] < 1 ||
            pop->hh_nhr[h] > CS_MAX_HOUSEHOLD_SIZE)
            return set_err(e, CS_EINVAL, "household %d out of range", h);
    e->flags0.assign(N, 0);
    e->h_age.assign(pop->age, pop->age + N);
    e->h_flags.assign(N, 0);
    for (int i = 0; i < N; i++) {  // the person flags with the household position (cs::PF_*)
        const int h = pop->hh[i], pos = i - pop->hh_first[h];
        if (pos < 0 || pos >= pop->hh_nh[h]) return set_err(e, CS_EINVAL, "person %d lies outside its household's member range", i);
        e->h_flags[i] = (uint8_t)((pop->flags[i] & cs::PF_INPUT) | (pos << cs::PF_HPOS_SHIFT) |
                                  (pop->hh_nhr[h] != pop->hh_nh[h] ? cs::PF_NHR_DIFFERS : 0));
    }
    for (int i = 0; i < N; i++)
        e->flags0[i] = (uint8_t)((no_place_links_host(pop->place_links + (size_t)i * CS_MAX_PLACE_TYPES, e->NPT) ? cs::HOT_NO_LINKS : 0) |
                                 (pop->hh_nh[pop->hh[i]] << cs::HOT_NH_SHIFT));
    std::vector<int32_t> padunit(N, 0);
    e->adunits_valid = pop->mcell_adunit != nullptr;
    for (int i = 0; i < N && e->adunits_valid; i++) {
        padunit[i] = pop->mcell_adunit[pop->mcell[i]];
        if (padunit[i] < 0 || padunit[i] >= e->NA) e->adunits_valid = false;
    }
    if (!e->adunits_valid) std::fill(padunit.begin(), padunit.end(), 0);
    std::vector<int32_t> mstart(e->NM + 1, 0), mmem(N);
    for (int i = 0; i < N; i++) mstart[pop->mcell[i] + 1]++;
    for (int m = 0; m < e->NM; m++) mstart[m + 1] += mstart[m];
    {
        std::vector<int32_t> cur(mstart.begin(), mstart.end() - 1);
        for (int i = 0; i < N; i++) mmem[cur[pop->mcell[i]]++] = i;
    }
    int rc;
    auto& L = e->pop_allocs;
    {
        std::vector<cs::PStat> ps(N);
        for (int i = 0; i < N; i++)
            ps[i] = cs::PStat{pop->hh[i], pop->age[i], e->h_flags[i], 0,
                              make_float2(pop->hh_loc[2 * (size_t)pop->hh[i]], pop->hh_loc[2 * (size_t)pop->hh[i] + 1])};
        std::vector<cs::HRec> hr(e->NH);
        for (int h = 0; h < e->NH; h++) hr[h] = cs::HRec{pop->hh_first[h], pop->hh_nh[h], pop->hh_nhr[h]};
        if ((rc = dupload(e, L, &d.pst, ps.data(), N))) return rc;
        if ((rc = dupload(e, L, &d.hrec, hr.data(), e->NH))) return rc;
        d.hh.base = d.age.base = d.pflags.base = d.pst;
        d.ploc.base = d.pst;
        d.hfirst.base = d.hnh.base = d.hnhr.base = d.hrec;
    }
    if ((rc = dupload(e, L, &d.pc, pc.data(), N))) return rc;
    if ((rc = dupload(e, L, &d.mcell, pop->mcell, N))) return rc;
    if ((rc = dupload(e, L, &d.padunit, padunit.data(), N))) return rc;
    {
        std::vector<cs::PLinks> lk(N);
        for (int i = 0; i < N; i++)
            for (int t = 0; t < CS_MAX_PLACE_TYPES; t++) {
                lk[i].lk[t] = pop->place_links[(size_t)i * CS_MAX_PLACE_TYPES + t];
                lk[i].g[t] = pop->place_group_links[(size_t)i * CS_MAX_PLACE_TYPES + t];
            }
        if ((rc = dupload(e, L, &d.plk, lk.data(), N))) return rc;
        d.plinks.base = d.pglinks.base = d.plk;
    }
    if ((rc = dupload(e, L, &d.cstart, pop->cell_member_start, NC + 1))) return rc;
    if ((rc = dupload(e, L, &d.cn, pop->cell_n, NC))) return rc;
    e->max_cn = 0;
    for (int j = 0; j < NC; j++) e->max_cn = std::max(e->max_cn, (int)pop->cell_n[j]);
    if ((rc = dupload(e, L, &d.cidx, pop->cell_index, NC))) return rc;
    {   // grid coordinates of the populated cells (dist2_cc_min without divisions)
        std::vector<int2> xy(NC);
        for (int j = 0; j < NC; j++) xy[j] = make_int2(pop->cell_index[j] / e->P.nch, pop->cell_index[j] % e->P.nch);
        if ((rc = dupload(e, L, &d.cxy, xy.data(), NC))) return rc;
        e->cxy_nch = e->P.nch;
    }
    if ((rc = dupload(e, L, &d.cmem, pop->cell_members, N))) return rc;
    if ((rc = dupload(e, L, &d.mstart, mstart.data(), e->NM + 1))) return rc;
    if ((rc = dupload(e, L, &d.mmem, mmem.data(), N))) return rc;
    int off = 0;
    for (int t = 0; t < CS_MAX_PLACE_TYPES; t++) {
        d.np_off[t] = off;
        int np = (t < e->NPT) ? pop->n_places[t] : 0;
        e->np[t] = d.np[t] = np;
        off += np;
        if (t >= e->NPT) {
            d.pl_rec[t] = d.pl_n[t].base = d.pl_ms[t].base = d.pl_gs0[t].base = nullptr;
            d.pl_grp[t] = d.pl_gst[t].base = d.pl_gsz[t].base = nullptr;
            d.pl_mem[t] = nullptr;
            continue;
        }
        const int32_t* ms = pop->place_member_start[t];
        const int32_t* gs0 = pop->place_groups_start[t];
        if (np < 0 || (np && (!ms || !gs0 || ms[0] != 0 || gs0[0] != 0))) return set_err(e, CS_EINVAL, "place arrays of type %d", t);
        for (int k = 0; k < np; k++) {
            // a place may have no members (StratifyPlaces leaves it without
            // groups, src/SetupModel.cpp:2518); no link points to it
            if (ms[k + 1] - ms[k] != pop->place_n[t][k] || pop->place_n[t][k] < 0)
                return set_err(e, CS_EINVAL, "place %d of type %d: member range and n disagree", k, t);
            for (int g = gs0[k]; g < gs0[k + 1]; g++)
                if (pop->place_group_start[t][g] < 0 || pop->place_group_size[t][g] < 0 ||
                    pop->place_group_start[t][g] + pop->place_group_size[t][g] > pop->place_n[t][k])
                    return set_err(e, CS_EINVAL, "group %d of place %d type %d outside the member list", g, k, t);
        }
        for (int q = 0; q < (np ? ms[np] : 0); q++)
            if (pop->place_members[t][q] < 0 || pop->place_members[t][q] >= N) return set_err(e, CS_EINVAL, "place member out of range");
        int ng = np ? gs0[np] : 0;
        std::vector<cs::PlRec> pr(np + 1);
        for (int k = 0; k <= np; k++) pr[k] = cs::PlRec{ms[k], gs0[k], k < np ? pop->place_n[t][k] : 0, 0};
        std::vector<cs::GRec> gr(ng);
        for (int g = 0; g < ng; g++) gr[g] = cs::GRec{pop->place_group_start[t][g], pop->place_group_size[t][g]};
        if ((rc = dupload(e, L, &d.pl_rec[t], pr.data(), np + 1))) return rc;
        if ((rc = dupload(e, L, &d.pl_grp[t], gr.data(), ng))) return rc;
        if ((rc = dupload(e, L, &d.pl_mem[t], pop->place_members[t], np ? ms[np] : 0))) return rc;
        d.pl_n[t].base = d.pl_ms[t].base = d.pl_gs0[t].base = d.pl_rec[t];
        d.pl_gst[t].base = d.pl_gsz[t].base = d.pl_grp[t];
    }
    d.np_off[CS_MAX_PLACE_TYPES] = off;
    if ((rc = dupload(e, L, &d.np_off_d, (const int32_t*)d.np_off, CS_MAX_PLACE_TYPES + 1))) return rc;
    {   // the places of each microcell (Microcell.places), by type then index
        std::vector<int32_t> ms(e->NM + 1, 0), ml(std::max(off, 1));
        for (int t = 0; t < e->NPT; t++)
            for (int k = 0; k < e->np[t]; k++) {
                const int m = pop->place_mcell[t] ? pop->place_mcell[t][k] : -1;
                if (m < 0 || m >= e->NM) return set_err(e, CS_EINVAL, "place %d of type %d: microcell out of range", k, t);
                if (k >= (1 << 27)) return set_err(e, CS_EINVAL, "more than 2^27 places of type %d", t);
                ms[m + 1]++;
            }
        for (int m = 0; m < e->NM; m++) ms[m + 1] += ms[m];
        std::vector<int32_t> cur(ms.begin(), ms.end() - 1);
        for (int t = 0; t < e->NPT; t++)
            for (int k = 0; k < e->np[t]; k++) ml[cur[pop->place_mcell[t][k]]++] = (t << 27) | k;
        if ((rc = dupload(e, L, &d.mpl_start, ms.data(), e->NM + 1))) return rc;
        if ((rc = dupload(e, L, &d.mpl, ml.data(), std::max(off, 1)))) return rc;
    }
    for (int i = 0; i < N; i++)
        for (int t = 0; t < e->NPT; t++) {
            int lk = pop->place_links[(size_t)i * CS_MAX_PLACE_TYPES + t];
            if (lk >= e->np[t]) return set_err(e, CS_EINVAL, "person %d links to place %d of type %d beyond %d", i, lk, t, e->np[t]);
            if (lk >= 0) {
                int g = pop->place_group_links[(size_t)i * CS_MAX_PLACE_TYPES + t];
                if (g >= pop->place_groups_start[t][lk + 1] - pop->place_groups_start[t][lk])
                    return set_err(e, CS_EINVAL, "person %d group link beyond the place's groups", i);
            }
        }
    // max_trans is built when full tables are chosen (cs_alloc_realisations)
    if (e->mt_buf) (void)hipFree(e->mt_buf);
    e->mt_buf = nullptr;
    d.max_trans = nullptr;
    d.ck = d.ck_nb = 0;
    d.ck_pre = nullptr;
    d.cum_trans = nullptr;
    d.invcdf = nullptr;
    HIPCHK(e, hipMemset(d.status, 0, sizeof(unsigned)));
    e->have_pop = true;
    return check_flags(e);
}

int cs_alloc_realisations(cs_engine* e, int n_real, const uint64_t* seeds)
{
    if (!e || n_real <= 0 || !seeds) return CS_EINVAL;
    if (!e->have_pop) return set_err(e, CS_ESTATE, "cs_set_population must come first");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    drop_graph(e);
    free_list(e->real_allocs);
    Dev& d = e->d;
    const size_t R = (size_t)n_real, N = (size_t)e->N, NC = (size_t)e->NC, NM = (size_t)e->NM;
    const size_t RN = R * N, RC = R * NC, PL = R * (size_t)d.np_off[CS_MAX_PLACE_TYPES];
    // the per-(realisation, microcell) and per-(realisation, place) passes
    // index in 32 bits
    if (R * NM >= ((size_t)1 << 32) || PL >= ((size_t)1 << 32))
        return set_err(e, CS_EUNSUPPORTED, "%d realisations x %zu microcells (or places) exceed 2^32", n_real, NM);
    e->R = d.R = n_real;
    auto& A = e->real_allocs;
    int rc = 0;
#define AL(f, n) if ((rc = dalloc(e, A, &d.f, (n)))) return rc
    uint64_t* sd;
    if ((rc = dalloc(e, A, &sd, R))) return rc;
    HIPCHK(e, hipMemcpy(sd, seeds, R * 8, hipMemcpyHostToDevice));
    d.seed = sd;
    AL(hot, RN);
    d.bind_hot(d.hot);
    AL(shot, RN);
    AL(gate, RN);
    AL(tim, RN);
    d.bind_tim(d.tim);
    AL(qc, RN); AL(dct, RN);
    AL(listpos, RN); AL(infector, RN); AL(susc, RN); AL(prob_care, RN);
    AL(t_ts, RN); AL(t_te, RN); AL(t_vs, RN); AL(t_is, RN); AL(t_qs, RN);
    AL(cS, RC); AL(cL, RC); AL(cI, RC); AL(cR, RC); AL(cD, RC); AL(cumTC, RC); AL(S0, RC);
    AL(tot_prob, RC); AL(perm, RN);
    {   // destination tables: full NC x NC rows while they fit, else checkpoints
        int stride = e->dest_stride;
        if (stride < 0 && (R + 1) * NC * NC * 4 <= CS_FULL_TABLE_BYTES) stride = 0;
        const double inv_bytes = (double)RC * 1025 * 4;
        const bool index = e->dest_index && inv_bytes <= (double)CS_CK_INDEX_BYTES;
        if (stride < 0)
            for (stride = 8; (double)R * NC * NC * 4 / stride + (index ? inv_bytes : 0.0) > (double)CS_CK_TABLE_BYTES;
                 stride *= 2) {}
        d.ck = stride;
        d.ck_nb = stride ? (int)((NC + stride - 1) / stride) : 0;
        d.cum_trans = nullptr; d.invcdf = nullptr; d.ck_pre = nullptr; d.ck_inv = nullptr;
        if (stride) {
            AL(ck_pre, RC * d.ck_nb);
            if (index) AL(ck_inv, RC * 1025);
        } else {
            AL(cum_trans, RC * NC); AL(invcdf, RC * 1025);
            if ((rc = ensure_max_trans(e))) return rc;
        }
    }
    AL(pl_cs, PL); AL(pl_ce, PL); AL(m_socdist, R * NM); AL(m_moverest, R * NM); AL(m_trig, R * NM * 6);
    AL(ist, R); AL(trig_hist, R * CS_TRIG_RING * 3); AL(mc_state, R * NM); AL(mc_end, R * NM); AL(pl_once, PL); AL(itv_n, 4);
    d.det = nullptr; d.close_l = nullptr; d.det_cap = d.close_cap = 0;  // sized when interventions first run (ensure_itv)
    AL(cnt, R); AL(prev_prop, R); AL(samp_prop, R); AL(loc_init, 2 * R);
    AL(si_buf, RN); AL(ci_buf, RN); AL(spK, RC);
    d.seg_cap = (unsigned)std::min<size_t>(RN / 64 + RC + 1, 0xffffffffu);
    d.draw_cap = d.seg_cap;
    AL(segl, d.seg_cap); AL(drawl, d.draw_cap); AL(cells, RC); AL(wl, cs::WL_N);
    size_t cap = std::min<size_t>(std::max<size_t>(RN / 16, (size_t)1 << 20), (size_t)1 << 28);
    cap -= cap % CS_EV_SHARDS;
    d.ev_cap = (unsigned)cap;
    AL(ev_key, cap); AL(ev_gp, cap); AL(ev_infector, cap); AL(ev_type, cap); AL(ev_count, CS_EV_SHARDS * CS_EV_STRIDE);
    AL(pc_count, CS_EV_SHARDS * CS_EV_STRIDE);
    d.pcl = nullptr;
    d.pc_cap = 0;
    AL(clr, RN); AL(batch, RN); AL(bcount, RC);
    d.mv_words = (e->max_cn + 31) / 32 + 1;
    AL(mv_over, RC); AL(mv_over_n, 2); AL(mv_bm, (size_t)CS_MOVE_BIG_BLOCKS * 2 * d.mv_words);
    d.move_cap = e->move_cap;
    d.bind_clr(d.clr);
    AL(care_gp, RN); AL(care_k, RN); AL(care_abs, RN); AL(care_count, 1);
    AL(last_tot, CS_TS_LAST * R);
    AL(cstage, (size_t)CS_TALLY_COPIES * R);
    d.due_scap = (unsigned)std::min<size_t>(std::max<size_t>(RN / 64, 4096) / CS_DUE_SHARDS + 1, 0x7fffffff);
    AL(due, (size_t)CS_DUE_RING * CS_DUE_SHARDS * d.due_scap); AL(due_n, CS_DUE_RING * CS_DUE_SHARDS * CS_DUE_STRIDE); AL(due_ovf, CS_DUE_RING);
    d.due_on = 0;
#undef AL
    if ((rc = size_contacts(e))) return rc;
    if ((rc = dalloc(e, A, &e->d_sel, R))) return rc;
    if ((rc = dalloc(e, A, &e->rbeta_buf, R))) return rc;
    if ((rc = dalloc(e, A, &e->rseed_buf, R))) return rc;
    if ((rc = dalloc(e, A, &e->rthr_buf, R))) return rc;
    d.rbeta = nullptr;
    d.rseed_n = nullptr;
    d.rthr = nullptr;
    cs::k_claim_init<<<grid_for((long long)RN, 256), 256, 0, e->stream>>>(d.clr, RN);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipMemset(d.bcount, 0, RC * 4));
    HIPCHK(e, hipMemset(d.mv_over_n, 0, 2 * 4));
    HIPCHK(e, hipMemset(d.wl, 0, cs::WL_N * 4));
    HIPCHK(e, hipMemset(d.cnt, 0, R * sizeof(cs::Counters)));
    HIPCHK(e, hipMemset(d.cstage, 0, (size_t)CS_TALLY_COPIES * R * sizeof(cs::Counters)));
    {   // counters and samples by administrative unit, when the units are known
        const bool units = e->P.do_ad_units && e->adunits_valid && e->NA > 0;
        d.NA = units ? e->NA : 0;
        d.ad_cnt = nullptr;
        if (units) {
            unsigned long long* ac;
            if ((rc = dalloc(e, A, &ac, (size_t)CS_AD_COPIES * R * d.NA * CS_TSU_FIELDS))) return rc;
            HIPCHK(e, hipMemset(ac, 0, (size_t)CS_AD_COPIES * R * d.NA * CS_TSU_FIELDS * 8));
            d.ad_cnt = ac;
        }
        d.ext_w = CS_TSA_FIELDS * CS_NUM_AGE_GROUPS + CS_TSU_FIELDS * d.NA;
        long long* le;
        if ((rc = dalloc(e, A, &le, R * (size_t)(CS_TSA_CUM * CS_NUM_AGE_GROUPS + CS_TSU_FIELDS * d.NA)))) return rc;
        d.last_ext = le;
        d.ts_ext = nullptr;
    }
    HIPCHK(e, hipMemset(d.m_trig, 0, R * NM * 6 * 4));
    HIPCHK(e, hipMemset(d.itv_n, 0, 4 * sizeof(unsigned)));
    cs::k_reset_interventions<<<grid_for(d.R, 64), 64, 0, e->stream>>>(d);
    d.ts_buf = nullptr;
    d.ts_max = 0;
    e->ts_n = 0;
    e->ts_fresh = true;
    return 0;
}

int cs_set_realisation_params(cs_engine* e, const double* local_beta, const int32_t* n_seed)
{
    if (!e) return CS_EINVAL;
    if (e->R == 0) return set_err(e, CS_ESTATE, "no realisations allocated");
    for (int r = 0; r < e->R; r++) {
        if (local_beta && !(local_beta[r] >= 0)) return set_err(e, CS_EINVAL, "local_beta[%d] must be >= 0", r);
        if (n_seed && n_seed[r] < 0) return set_err(e, CS_EINVAL, "n_seed[%d] must be >= 0", r);
    }
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    Dev& d = e->d;
    if (local_beta) HIPCHK(e, hipMemcpy(e->rbeta_buf, local_beta, (size_t)e->R * 8, hipMemcpyHostToDevice));
    if (n_seed) HIPCHK(e, hipMemcpy(e->rseed_buf, n_seed, (size_t)e->R * 4, hipMemcpyHostToDevice));
    d.rbeta = local_beta ? e->rbeta_buf : nullptr;
    d.rseed_n = n_seed ? e->rseed_buf : nullptr;
    drop_graph(e);  // the captured launches hold the old pointers
    return 0;
}

int cs_set_realisation_triggers(cs_engine* e, const double* threshold)
{
    if (!e) return CS_EINVAL;
    if (e->R == 0) return set_err(e, CS_ESTATE, "no realisations allocated");
    for (int r = 0; threshold && r < e->R; r++)
        if (!(threshold[r] >= 0)) return set_err(e, CS_EINVAL, "threshold[%d] must be >= 0", r);
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    if (threshold) HIPCHK(e, hipMemcpy(e->rthr_buf, threshold, (size_t)e->R * 8, hipMemcpyHostToDevice));
    e->d.rthr = threshold ? e->rthr_buf : nullptr;
    drop_graph(e);  // the captured launches hold the old pointer
    return 0;
}

int cs_reset_realisations(cs_engine* e, int n_seed)
{
    if (!e) return CS_EINVAL;
    if (e->R == 0) return set_err(e, CS_ESTATE, "no realisations allocated");
    int rc = check_path(e);
    if (rc) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    Dev& d = e->d;
    hipStream_t st = e->stream;
    e->ts_fresh = true;
    HIPCHK(e, hipMemsetAsync(d.cnt, 0, (size_t)d.R * sizeof(cs::Counters), st));
    if (d.ad_cnt) HIPCHK(e, hipMemsetAsync(d.ad_cnt, 0, (size_t)CS_AD_COPIES * d.R * d.NA * CS_TSU_FIELDS * 8, st));
    HIPCHK(e, hipMemsetAsync(d.due_n, 0, CS_DUE_RING * CS_DUE_SHARDS * CS_DUE_STRIDE * sizeof(unsigned), st));
    HIPCHK(e, hipMemsetAsync(d.due_ovf, 0, CS_DUE_RING * sizeof(unsigned), st));
    if (d.due_on != (e->due_ok ? 1 : 0)) {  // fresh state: every latent person is listed from here on
        HIPCHK(e, hipStreamSynchronize(st));
        drop_graph(e);
        d.due_on = e->due_ok ? 1 : 0;
    }
    cs::k_reset_persons<<<grid_for((long long)d.R * d.N, 256), 256, 0, st>>>(d);
    cs::k_reset_cells<<<grid_for((long long)d.R * d.NC, 64), 64, 0, st>>>(d);
    cs::k_reset_misc<<<grid_for(std::max((long long)d.R * d.np_off[CS_MAX_PLACE_TYPES], (long long)d.R * d.NM), 256), 256, 0, st>>>(d);
    cs::k_reset_interventions<<<grid_for(d.R, 64), 64, 0, st>>>(d);
    if ((rc = ensure_itv(e))) return rc;
    if ((rc = launch_update_probs(e, nullptr))) return rc;
    cs::k_zero_counts<<<1, 64, 0, st>>>(d.wl, cs::WL_N);
    cs::k_seed<<<grid_for(d.R, 64), 64, 0, st>>>(d, n_seed);
    launch_list_move(e, st, 0);
    HIPCHK(e, hipGetLastError());
    return check_flags(e);
}

int cs_load_state(cs_engine* e, int real, const cs_state* s)
{
    if (!e || !s) return CS_EINVAL;
    if (real < 0 || real >= e->R) return set_err(e, CS_EINVAL, "realisation %d out of range", real);
    HIPCHK(e, hipSetDevice(e->device));
    Dev& d = e->d;
    const size_t N = e->N, NC = e->NC, o = (size_t)real * N, oc = (size_t)real * NC;
    e->ts_fresh = true;
    if (d.due_on) {  // loaded latent persons are on no due list: scan them all until the next reset
        HIPCHK(e, hipStreamSynchronize(e->stream));
        drop_graph(e);
        d.due_on = 0;
    }
    {
        // the packed fields: read the records, overwrite the fields the state gives, write back
        std::vector<cs::Hot> h(N);
        HIPCHK(e, hipMemcpy(h.data(), d.hot + o, N * sizeof(cs::Hot), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; i++) {
            cs::Hot& x = h[i];
            if (s->inf) x.inf = s->inf[i];
            if (s->infect_type) x.itype = s->infect_type[i];
            if (s->infectiousness) x.infness = s->infectiousness[i];
            if (s->latent_time) x.t_lat = s->latent_time[i];
            if (s->absent_start_time) x.t_as = s->absent_start_time[i];
            if (s->absent_stop_time) x.t_ae = s->absent_stop_time[i];
            x.flags = e->flags0[i];  // and every household
            x.age = e->h_age[i];
            x.pflags = e->h_flags[i];
        }
        HIPCHK(e, hipMemcpy(d.hot + o, h.data(), N * sizeof(cs::Hot), hipMemcpyHostToDevice));
        HIPCHK(e, hipMemset(d.gate + o, 0, N * sizeof(uint16_t)));  // loaded timers: look at every one
        std::vector<cs::Tim> m(N);
        HIPCHK(e, hipMemcpy(m.data(), d.tim + o, N * sizeof(cs::Tim), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; i++) {
            cs::Tim& x = m[i];
            if (s->infection_time) x.t_inf = s->infection_time[i];
            if (s->recovery_or_death_time) x.t_rod = s->recovery_or_death_time[i];
            if (s->sari_time) x.t_sari = s->sari_time[i];
            if (s->critical_time) x.t_crit = s->critical_time[i];
            if (s->stepdown_time) x.t_step = s->stepdown_time[i];
            if (s->detected_time) x.t_det = s->detected_time[i];
            if (s->severity_current) x.sev_cur = s->severity_current[i];
            if (s->severity_final) x.sev_final = s->severity_final[i];
            if (s->bits) x.bits = s->bits[i];
        }
        HIPCHK(e, hipMemcpy(d.tim + o, m.data(), N * sizeof(cs::Tim), hipMemcpyHostToDevice));
    }
#define UP(dst, src, off, n) if (src) HIPCHK(e, hipMemcpy(d.dst + (off), src, (n) * sizeof(*d.dst), hipMemcpyHostToDevice))
    UP(listpos, s->listpos, o, N); UP(infector, s->infector, o, N);
    UP(susc, s->susc, o, N); UP(prob_care, s->prob_care, o, N);
    UP(t_ts, s->treat_start_time, o, N); UP(t_te, s->treat_stop_time, o, N); UP(t_vs, s->vacc_start_time, o, N);
    UP(t_is, s->isolation_start_time, o, N); UP(t_qs, s->quar_start_time, o, N); UP(qc, s->quar_comply, o, N);
    UP(dct, s->dct_traced, o, N);
    UP(cS, s->cell_S, oc, NC); UP(cL, s->cell_L, oc, NC); UP(cI, s->cell_I, oc, NC); UP(cR, s->cell_R, oc, NC);
    UP(cD, s->cell_D, oc, NC); UP(cumTC, s->cell_cumTC, oc, NC); UP(perm, s->cell_perm, o, N);
    const size_t tot = d.np_off[CS_MAX_PLACE_TYPES];
    for (int t = 0; t < e->NPT; t++) {
        UP(pl_cs, s->place_close_start[t], real * tot + d.np_off[t], (size_t)e->np[t]);
        UP(pl_ce, s->place_close_end[t], real * tot + d.np_off[t], (size_t)e->np[t]);
    }
    UP(m_socdist, s->mcell_socdist, (size_t)real * e->NM, (size_t)e->NM);
    UP(m_moverest, s->mcell_moverest, (size_t)real * e->NM, (size_t)e->NM);
#undef UP
    // derived per-realisation values, as a host run would hold them
    std::vector<int32_t> S(NC);
    HIPCHK(e, hipMemcpy(S.data(), d.cS + oc, NC * 4, hipMemcpyDeviceToHost));
    long long st = 0;
    for (int32_t v : S) st += v;
    double prev = (double)st / (double)N;
    double loc[2] = {e->P.location_initial_infection[0], e->P.location_initial_infection[1]};
    HIPCHK(e, hipMemcpy(d.prev_prop + real, &prev, 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(d.loc_init + 2 * real, loc, 16, hipMemcpyHostToDevice));
    std::vector<int> sel(e->R, 0);
    sel[real] = 1;
    HIPCHK(e, hipMemcpy(e->d_sel, sel.data(), e->R * sizeof(int), hipMemcpyHostToDevice));
    int rc = launch_update_probs(e, e->d_sel);
    if (rc) return rc;
    cs::k_shot_build<<<grid_for(N, 256), 256, 0, e->stream>>>(d, real);  // the infectious persons' slot copies
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return 0;
}

int cs_store_state(cs_engine* e, int real, cs_state* s)
{
    if (!e || !s) return CS_EINVAL;
    if (real < 0 || real >= e->R) return set_err(e, CS_EINVAL, "realisation %d out of range", real);
    HIPCHK(e, hipSetDevice(e->device));
    Dev& d = e->d;
    const size_t N = e->N, NC = e->NC, o = (size_t)real * N, oc = (size_t)real * NC;
    const size_t tot = d.np_off[CS_MAX_PLACE_TYPES];
    // every array the caller asked for is copied into one pinned staging
    // buffer (one asynchronous copy each, one wait), then unpacked: a
    // drop-in caller stores the state after every timestep
    struct Part { const void* src; void* dst; size_t bytes, off; };
    std::vector<Part> parts;
    size_t need = 0;
    auto add = [&](const void* src, void* dst, size_t bytes) {
        parts.push_back(Part{src, dst, bytes, need});
        need += (bytes + 255) & ~(size_t)255;
    };
    add(d.hot + o, nullptr, N * sizeof(cs::Hot));
    add(d.tim + o, nullptr, N * sizeof(cs::Tim));
#define DN(src, dst, off, n) if (dst) add(d.src + (off), dst, (n) * sizeof(*d.src))
    DN(listpos, s->listpos, o, N); DN(infector, s->infector, o, N);
    DN(susc, s->susc, o, N); DN(prob_care, s->prob_care, o, N);
    DN(t_ts, s->treat_start_time, o, N); DN(t_te, s->treat_stop_time, o, N); DN(t_vs, s->vacc_start_time, o, N);
    DN(t_is, s->isolation_start_time, o, N); DN(t_qs, s->quar_start_time, o, N); DN(qc, s->quar_comply, o, N);
    DN(dct, s->dct_traced, o, N);
    DN(cS, s->cell_S, oc, NC); DN(cL, s->cell_L, oc, NC); DN(cI, s->cell_I, oc, NC); DN(cR, s->cell_R, oc, NC);
    DN(cD, s->cell_D, oc, NC); DN(cumTC, s->cell_cumTC, oc, NC); DN(perm, s->cell_perm, o, N);
    for (int t = 0; t < e->NPT; t++) {
        DN(pl_cs, s->place_close_start[t], real * tot + d.np_off[t], (size_t)e->np[t]);
        DN(pl_ce, s->place_close_end[t], real * tot + d.np_off[t], (size_t)e->np[t]);
    }
    DN(m_socdist, s->mcell_socdist, (size_t)real * e->NM, (size_t)e->NM);
    DN(m_moverest, s->mcell_moverest, (size_t)real * e->NM, (size_t)e->NM);
#undef DN
    if (need > e->pin_cap) {
        if (e->pin) HIPCHK(e, hipHostFree(e->pin));
        e->pin = nullptr;
        e->pin_cap = 0;
        HIPCHK(e, hipHostMalloc((void**)&e->pin, need, hipHostMallocDefault));
        e->pin_cap = need;
    }
    for (const Part& q : parts)
        if (q.bytes) HIPCHK(e, hipMemcpyAsync(e->pin + q.off, q.src, q.bytes, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    const cs::Hot* h = reinterpret_cast<const cs::Hot*>(e->pin + parts[0].off);
    for (size_t i = 0; i < N; i++) {
        const cs::Hot& x = h[i];
        if (s->inf) s->inf[i] = x.inf;
        if (s->infect_type) s->infect_type[i] = x.itype;
        if (s->infectiousness) s->infectiousness[i] = x.infness;
        if (s->latent_time) s->latent_time[i] = x.t_lat;
        if (s->absent_start_time) s->absent_start_time[i] = x.t_as;
        if (s->absent_stop_time) s->absent_stop_time[i] = x.t_ae;
    }
    const cs::Tim* m = reinterpret_cast<const cs::Tim*>(e->pin + parts[1].off);
    for (size_t i = 0; i < N; i++) {
        const cs::Tim& x = m[i];
        if (s->infection_time) s->infection_time[i] = x.t_inf;
        if (s->recovery_or_death_time) s->recovery_or_death_time[i] = x.t_rod;
        if (s->sari_time) s->sari_time[i] = x.t_sari;
        if (s->critical_time) s->critical_time[i] = x.t_crit;
        if (s->stepdown_time) s->stepdown_time[i] = x.t_step;
        if (s->detected_time) s->detected_time[i] = x.t_det;
        if (s->severity_current) s->severity_current[i] = x.sev_cur;
        if (s->severity_final) s->severity_final[i] = x.sev_final;
        if (s->bits) s->bits[i] = x.bits;
    }
    for (size_t k = 2; k < parts.size(); k++)
        if (parts[k].bytes) memcpy(parts[k].dst, e->pin + parts[k].off, parts[k].bytes);
    return 0;
}

int cs_infect_sweep(cs_engine* e, double t)
{
    if (!e) return CS_EINVAL;
    if (e->R == 0) return set_err(e, CS_ESTATE, "no realisations allocated");
    int rc = check_sweep(e);
    if (rc) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    return infect_sweep(e, t);
}

int cs_incub_recovery_sweep(cs_engine* e, double t)
{
    if (!e) return CS_EINVAL;
    if (e->R == 0) return set_err(e, CS_ESTATE, "no realisations allocated");
    int rc = check_sweep(e);
    if (rc) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    cs::k_step_begin<<<1, 64, 0, e->stream>>>(e->d, t, 0);
    if ((rc = launch_incub(e))) return rc;
    return check_flags(e);
}

int cs_update_probs(cs_engine* e)
{
    if (!e) return CS_EINVAL;
    if (e->R == 0) return set_err(e, CS_ESTATE, "no realisations allocated");
    HIPCHK(e, hipSetDevice(e->device));
    int rc = launch_update_probs(e, nullptr);
    if (rc) return rc;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return 0;
}

// a run's set-up (cs_run_steps and batches): the first sample of a new
// series, room for the run's samples, the device clock at t0
static int run_prepare(cs_engine* e, double t0, int n_steps, int sps)
{
    if (!e || n_steps < 0 || sps <= 0) return CS_EINVAL;
    if (e->R == 0) return set_err(e, CS_ESTATE, "no realisations allocated");
    int rc = check_path(e);
    if (rc) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    if ((rc = ensure_itv(e))) return rc;
    Dev& d = e->d;
    const int ns = n_steps / sps;
    if (e->ts_fresh) {  // a new series: RunModel samples before its first timestep
        e->ts_n = 0;
        cs::k_record_init<<<grid_for(d.R, 64), 64, 0, e->stream>>>(d, t0);
        e->ts_fresh = false;
    }
    if (e->ts_n + ns > d.ts_max || !d.ts_buf) {  // grow, keeping the samples taken so far
        const int cap = std::max(std::max(e->ts_n + ns, 2 * d.ts_max), 1);
        int32_t *tb, *tx;
        if ((rc = dalloc(e, e->real_allocs, &tb, (size_t)d.R * cap * CS_TS_FIELDS))) return rc;
        if ((rc = dalloc(e, e->real_allocs, &tx, (size_t)d.R * cap * d.ext_w))) return rc;
        HIPCHK(e, hipMemsetAsync(tb, 0, (size_t)d.R * cap * CS_TS_FIELDS * 4, e->stream));
        HIPCHK(e, hipMemsetAsync(tx, 0, (size_t)d.R * cap * d.ext_w * 4, e->stream));
        if (d.ts_buf && e->ts_n) {
            HIPCHK(e, hipMemcpy2DAsync(tb, (size_t)cap * CS_TS_FIELDS * 4, d.ts_buf, (size_t)d.ts_max * CS_TS_FIELDS * 4,
                                       (size_t)e->ts_n * CS_TS_FIELDS * 4, (size_t)d.R, hipMemcpyDeviceToDevice, e->stream));
            HIPCHK(e, hipMemcpy2DAsync(tx, (size_t)cap * d.ext_w * 4, d.ts_ext, (size_t)d.ts_max * d.ext_w * 4,
                                       (size_t)e->ts_n * d.ext_w * 4, (size_t)d.R, hipMemcpyDeviceToDevice, e->stream));
        }
        if (d.ts_buf) {
            HIPCHK(e, hipStreamSynchronize(e->stream));
            drop_graph(e);  // the captured record launches hold the old buffers
            for (void* old : {(void*)d.ts_buf, (void*)d.ts_ext}) {
                (void)hipFree(old);
                e->real_allocs.erase(std::remove(e->real_allocs.begin(), e->real_allocs.end(), old), e->real_allocs.end());
            }
        }
        d.ts_buf = tb;
        d.ts_ext = tx;
        d.ts_max = cap;
    }
    const cs::Clock c0 = {t0, 0, e->ts_n};
    HIPCHK(e, hipMemcpyAsync(d.clock, &c0, sizeof(c0), hipMemcpyHostToDevice, e->stream));
    return 0;
}

// after replays that ended in the halt an overflow causes (the device clock
// at c): grow the lists and finish the run's steps to `upto` directly
static int run_recover(cs_engine* e, const cs::Clock& c, int upto, int sps)
{
    int rc;
    if ((rc = grow_events(e))) return rc;
    return run_direct(e, c.t, c.step, upto, sps);
}

// the run's last steps (fewer than an output step, or all of them without
// replays) from step `done` at time t, and the series' sample count
static int run_finish(cs_engine* e, double t, int done, int n_steps, int sps)
{
    int rc;
    if ((rc = run_direct(e, t, done, n_steps, sps))) return rc;
    cs::Clock c;
    if ((rc = read_clock(e, &c))) return rc;
    e->ts_n = c.sample;
    HIPCHK(e, hipGetLastError());
    return check_flags(e);
}

int cs_run_steps(cs_engine* e, double t0, int n_steps, int sps)
{
    int rc = run_prepare(e, t0, n_steps, sps);
    if (rc) return rc;
    Dev& d = e->d;
    const int periods = n_steps / sps;
    int done = 0;
    double t = t0;
    if (e->use_graphs && !e->prof && periods > 0) {  // profiling times every launch: direct launches then
        if ((!e->gexec || e->graph_sps != sps) && (rc = capture_graph(e, sps))) return rc;
        int p = 0;
        for (int guard = 0; p < periods && guard < 64;) {
            const int batch = std::min(periods - p, 8);  // output steps between flag checks
            for (int b = 0; b < batch; b++) HIPCHK(e, hipGraphLaunch(e->gexec, e->stream));
            cs::Clock c;
            if ((rc = read_clock(e, &c))) return rc;
            unsigned f = 0;
            HIPCHK(e, hipMemcpy(&f, d.status, sizeof(f), hipMemcpyDeviceToHost));
            if (f & cs::FLAG_OVERFLOW) {
                // the replay halted at step c.step: grow the lists, finish that
                // output step directly, then replay again
                guard++;
                const int end = (c.step / sps + 1) * sps;
                if ((rc = run_recover(e, c, end, sps))) return rc;
                p = end / sps;
                if ((rc = capture_graph(e, sps))) return rc;
                continue;
            }
            if ((rc = check_flags(e))) return rc;
            p += batch;
        }
        if (p < periods) return set_err(e, CS_EOVERFLOW, "work lists keep overflowing");
        cs::Clock c;
        if ((rc = read_clock(e, &c))) return rc;
        done = c.step;
        t = c.t;
    }
    return run_finish(e, t, done, n_steps, sps);
}

int cs_set_queue_capacity(cs_engine* e, uint32_t events, uint32_t draw_chunks)
{
    if (!e) return CS_EINVAL;
    if (e->R == 0) return set_err(e, CS_ESTATE, "no realisations allocated");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    drop_graph(e);
    Dev& d = e->d;
    auto drop = [&](void* p) { e->real_allocs.erase(std::remove(e->real_allocs.begin(), e->real_allocs.end(), p), e->real_allocs.end()); };
    int rc;
    if (events) {
        if (events > 0xffffffffu - CS_EV_SHARDS) return set_err(e, CS_EINVAL, "event capacity %u too large", events);
        events += (CS_EV_SHARDS - events % CS_EV_SHARDS) % CS_EV_SHARDS;  // whole shards
        (void)hipFree(d.ev_key); (void)hipFree(d.ev_gp); (void)hipFree(d.ev_infector); (void)hipFree(d.ev_type);
        drop(d.ev_key); drop(d.ev_gp); drop(d.ev_infector); drop(d.ev_type);
        if ((rc = dalloc(e, e->real_allocs, &d.ev_key, events)) || (rc = dalloc(e, e->real_allocs, &d.ev_gp, events)) ||
            (rc = dalloc(e, e->real_allocs, &d.ev_infector, events)) || (rc = dalloc(e, e->real_allocs, &d.ev_type, events)))
            return rc;
        d.ev_cap = events;
        if ((rc = size_contacts(e))) return rc;
    }
    if (draw_chunks) {
        (void)hipFree(d.drawl);
        drop(d.drawl);
        if ((rc = dalloc(e, e->real_allocs, &d.drawl, draw_chunks))) return rc;
        d.draw_cap = draw_chunks;
    }
    return 0;
}

int cs_set_due_capacity(cs_engine* e, uint32_t per_step)
{
    if (!e) return CS_EINVAL;
    if (e->R == 0) return set_err(e, CS_ESTATE, "no realisations allocated");
    if (!per_step) return 0;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    drop_graph(e);
    Dev& d = e->d;
    (void)hipFree(d.due);
    e->real_allocs.erase(std::remove(e->real_allocs.begin(), e->real_allocs.end(), (void*)d.due), e->real_allocs.end());
    const unsigned scap = (per_step + CS_DUE_SHARDS - 1) / CS_DUE_SHARDS;
    int rc = dalloc(e, e->real_allocs, &d.due, (size_t)CS_DUE_RING * CS_DUE_SHARDS * scap);
    if (rc) return rc;
    d.due_scap = scap;
    // the lists held so far are gone: scan until the next reset lists afresh
    HIPCHK(e, hipMemset(d.due_n, 0, CS_DUE_RING * CS_DUE_SHARDS * CS_DUE_STRIDE * sizeof(unsigned)));
    HIPCHK(e, hipMemset(d.due_ovf, 0, CS_DUE_RING * sizeof(unsigned)));
    d.due_on = 0;
    return 0;
}

int cs_get_queue_capacity(cs_engine* e, uint32_t* events, uint32_t* draw_chunks)
{
    if (!e) return CS_EINVAL;
    if (events) *events = e->d.ev_cap;
    if (draw_chunks) *draw_chunks = e->d.draw_cap;
    return 0;
}

int cs_set_launch_shape(cs_engine* e, uint32_t grid_waves, uint32_t grid_cells)
{
    if (!e) return CS_EINVAL;
    if (grid_waves) e->grid_waves = grid_waves;
    if (grid_cells) e->grid_cells = grid_cells;
    drop_graph(e);
    return 0;
}

int cs_set_move_capacity(cs_engine* e, int per_cell)
{
    if (!e) return CS_EINVAL;
    if (per_cell < 0 || per_cell > CS_MOVE_CAP)
        return set_err(e, CS_EINVAL, "move capacity %d outside [0, %d]", per_cell, CS_MOVE_CAP);
    e->move_cap = e->d.move_cap = per_cell;
    drop_graph(e);
    return 0;
}

int cs_autotune(cs_engine* e, double t, int reps, cs_tuning* out)
{
    if (!e || reps <= 0) return CS_EINVAL;
    if (e->R == 0) return set_err(e, CS_ESTATE, "no realisations allocated");
    int rc = check_path(e);
    if (rc) return rc;
    HIPCHK(e, hipSetDevice(e->device));
    const bool prof = e->prof;
    e->prof = false;
    hipEvent_t a = pool_event(e), b = pool_event(e);
    // the event phase reads the state and writes only scratch, events and
    // work lists, so it can be timed repeatedly on the current state
    auto time_events = [&](float* ms_out) -> int {
        float best = 1e30f;
        for (int k = 0; k < reps; k++) {
            cs::k_step_begin<<<1, 64, 0, e->stream>>>(e->d, t, 0);
            HIPCHK(e, hipEventRecord(a, e->stream));
            int r2 = launch_infect_events(e);
            if (r2) return r2;
            HIPCHK(e, hipEventRecord(b, e->stream));
            HIPCHK(e, hipEventSynchronize(b));
            float ms = 0;
            HIPCHK(e, hipEventElapsedTime(&ms, a, b));
            best = std::min(best, ms);
            unsigned f = 0;
            HIPCHK(e, hipMemcpy(&f, e->d.status, sizeof(f), hipMemcpyDeviceToHost));
            if (f & cs::FLAG_OVERFLOW) {  // grow and measure again
                if ((r2 = grow_events(e))) return r2;
                k--;
                best = 1e30f;
            } else if ((r2 = check_flags(e))) return r2;
        }
        *ms_out = best;
        return 0;
    };
    const uint32_t waves[] = {1024, 2048, 4096, 8192, 16384};
    const uint32_t cells[] = {512, 1024, 2048, 4096};
    // each candidate's best of `reps`; the largest grid within 3% of the
    // fastest wins, so timing noise on a small state cannot pick a grid too
    // small for the larger states that follow
    auto pick = [](const float* ms, const uint32_t* g, int n) {
        float best = 1e30f;
        for (int i = 0; i < n; i++) best = std::min(best, ms[i]);
        uint32_t choice = g[0];
        for (int i = 0; i < n; i++)
            if (ms[i] <= 1.03f * best) choice = std::max(choice, g[i]);
        return choice;
    };
    float ms_w[5], ms_c[4];
    for (int i = 0; i < 5; i++) {
        e->grid_waves = waves[i];
        if ((rc = time_events(&ms_w[i]))) { e->prof = prof; return rc; }
    }
    const uint32_t bw = pick(ms_w, waves, 5);
    e->grid_waves = bw;
    for (int i = 0; i < 4; i++) {
        e->grid_cells = cells[i];
        if ((rc = time_events(&ms_c[i]))) { e->prof = prof; return rc; }
    }
    const uint32_t bc = pick(ms_c, cells, 4);
    e->grid_cells = bc;
    float best = 1e30f;
    for (int i = 0; i < 4; i++) best = std::min(best, ms_c[i]);
    // the queue holds 4 times the events of this state (it still grows on demand)
    unsigned cnt[CS_EV_SHARDS * CS_EV_STRIDE], nev = 0;
    HIPCHK(e, hipMemcpy(cnt, e->d.ev_count, sizeof(cnt), hipMemcpyDeviceToHost));
    for (int s = 0; s < CS_EV_SHARDS; s++) nev += cnt[s * CS_EV_STRIDE];
    const uint64_t want = std::min<uint64_t>((uint64_t)nev * 4u, 0xffffffffull);
    if (want > e->d.ev_cap && (rc = cs_set_queue_capacity(e, (uint32_t)want, 0))) { e->prof = prof; return rc; }
    e->ev_pool.push_back(a);
    e->ev_pool.push_back(b);
    e->prof = prof;
    drop_graph(e);
    if (out) {
        out->grid_waves = bw;
        out->grid_cells = bc;
        out->event_capacity = e->d.ev_cap;
        out->draw_capacity = e->d.draw_cap;
        out->events_at_state = nev;
        out->event_phase_ms = best;
    }
    return 0;
}

int cs_set_dest_tables(cs_engine* e, int stride)
{
    if (!e || stride < -1) return CS_EINVAL;
    e->dest_stride = stride;
    return 0;
}

int cs_set_dest_index(cs_engine* e, int enable)
{
    if (!e) return CS_EINVAL;
    e->dest_index = enable != 0;
    return 0;
}

int cs_get_dest_tables(const cs_engine* e, int* stride)
{
    if (!e || !stride) return CS_EINVAL;
    *stride = e->R ? e->d.ck : e->dest_stride;
    return 0;
}

int cs_set_step_graphs(cs_engine* e, int enable)
{
    if (!e) return CS_EINVAL;
    e->use_graphs = enable != 0;
    if (!e->use_graphs) drop_graph(e);
    return 0;
}

int cs_get_counters(cs_engine* e, int real, cs_counters* out, int reset)
{
    if (!e || !out) return CS_EINVAL;
    if (real < 0 || real >= e->R) return set_err(e, CS_EINVAL, "realisation %d out of range", real);
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    Dev& d = e->d;
    cs::Counters c;
    HIPCHK(e, hipMemcpy(&c, d.cnt + real, sizeof(c), hipMemcpyDeviceToHost));
    memset(out, 0, sizeof(*out));
    out->cumI = (int64_t)c.cumI; out->cumC = (int64_t)c.cumC; out->cumDC = (int64_t)c.cumDC; out->cumFC = (int64_t)c.cumFC;
    out->cumTG = (int64_t)c.cumTG; out->cumSI = (int64_t)c.cumSI; out->nTG = (int64_t)c.nTG;
    out->cumAC = (int64_t)c.cumAC; out->cumACS = (int64_t)c.cumACS; out->cumAA = (int64_t)c.cumAA;
    out->cumHQ = (int64_t)c.cumHQ; out->cumAH = (int64_t)c.cumAH; out->cumAPC = (int64_t)c.cumAPC;
    out->cumAPA = (int64_t)c.cumAPA; out->cumAPCS = (int64_t)c.cumAPCS;
    out->sumRad2 = c.sumRad2;
    memcpy(&out->maxRad2, &c.maxRad2_bits, 8);
    for (int i = 0; i < CS_INFECT_TYPE_MASK; i++) out->cumItype[i] = (int64_t)c.cumItype[i];
    for (int a = 0; a < CS_NUM_AGE_GROUPS; a++) {
        out->cumIa[a] = (int64_t)c.cumIa[a]; out->cumCa[a] = (int64_t)c.cumCa[a]; out->cumDa[a] = (int64_t)c.cumDa[a];
    }
    for (int i = 0; i < 2; i++) { out->cumI_keyworker[i] = (int64_t)c.cumI_keyworker[i]; out->cumC_keyworker[i] = (int64_t)c.cumC_keyworker[i]; }
    for (int k = 0; k < CS_NUM_SEV_KINDS; k++) {
        out->sev_prev[k] = (int64_t)c.sev_prev[k]; out->sev_cum[k] = (int64_t)c.sev_cum[k];
        for (int a = 0; a < CS_NUM_AGE_GROUPS; a++) {
            out->sev_prev_age[k][a] = (int64_t)c.sev_prev_age[k][a];
            out->sev_cum_age[k][a] = (int64_t)c.sev_cum_age[k][a];
        }
    }
    const size_t NC = e->NC, oc = (size_t)real * NC;
    std::vector<int32_t> v(NC);
    int64_t* tot[5] = {&out->S, &out->L, &out->I, &out->R, &out->D};
    int32_t* src[5] = {d.cS, d.cL, d.cI, d.cR, d.cD};
    for (int q = 0; q < 5; q++) {
        HIPCHK(e, hipMemcpy(v.data(), src[q] + oc, NC * 4, hipMemcpyDeviceToHost));
        int64_t s = 0;
        for (int32_t x : v) s += x;
        *tot[q] = s;
    }
    if (reset) {
        unsigned long long keep = c.maxRad2_bits;
        memset(&c, 0, sizeof(c));
        c.maxRad2_bits = keep;
        HIPCHK(e, hipMemcpy(d.cnt + real, &c, sizeof(c), hipMemcpyHostToDevice));
    }
    return 0;
}

int cs_get_intervention_state(cs_engine* e, int real, cs_intervention_state* out)
{
    if (!e || !out) return CS_EINVAL;
    if (real < 0 || real >= e->R) return set_err(e, CS_EINVAL, "realisation %d out of range", real);
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    Dev& d = e->d;
    cs::IntState s;
    HIPCHK(e, hipMemcpy(&s, d.ist + real, sizeof(s), hipMemcpyDeviceToHost));
    std::vector<uint8_t> mc(e->NM);
    HIPCHK(e, hipMemcpy(mc.data(), d.mc_state + (size_t)real * e->NM, e->NM, hipMemcpyDeviceToHost));
    memset(out, 0, sizeof(*out));
    out->alert = s.alert;
    out->trigger = s.trig;
    out->control_prop_cases_id = e->P.interventions_active ? s.ctrl : e->P.control_prop_cases_id;
    out->case_isolation_start = s.ci_start;
    out->hquarantine_start = s.hq_start;
    out->place_close_start = s.pc_start;
    for (uint8_t v : mc) out->mcells_closed += v == cs::MC_TREATED;
    return 0;
}

int cs_get_timeseries(cs_engine* e, int32_t* out, int max_samples, int* n_samples)
{
    if (!e || !out || !n_samples) return CS_EINVAL;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    Dev& d = e->d;
    int n = std::min(e->ts_n, max_samples);
    for (int r = 0; r < d.R; r++)
        if (n) HIPCHK(e, hipMemcpy(out + (size_t)r * max_samples * CS_TS_FIELDS, d.ts_buf + (size_t)r * d.ts_max * CS_TS_FIELDS,
                                   (size_t)n * CS_TS_FIELDS * 4, hipMemcpyDeviceToHost));
    *n_samples = n;
    return 0;
}

int cs_get_timeseries_age(cs_engine* e, int32_t* out, int max_samples, int* n_samples)
{
    if (!e || !out || !n_samples || max_samples < 0) return CS_EINVAL;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    Dev& d = e->d;
    const int n = std::min(e->ts_n, max_samples);
    const size_t W = CS_TSA_FIELDS * CS_NUM_AGE_GROUPS;
    for (int r = 0; r < d.R && n; r++)
        HIPCHK(e, hipMemcpy2D(out + (size_t)r * max_samples * W, W * 4, d.ts_ext + (size_t)r * d.ts_max * d.ext_w,
                              (size_t)d.ext_w * 4, W * 4, (size_t)n, hipMemcpyDeviceToHost));
    *n_samples = n;
    return 0;
}

int cs_get_timeseries_adunit(cs_engine* e, int32_t* out, int max_samples, int* n_samples)
{
    if (!e || !out || !n_samples || max_samples < 0) return CS_EINVAL;
    Dev& d = e->d;
    if (e->R == 0 || !d.ad_cnt) return set_err(e, CS_ESTATE, "no series by administrative unit (do_ad_units off or units unknown)");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    const int n = std::min(e->ts_n, max_samples);
    const size_t W = (size_t)CS_TSU_FIELDS * d.NA, off = CS_TSA_FIELDS * CS_NUM_AGE_GROUPS;
    for (int r = 0; r < d.R && n; r++)
        HIPCHK(e, hipMemcpy2D(out + (size_t)r * max_samples * W, W * 4, d.ts_ext + (size_t)r * d.ts_max * d.ext_w + off,
                              (size_t)d.ext_w * 4, W * 4, (size_t)n, hipMemcpyDeviceToHost));
    *n_samples = n;
    return 0;
}

int cs_get_adunit_counters(cs_engine* e, int real, int64_t* out, int reset)
{
    if (!e || !out) return CS_EINVAL;
    Dev& d = e->d;
    if (real < 0 || real >= e->R) return set_err(e, CS_EINVAL, "realisation %d out of range", real);
    if (!d.ad_cnt) return set_err(e, CS_ESTATE, "no counters by administrative unit (do_ad_units off or units unknown)");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    const size_t n = (size_t)d.NA * CS_TSU_FIELDS;
    std::vector<unsigned long long> h(n);
    for (size_t i = 0; i < n; i++) out[i] = 0;
    for (int k = 0; k < CS_AD_COPIES; k++) {  // the sum of the copies (ad_counter)
        unsigned long long* src = d.ad_cnt + ((size_t)k * e->R + real) * n;
        HIPCHK(e, hipMemcpy(h.data(), src, n * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) out[i] += (int64_t)h[i];
        if (reset) HIPCHK(e, hipMemset(src, 0, n * 8));
    }
    return 0;
}

int cs_get_work_counts(cs_engine* e, cs_work_counts* out, int reset)
{
    if (!e || !out) return CS_EINVAL;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    cs::Work w;
    HIPCHK(e, hipMemcpy(&w, e->d.wc, sizeof(w), hipMemcpyDeviceToHost));
    using cs::wsum;
    out->n_inf = (int64_t)wsum(w.n_inf); out->n_hh_inf = (int64_t)wsum(w.n_hh_inf); out->n_hh_member = (int64_t)wsum(w.n_hh_member);
    out->n_hh_scan = (int64_t)wsum(w.n_hh_scan);
    out->n_place_link = (int64_t)wsum(w.n_place_link); out->n_place_cand = (int64_t)wsum(w.n_place_cand);
    out->n_sp_draw = (int64_t)wsum(w.n_sp_draw); out->n_sp_susc = (int64_t)wsum(w.n_sp_susc); out->n_new = (int64_t)wsum(w.n_new);
    out->n_active = (int64_t)wsum(w.n_active); out->n_transition = (int64_t)wsum(w.n_transition); out->n_cells = (int64_t)e->NC;
    out->n_sp_scan = (int64_t)wsum(w.n_sp_scan); out->n_sp_probe = (int64_t)wsum(w.n_sp_probe);
    if (reset) HIPCHK(e, hipMemset(e->d.wc, 0, sizeof(cs::Work)));
    return 0;
}

int cs_set_profiling(cs_engine* e, int enable)
{
    if (!e) return CS_EINVAL;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    resolve_profile(e);
    e->prof = enable != 0;
    if (e->prof)  // enabling starts a new account; disabling keeps it for cs_get_profile
        for (int k = 0; k < CS_NUM_KERNEL_KINDS; k++) { e->prof_ms[k] = 0; e->prof_n[k] = 0; }
    return 0;
}

int cs_get_profile(cs_engine* e, double* ms, int64_t* launches, int reset)
{
    if (!e || !ms || !launches) return CS_EINVAL;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    resolve_profile(e);
    for (int k = 0; k < CS_NUM_KERNEL_KINDS; k++) {
        ms[k] = e->prof_ms[k];
        launches[k] = e->prof_n[k];
        if (reset) { e->prof_ms[k] = 0; e->prof_n[k] = 0; }
    }
    return 0;
}

int cs_sync(cs_engine* e)
{
    if (!e) return CS_EINVAL;
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return 0;
}

// ---------------------------------------------------------- components


#define TMP_UP(e, buf, src, n)                                                   \
    HIPCHK(e, hipMalloc(&buf.p, std::max<size_t>((n), 1) * sizeof(*buf.p)));    \
    if (src) HIPCHK(e, hipMemcpy(buf.p, src, (n) * sizeof(*buf.p), hipMemcpyHostToDevice))

int cs_kernel_num(cs_engine* e, const double* r2, double* out, int n)
{
    if (!e || !r2 || !out || n < 0) return CS_EINVAL;
    if (!e->have_params) return set_err(e, CS_ESTATE, "no parameters");
    HIPCHK(e, hipSetDevice(e->device));
    DBuf<double> a, b;
    TMP_UP(e, a, r2, (size_t)n);
    TMP_UP(e, b, (const double*)nullptr, (size_t)n);
    HIPCHK(e, hipMemset(e->d.status, 0, 4));
    if (n) cs::k_t_kernel_num<<<grid_for(n, 256, 1u << 30), 256, 0, e->stream>>>(e->d, a.p, b.p, n);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    HIPCHK(e, hipMemcpy(out, b.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    return check_flags(e);
}

int cs_dist2_persons(cs_engine* e, const int32_t* pa, const int32_t* pb, double* out, int n)
{
    if (!e || !pa || !pb || !out || n < 0) return CS_EINVAL;
    if (!e->have_pop) return set_err(e, CS_ESTATE, "no population");
    for (int i = 0; i < n; i++)
        if (pa[i] < 0 || pa[i] >= e->N || pb[i] < 0 || pb[i] >= e->N) return set_err(e, CS_EINVAL, "person index out of range");
    HIPCHK(e, hipSetDevice(e->device));
    DBuf<int32_t> a, b;
    DBuf<double> o;
    TMP_UP(e, a, pa, (size_t)n);
    TMP_UP(e, b, pb, (size_t)n);
    TMP_UP(e, o, (const double*)nullptr, (size_t)n);
    if (n) cs::k_t_dist2<<<grid_for(n, 256, 1u << 30), 256, 0, e->stream>>>(e->d, a.p, b.p, o.p, n);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    HIPCHK(e, hipMemcpy(out, o.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    return 0;
}

int cs_get_cell_tables(cs_engine* e, int real, float* mt, float* ct, int32_t* inv, float* tp, int32_t* S0)
{
    if (!e) return CS_EINVAL;
    if (!e->have_pop) return set_err(e, CS_ESTATE, "no population");
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    const size_t NC = e->NC;
    Dev& d = e->d;
    if (d.ck || (mt && !e->mt_buf && e->R == 0)) {  // no stored tables: expand them
        DBuf<float> m, c;
        DBuf<int32_t> v;
        if (mt) {
            TMP_UP(e, m, (const float*)nullptr, NC * NC);
            cs::k_build_max_trans<<<grid_for((long long)NC * NC, 256, 1u << 30), 256, 0, e->stream>>>(d, m.p);
            HIPCHK(e, hipStreamSynchronize(e->stream));
            HIPCHK(e, hipMemcpy(mt, m.p, NC * NC * 4, hipMemcpyDeviceToHost));
        }
        if (real >= 0 && real < e->R && (ct || inv)) {
            TMP_UP(e, c, (const float*)nullptr, NC * NC);
            TMP_UP(e, v, (const int32_t*)nullptr, NC * 1025);
            cs::k_expand_rows<<<grid_for((long long)NC, 64), 64, 0, e->stream>>>(d, real, nullptr, c.p, v.p);
            HIPCHK(e, hipStreamSynchronize(e->stream));
            if (ct) HIPCHK(e, hipMemcpy(ct, c.p, NC * NC * 4, hipMemcpyDeviceToHost));
            if (inv) HIPCHK(e, hipMemcpy(inv, v.p, NC * 1025 * 4, hipMemcpyDeviceToHost));
        }
        if (real >= 0 && real < e->R) {
            const size_t oc = (size_t)real * NC;
            if (tp) HIPCHK(e, hipMemcpy(tp, d.tot_prob + oc, NC * 4, hipMemcpyDeviceToHost));
            if (S0) HIPCHK(e, hipMemcpy(S0, d.S0 + oc, NC * 4, hipMemcpyDeviceToHost));
        }
        int rc = check_flags(e);
        return rc;
    }
    if (mt) HIPCHK(e, hipMemcpy(mt, d.max_trans, NC * NC * 4, hipMemcpyDeviceToHost));
    if (real >= 0 && real < e->R) {
        const size_t oc = (size_t)real * NC;
        if (ct) HIPCHK(e, hipMemcpy(ct, d.cum_trans + oc * NC, NC * NC * 4, hipMemcpyDeviceToHost));
        if (inv) {
            HIPCHK(e, hipMemcpy(inv, d.invcdf + oc * 1025, NC * 1025 * 4, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < NC * 1025; k++) inv[k] &= 0x7fffffff;  // the InvCDF cell (the flag is the library's own)
        }
        if (tp) HIPCHK(e, hipMemcpy(tp, d.tot_prob + oc, NC * 4, hipMemcpyDeviceToHost));
        if (S0) HIPCHK(e, hipMemcpy(S0, d.S0 + oc, NC * 4, hipMemcpyDeviceToHost));
    }
    return 0;
}

int cs_eval_hazards(cs_engine* e, int real, int kind, uint16_t ts, const int32_t* person, const int32_t* other, double* out, int n)
{
    if (!e || !person || !out || n < 0 || kind < 0 || kind > 7) return CS_EINVAL;
    if (real < 0 || real >= e->R) return set_err(e, CS_EINVAL, "realisation %d out of range", real);
    for (int i = 0; i < n; i++) {
        if (person[i] < 0 || person[i] >= e->N) return set_err(e, CS_EINVAL, "person index out of range");
        if (other) {
            int lim = (kind == 1 || kind == 5) ? e->NPT : e->N;
            if (other[i] < 0 || other[i] >= lim) return set_err(e, CS_EINVAL, "second index out of range");
        }
    }
    HIPCHK(e, hipSetDevice(e->device));
    DBuf<int32_t> a, b;
    DBuf<double> o;
    TMP_UP(e, a, person, (size_t)n);
    TMP_UP(e, b, other, (size_t)n);
    TMP_UP(e, o, (const double*)nullptr, (size_t)n);
    if (n) cs::k_t_hazards<<<grid_for(n, 256, 1u << 30), 256, 0, e->stream>>>(e->d, real, kind, ts, a.p, other ? b.p : nullptr, o.p, n);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    HIPCHK(e, hipMemcpy(out, o.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    return 0;
}

int cs_icdf_choose(cs_engine* e, int table, const double* u, const double* mean, uint16_t* out, int n)
{
    if (!e || !u || !mean || !out || n < 0 || table < 0 || table >= CS_NUM_ICDF) return CS_EINVAL;
    if (!e->have_params) return set_err(e, CS_ESTATE, "no parameters");
    for (int i = 0; i < n; i++)
        if (!(u[i] > 0.0 && u[i] < 1.0)) return set_err(e, CS_EINVAL, "uniform outside (0,1)");
    HIPCHK(e, hipSetDevice(e->device));
    DBuf<double> a, b;
    DBuf<uint16_t> o;
    TMP_UP(e, a, u, (size_t)n);
    TMP_UP(e, b, mean, (size_t)n);
    TMP_UP(e, o, (const uint16_t*)nullptr, (size_t)n);
    if (n) cs::k_t_choose<<<grid_for(n, 256, 1u << 30), 256, 0, e->stream>>>(e->d, table, a.p, b.p, o.p, n);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    HIPCHK(e, hipMemcpy(out, o.p, (size_t)n * 2, hipMemcpyDeviceToHost));
    return 0;
}

int cs_select_infector(cs_engine* e, const float* cell_inf, int n_inf, const double* u, int32_t* out, int n)
{
    if (!e || !cell_inf || !u || !out || n < 0 || n_inf <= 0) return CS_EINVAL;
    HIPCHK(e, hipSetDevice(e->device));
    DBuf<float> c;
    DBuf<double> a;
    DBuf<int32_t> o;
    TMP_UP(e, c, cell_inf, (size_t)n_inf);
    TMP_UP(e, a, u, (size_t)n);
    TMP_UP(e, o, (const int32_t*)nullptr, (size_t)n);
    if (n) cs::k_t_select_infector<<<grid_for(n, 256, 1u << 30), 256, 0, e->stream>>>(c.p, n_inf, a.p, o.p, n);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    HIPCHK(e, hipMemcpy(out, o.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    return 0;
}

int cs_select_dest_cell(cs_engine* e, int real, int cell, const double* u, int32_t* out, int n)
{
    if (!e || !u || !out || n < 0) return CS_EINVAL;
    if (real < 0 || real >= e->R || cell < 0 || cell >= e->NC) return set_err(e, CS_EINVAL, "realisation or cell out of range");
    for (int i = 0; i < n; i++)
        if (!(u[i] > 0.0 && u[i] < 1.0)) return set_err(e, CS_EINVAL, "uniform outside (0,1)");
    HIPCHK(e, hipSetDevice(e->device));
    DBuf<double> a;
    DBuf<int32_t> o;
    TMP_UP(e, a, u, (size_t)n);
    TMP_UP(e, o, (const int32_t*)nullptr, (size_t)n);
    if (n) cs::k_t_select_dest<<<grid_for(n, 256, 1u << 30), 256, 0, e->stream>>>(e->d, real, cell, a.p, o.p, n);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    HIPCHK(e, hipMemcpy(out, o.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    return 0;
}

int cs_philox(cs_engine* e, const uint32_t* ck, uint32_t* out, int n)
{
    if (!e || !ck || !out || n < 0) return CS_EINVAL;
    HIPCHK(e, hipSetDevice(e->device));
    DBuf<uint32_t> a, o;
    TMP_UP(e, a, ck, (size_t)n * 6);
    TMP_UP(e, o, (const uint32_t*)nullptr, (size_t)n * 4);
    if (n) cs::k_t_philox<<<grid_for(n, 256, 1u << 30), 256, 0, e->stream>>>(a.p, o.p, n);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    HIPCHK(e, hipMemcpy(out, o.p, (size_t)n * 16, hipMemcpyDeviceToHost));
    return 0;
}

// sampler draws for distribution tests: out[i] from stream {key, c0, c1, i}
}  // extern "C"

namespace {

// InitTransmissionCoeffs' sums from per-person draws already on the device
int itc_sums(cs_engine* e, const double* prop_age_group, const float* d_inf, const uint16_t* d_rec, const int8_t* d_st,
             cs_transmission_coeffs* out)
{
    const cs_params& P = e->P;
    DBuf<double> pa, part, res;
    if (P.got_waifw_spatial) {
        if (!prop_age_group) return set_err(e, CS_EINVAL, "prop_age_group is needed with a spatial WAIFW matrix");
        if (!e->adunits_valid) return set_err(e, CS_EINVAL, "administrative units of the population are missing or out of range");
        TMP_UP(e, pa, prop_age_group, (size_t)e->NA * CS_NUM_AGE_GROUPS);
    }
    const int nb = std::max(1, std::min(2048, (e->N + ITC_BLOCK - 1) / ITC_BLOCK));
    TMP_UP(e, part, (const double*)nullptr, (size_t)nb * cs::ITC_NSUM);
    TMP_UP(e, res, (const double*)nullptr, (size_t)cs::ITC_NSUM);
    cs::k_itc_person<<<nb, ITC_BLOCK, 0, e->stream>>>(e->d, d_inf, d_rec, d_st, pa.p, part.p);
    cs::k_itc_final<<<1, 64, 0, e->stream>>>(part.p, nb, res.p);
    HIPCHK(e, hipGetLastError());
    double v[cs::ITC_NSUM];
    HIPCHK(e, hipMemcpyAsync(v, res.p, sizeof(v), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    // the scalar tail of src/SetupModel.cpp:748-845
    memset(out, 0, sizeof(*out));
    const double N = (double)e->N;
    double place = 0;
    for (int pt = 0; pt < CS_MAX_PLACE_TYPES; pt++) {
        out->r0_places_by_type[pt] = v[cs::ITC_PLACE0 + pt] / N;
        place += v[cs::ITC_PLACE0 + pt];
    }
    out->household_sar = v[cs::ITC_HH] / v[cs::ITC_HH_DENOM];
    out->r0_household = v[cs::ITC_HH] / N;
    out->r0_places = place / N;
    out->mean_infectious_days = v[cs::ITC_REC_DAYS] / N;
    out->mean_infectious_steps = v[cs::ITC_REC_STEPS] / N;
    out->spatial_r0 = v[cs::ITC_SPATIAL] / N;
    double beta = P.local_beta;
    if (!P.fix_local_beta) {
        beta = P.do_si ? (P.r0 / out->spatial_r0 - out->r0_household - out->r0_places)
                       : (P.r0 - out->r0_household - out->r0_places) / out->spatial_r0;
        if (beta < 0) beta = 0;
    }
    out->local_beta = beta;
    out->r0_spatial = out->spatial_r0 * beta;
    out->r0 = out->r0_household + out->r0_places + out->r0_spatial;
    return 0;
}

}  // namespace

extern "C" {

int cs_init_transmission_coeffs(cs_engine* e, const double* prop_age_group, uint64_t seed, cs_transmission_coeffs* out)
{
    if (!e || !out) return CS_EINVAL;
    if (!e->have_params || !e->have_pop) return set_err(e, CS_ESTATE, "parameters and population must come first");
    HIPCHK(e, hipSetDevice(e->device));
    DBuf<float> inf;
    DBuf<uint16_t> rec;
    TMP_UP(e, inf, (const float*)nullptr, (size_t)e->N);
    TMP_UP(e, rec, (const uint16_t*)nullptr, (size_t)e->N);
    cs::k_itc_draw<<<grid_for(e->N, 256), 256, 0, e->stream>>>(e->d, seed, inf.p, rec.p);
    HIPCHK(e, hipGetLastError());
    int rc = itc_sums(e, prop_age_group, inf.p, rec.p, nullptr, out);
    if (rc) return rc;
    if (!e->P.fix_local_beta) {  // the reference sets P.LocalBeta
        e->P.local_beta = out->local_beta;
        HIPCHK(e, hipMemcpy(e->dP, &e->P, sizeof(cs_params), hipMemcpyHostToDevice));
    }
    return 0;
}

int cs_transmission_coeffs_from_draws(cs_engine* e, const double* prop_age_group, const float* infectiousness,
                                      const uint16_t* infectious_steps, const int8_t* infstat, cs_transmission_coeffs* out)
{
    if (!e || !out || !infectiousness || !infectious_steps) return CS_EINVAL;
    if (!e->have_params || !e->have_pop) return set_err(e, CS_ESTATE, "parameters and population must come first");
    HIPCHK(e, hipSetDevice(e->device));
    DBuf<float> inf;
    DBuf<uint16_t> rec;
    DBuf<int8_t> st;
    TMP_UP(e, inf, infectiousness, (size_t)e->N);
    TMP_UP(e, rec, infectious_steps, (size_t)e->N);
    if (infstat) TMP_UP(e, st, infstat, (size_t)e->N);
    return itc_sums(e, prop_age_group, inf.p, rec.p, st.p, out);
}

int cs_sample(cs_engine* e, int which, uint64_t key, uint32_t c0, uint32_t c1, double a, double b, double* out, int n)
{
    if (!e || !out || n < 0 || which < 0 || which > 4) return CS_EINVAL;
    HIPCHK(e, hipSetDevice(e->device));
    DBuf<double> o;
    TMP_UP(e, o, (const double*)nullptr, (size_t)n);
    if (n) cs::k_t_sample<<<grid_for(n, 256, 1u << 30), 256, 0, e->stream>>>(which, key, c0, c1, a, b, o.p, n);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    HIPCHK(e, hipMemcpy(out, o.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------ batches
// engines of one device (a region each: its own population, parameters and
// realisations) advanced together: every output step, each engine's
// captured step graph is launched on the engine's own stream, all before
// any wait, so the device's hardware queues run the regions side by side

struct cs_batch {
    std::vector<cs_engine*> eng;
    int device = 0;
    char err[512] = {0};
};

namespace {

int bset_err(cs_batch* b, int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(b->err, sizeof(b->err), fmt, ap);
    va_end(ap);
    return code;
}

// an engine's error, as the batch's
int bfrom(cs_batch* b, size_t i, int code)
{
    return bset_err(b, code, "engine %zu: %s", i, b->eng[i]->err);
}

}  // namespace

extern "C" {

int cs_batch_create(cs_batch** out, cs_engine* const* engines, int n)
{
    if (!out || !engines || n < 1) return CS_EINVAL;
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!engines[i]) return CS_EINVAL;
        if (engines[i]->device != engines[0]->device) return CS_EINVAL;
        for (int j = 0; j < i; j++)
            if (engines[j] == engines[i]) return CS_EINVAL;
    }
    cs_batch* b = new cs_batch;
    b->eng.assign(engines, engines + n);
    b->device = engines[0]->device;
    *out = b;
    return 0;
}

void cs_batch_destroy(cs_batch* b) { delete b; }

const char* cs_batch_last_error(const cs_batch* b) { return b ? b->err : "null batch"; }

int cs_batch_run_steps(cs_batch* b, double t0, int n_steps, int sps)
{
    if (!b || n_steps < 0 || sps <= 0) return CS_EINVAL;
    const size_t n = b->eng.size();
    int rc;
    for (size_t i = 0; i < n; i++)
        if ((rc = run_prepare(b->eng[i], t0, n_steps, sps))) return bfrom(b, i, rc);
    const int periods = n_steps / sps;
    bool graphs = periods > 0;
    for (size_t i = 0; i < n; i++) graphs = graphs && b->eng[i]->use_graphs && !b->eng[i]->prof;
    std::vector<int> done(n, 0);
    std::vector<double> tt(n, t0);
    if (graphs) {
        if (hipSetDevice(b->device) != hipSuccess) return bset_err(b, CS_EHIP, "hipSetDevice");
        int p = 0;
        for (int guard = 0; p < periods && guard < 64;) {
            for (size_t i = 0; i < n; i++) {
                cs_engine* e = b->eng[i];
                if ((!e->gexec || e->graph_sps != sps) && (rc = capture_graph(e, sps))) return bfrom(b, i, rc);
            }
            const int nb = std::min(periods - p, 8);  // output steps between flag checks
            // every engine's replays are issued before any wait (each on its
            // own stream, behind that engine's set-up)
            for (int k = 0; k < nb; k++)
                for (size_t i = 0; i < n; i++)
                    if (hipGraphLaunch(b->eng[i]->gexec, b->eng[i]->stream) != hipSuccess)
                        return bset_err(b, CS_EHIP, "engine %zu graph launch", i);
            for (size_t i = 0; i < n; i++) {
                hipError_t st = hipStreamSynchronize(b->eng[i]->stream);
                if (st != hipSuccess) return bset_err(b, CS_EHIP, "engine %zu replay: %s", i, hipGetErrorString(st));
            }
            for (size_t i = 0; i < n; i++) {
                cs_engine* e = b->eng[i];
                cs::Clock c;
                if ((rc = read_clock(e, &c))) return bfrom(b, i, rc);
                unsigned f = 0;
                if (hipMemcpy(&f, e->d.status, sizeof(f), hipMemcpyDeviceToHost) != hipSuccess) return bset_err(b, CS_EHIP, "status");
                if (f & cs::FLAG_OVERFLOW) {
                    // e halted at step c.step and its later replays did
                    // nothing: grow its lists (which drops its graph, to be
                    // captured afresh above) and bring it to the others'
                    // step directly
                    guard++;
                    if ((rc = run_recover(e, c, (p + nb) * sps, sps))) return bfrom(b, i, rc);
                } else if ((rc = check_flags(e))) {
                    return bfrom(b, i, rc);
                }
            }
            p += nb;
        }
        if (p < periods) return bset_err(b, CS_EOVERFLOW, "work lists keep overflowing");
        for (size_t i = 0; i < n; i++) {
            cs::Clock c;
            if ((rc = read_clock(b->eng[i], &c))) return bfrom(b, i, rc);
            done[i] = c.step;
            tt[i] = c.t;
        }
    }
    for (size_t i = 0; i < n; i++)
        if ((rc = run_finish(b->eng[i], tt[i], done[i], n_steps, sps))) return bfrom(b, i, rc);
    return 0;
}

}  // extern "C"
