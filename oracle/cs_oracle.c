/*
 * cs_oracle.c -- CPU oracle for the infection step.  TEST INFRASTRUCTURE ONLY
 * (see cs_oracle.h): the checker for the GPU library and the CPU baseline.
 *
 * Each function cites the reference code it restates.  Where the GPU design
 * fixes an order the reference leaves to thread scheduling (event
 * resolution, per-cell list maintenance, random streams), the oracle follows
 * the GPU's documented rule so both produce identical results:
 *   - random streams: Philox4x32-10 (Salmon et al., SC'11), key = realisation
 *     seed, counter = {timestep | pass<<16, entity, sub, block}; uniforms
 *     ((hi:lo)>>11 + 0.5) * 2^-53, two per block;
 *   - event order: key = infector<<32 | pass<<27 | sub; per infectee the
 *     smallest key does DoInfect (src/Sweep.cpp:860-880 "first wins"), the
 *     largest key's infector/type is what remains recorded;
 *   - list maintenance: a batch X leaving a segment [lo,hi) is swapped into
 *     its tail: sorted positions of X below hi-|X| are paired, in order,
 *     with the sorted positions in [hi-|X|,hi) not holding X;
 *   - logarithm and exponential: the fixed polynomial evaluations below,
 *     identical on host and device (only +,-,*,/ and exact scaling).
 */
#include "cs_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ math */

static double ld_exp2(int k)
{
    /* 2^k for -1022 <= k <= 1023 by construction of the bit pattern */
    uint64_t b = (uint64_t)(k + 1023) << 52;
    double d;
    memcpy(&d, &b, 8);
    return d;
}

double or_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return x;
    uint64_t b;
    memcpy(&b, &x, 8);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) { /* subnormal */
        x *= 18014398509481984.0; /* 2^54 */
        memcpy(&b, &x, 8);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m;
    memcpy(&m, &b, 8);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double p = 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    double lm = 2.0 * s * p;
    double de = (double)e;
    return de * 6.93147180369123816490e-01 + (lm + de * 1.90821492927058770002e-10);
}

double or_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.1332191019412) return 0.0;
    double kf = floor(x * 1.4426950408889634074 + 0.5);
    int k = (int)kf;
    double r = (x - kf * 6.93147180369123816490e-01) - kf * 1.90821492927058770002e-10;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    if (k > 1023) return p * ld_exp2(1023) * ld_exp2(k - 1023);
    if (k < -1021) return p * ld_exp2(k + 1000) * ld_exp2(-1000);
    return p * ld_exp2(k);
}

static double pow_int(double q, int32_t n)
{
    double r = 1.0, b = q;
    while (n > 0) {
        if (n & 1) r *= b;
        b *= b;
        n >>= 1;
    }
    return r;
}

/* --------------------------------------------------------------- philox */

static void philox4x32_10(const uint32_t in[4], uint32_t k0, uint32_t k1, uint32_t out[4])
{
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0;
        uint32_t h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
        c0 = h1 ^ c1 ^ k0;
        c1 = l1;
        c2 = h0 ^ c3 ^ k1;
        c3 = l0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void or_philox(const uint32_t* ck, uint32_t* out, int n)
{
    for (int i = 0; i < n; i++) philox4x32_10(ck + 6 * i, ck[6 * i + 4], ck[6 * i + 5], out + 4 * i);
}

typedef struct {
    uint32_t k0, k1, c0, c1, c2, c3;
    double buf[2];
    int idx;
} rs;

static void rs_init(rs* s, uint64_t key, uint32_t c0, uint32_t c1, uint32_t c2)
{
    s->k0 = (uint32_t)key; s->k1 = (uint32_t)(key >> 32);
    s->c0 = c0; s->c1 = c1; s->c2 = c2; s->c3 = 0; s->idx = 2;
}

static double u53(uint32_t hi, uint32_t lo)
{
    uint64_t v = (((uint64_t)hi << 32) | lo) >> 11;
    return ((double)v + 0.5) * (1.0 / 9007199254740992.0);
}

static double rs_next(rs* s)
{
    if (s->idx == 2) {
        uint32_t in[4] = {s->c0, s->c1, s->c2, s->c3}, o[4];
        s->c3++;
        philox4x32_10(in, s->k0, s->k1, o);
        s->buf[0] = u53(o[0], o[1]);
        s->buf[1] = u53(o[2], o[3]);
        s->idx = 0;
    }
    return s->buf[s->idx++];
}

/* The cell's running sum of spatial infectiousness (src/Sweep.cpp:649-690,
 * a sequential double sum there) is kept exactly in 64.64 fixed point: the
 * sum is then independent of the order of additions, so a parallel scan
 * reproduces it, and each prefix is rounded to double once. */
typedef struct { uint64_t hi, lo; } fx128;
static fx128 fx_from(double v)
{
    fx128 r;
    double f = floor(v);
    r.hi = (uint64_t)f;
    r.lo = (uint64_t)((v - f) * 18446744073709551616.0); /* 2^64 */
    return r;
}
static fx128 fx_add(fx128 a, fx128 b)
{
    fx128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo);
    return r;
}
static double fx_to_double(fx128 a) { return (double)a.hi + (double)a.lo * 5.42101086242752217e-20; /* 2^-64 */ }

/* bound on the spatial rejection loop (src/Sweep.cpp:753-854 has none; a
 * cell whose first S0 slots hold only dead persons would loop for ever) */
#define OR_MAX_ATTEMPTS (1 << 20)

/* pass identifiers (high half of counter word 0) */
enum { PS_HH = 1, PS_PW_SAMPLE, PS_PW_ACCEPT, PS_PB_SAMPLE, PS_PB_ACCEPT, PS_SP_COUNT,
       PS_SP_DRAW, PS_INFECT, PS_INCUB, PS_CASE, PS_RECOVER, PS_HOLIDAY, PS_INIT, PS_SEED, PS_ITC,
       PS_HQ, PS_ISO, PS_PLACE, PS_POPGEN };
/* pass identifiers inside the event order key */
enum { KP_HH = 0, KP_PW = 1, KP_PB = 6, KP_SP = 11 };

static uint32_t cw0(unsigned ts, int pass) { return (uint32_t)ts | ((uint32_t)pass << 16); }

/* ------------------------------------------------------------- samplers */

/* Stirling-series remainder log(k!) - [(k+.5)log(k+1) - (k+1) + .5 log(2pi)] */
static double fc_corr(int32_t k)
{
    static const double tab[10] = {0.08106146679532726, 0.04134069595540929, 0.02767792568499834,
        0.02079067210376509, 0.01664469118982119, 0.01387612882307075, 0.01189670994589177,
        0.01041126526197209, 0.009255462182712733, 0.008330563433362871};
    if (k < 10) return tab[k];
    double ik = 1.0 / (double)(k + 1);
    double ik2 = ik * ik;
    return (1.0 / 12.0 - (1.0 / 360.0 - ik2 / 1260.0) * ik2) * ik;
}

/* Binomial(n, p): inversion when n*min(p,1-p) < 10, else BTRS (Hormann
 * 1993, transformed rejection with squeeze).  Same distribution as
 * ignbin_mt (src/Rand.cpp:404-662), which uses inversion / BTPE. */
static int32_t smp_binomial(rs* s, int32_t n, double pp)
{
    if (n <= 0 || pp <= 0.0) return 0;
    if (pp >= 1.0) return n;
    double p = pp < 0.5 ? pp : 1.0 - pp;
    double q = 1.0 - p;
    int32_t x;
    if ((double)n * p < 10.0) {
        double f = pow_int(q, n);
        double r = p / q;
        double u = rs_next(s);
        x = 0;
        while (u > f) {
            u -= f;
            x++;
            if (x >= n) { x = n; break; }
            f *= r * (double)(n - x + 1) / (double)x;
        }
    } else {
        double spq = sqrt((double)n * p * q);
        double b = 1.15 + 2.53 * spq;
        double a = -0.0873 + 0.0248 * b + 0.01 * p;
        double c = (double)n * p + 0.5;
        double vr = 0.92 - 4.2 / b;
        double r = p / q;
        double alpha = (2.83 + 5.1 / b) * spq;
        double m = floor((double)(n + 1) * p);
        for (;;) {
            double u = rs_next(s) - 0.5;
            double v = rs_next(s);
            double us = 0.5 - fabs(u);
            double kf = floor((2.0 * a / us + b) * u + c);
            if (kf < 0.0 || kf > (double)n) continue;
            if (us >= 0.07 && v <= vr) { x = (int32_t)kf; break; }
            double lv = or_log(v * alpha / (a / (us * us) + b));
            double ub = (m + 0.5) * or_log((m + 1.0) / (r * ((double)n - m + 1.0)))
                      + ((double)n + 1.0) * or_log(((double)n - m + 1.0) / ((double)n - kf + 1.0))
                      + (kf + 0.5) * or_log(r * ((double)n - kf + 1.0) / (kf + 1.0))
                      + fc_corr((int32_t)m) + fc_corr(n - (int32_t)m)
                      - fc_corr((int32_t)kf) - fc_corr(n - (int32_t)kf);
            if (lv <= ub) { x = (int32_t)kf; break; }
        }
    }
    return pp > 0.5 ? n - x : x;
}

/* log(k!) through the Stirling remainder above */
static double log_fact(int32_t k)
{
    double kk = (double)k;
    return fc_corr(k) + (kk + 0.5) * or_log(kk + 1.0) - (kk + 1.0) + 0.91893853320467274178;
}

/* Poisson(mu): inversion for mu < 10, else PTRS (Hormann 1993).  Same
 * distribution as ignpoi_mt (src/Rand.cpp:170-403). */
static int32_t smp_poisson(rs* s, double mu)
{
    if (!(mu > 0.0)) return 0;
    if (mu < 10.0) {
        double pk = or_exp(-mu);
        double u = rs_next(s);
        int32_t x = 0;
        while (u > pk) {
            u -= pk;
            x++;
            if (x > 10000) break;
            pk *= mu / (double)x;
        }
        return x;
    }
    double slam = sqrt(mu);
    double loglam = or_log(mu);
    double b = 0.931 + 2.53 * slam;
    double a = -0.059 + 0.02483 * b;
    double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        double u = rs_next(s) - 0.5;
        double v = rs_next(s);
        double us = 0.5 - fabs(u);
        double kf = floor((2.0 * a / us + b) * u + mu + 0.43);
        if (us >= 0.07 && v <= vr) return (int32_t)kf;
        if (kf < 0.0 || (us < 0.013 && v > us)) continue;
        if (or_log(v) + or_log(invalpha) - or_log(a / (us * us) + b) <= -mu + kf * loglam - log_fact((int32_t)kf))
            return (int32_t)kf;
    }
}

/* standard normal, Marsaglia polar method (one value per accepted pair) */
static double smp_normal(rs* s)
{
    for (;;) {
        double a = 2.0 * rs_next(s) - 1.0;
        double b = 2.0 * rs_next(s) - 1.0;
        double q = a * a + b * b;
        if (q >= 1.0 || q == 0.0) continue;
        return a * sqrt(-2.0 * or_log(q) / q);
    }
}

/* gen_gamma_mt(beta = rate, alpha = shape) restated: Marsaglia-Tsang with the
 * reference's acceptance test (src/Rand.cpp:1025-1086). */
static double smp_gamma(rs* s, double beta, double alpha)
{
    double al = alpha >= 1.0 ? alpha : alpha + 1.0;
    double d = al - 1.0 / 3.0;
    double c = 1.0 / (3.0 * sqrt(d));
    double u, z, v;
    for (;;) {
        u = rs_next(s);
        z = smp_normal(s);
        v = 1.0 + z * c;
        v = v * v * v;
        if (z <= -1.0 / c) continue;
        if (or_log(u) >= 0.5 * z * z + d - d * v + d * or_log(v)) continue;
        break;
    }
    double g = (beta != 1.0) ? (d * v) / beta : d * v;
    if (alpha < 1.0) g *= or_exp(or_log(rs_next(s)) / alpha);
    return g;
}

/* k distinct indices of [0,n): all of them when k == n, Floyd's algorithm
 * for k <= 32, selection sampling (Knuth's algorithm S) above.  A uniform
 * k-subset, as SampleWithoutReplacement (src/Rand.cpp:1109-1243). */
#define OR_FLOYD_MAX 32
static int smp_sample_wor(rs* s, int k, int n, int32_t* out)
{
    if (k >= n) {
        for (int i = 0; i < n; i++) out[i] = i;
        return n;
    }
    if (k <= OR_FLOYD_MAX) {
        int m = 0;
        for (int j = n - k; j < n; j++) {
            int t = (int)(rs_next(s) * (double)(j + 1));
            int dup = 0;
            for (int q = 0; q < m; q++) if (out[q] == t) { dup = 1; break; }
            out[m++] = dup ? j : t;
        }
        return k;
    }
    int sel = 0;
    for (int i = 0; i < n && sel < k; i++)
        if ((double)(n - i) * rs_next(s) < (double)(k - sel)) out[sel++] = i;
    return sel;
}

/* --------------------------------------------------------------- engine */

typedef struct {
    uint64_t key;
    int32_t infectee, infector;
    int16_t type;
} or_event;

typedef struct {
    uint64_t seed;
    int8_t* inf;
    uint8_t *sev_cur, *sev_final, *bits, *qc, *dct;
    uint8_t* hh_done; /* the person's household pass found no susceptible member (work count only) */
    int32_t *listpos, *infector;
    int16_t* itype;
    float *infness, *susc, *prob_care;
    uint16_t *t_inf, *t_lat, *t_rod, *t_sari, *t_crit, *t_step, *t_det, *t_as, *t_ae;
    uint16_t *t_ts, *t_te, *t_vs, *t_is, *t_qs;
    int32_t *cS, *cL, *cI, *cR, *cD, *cumTC, *S0;
    float* tot_prob;
    float* cum_trans;
    int32_t* invcdf;
    int32_t* perm;
    uint16_t *pl_cs[CS_MAX_PLACE_TYPES], *pl_ce[CS_MAX_PLACE_TYPES];
    uint8_t *m_socdist, *m_moverest;
    uint32_t* m_trig; /* [n_mcells][6] */
    cs_counters cnt;
    int64_t* ad_cnt;       /* [NAU][CS_TSU_FIELDS] cumI, cumC, cumDC, cumD by administrative unit (StateT's *_adunit) */
    double prev_prop_susc;
    double samp_prop_susc; /* State.S / PopSize as of the last sample */
    int has_beta;          /* per-realisation LocalBeta override */
    int has_thr;           /* per-realisation trigger threshold override */
    double thr;
    double local_beta;
    int n_seed;            /* per-realisation seed count, or -1 */
    double loc_init[2];
    /* interventions (one copy of the P / AdUnits fields the trigger logic
     * sets, src/CovidSim.cpp:3485-3689, per realisation) */
    int alert;             /* the alert condition has held at a sample */
    int32_t trig;          /* State.trigDetectedCases of the last sample */
    double ctrl;           /* P.ControlPropCasesId */
    double ci_start, hq_start, pc_start, pc_prev;
    int32_t* hist;         /* [OR_TRIG_RING][3]: cumDC, cumCritical, D at sample m (slot m % OR_TRIG_RING) */
    uint8_t* mc_state;     /* [NM] Microcell.placeclose: 0 untreated, 1 treated, 2 don't treat again */
    uint16_t* mc_end;      /* [NM] Microcell.place_end_time */
    uint8_t* pl_once[CS_MAX_PLACE_TYPES]; /* Place.control_trig at USHRT_MAX - 1 (closed under once-only) */
} or_real;

/* samples of trigger history kept per realisation: a trigger window
 * (TriggersSamplingInterval, WindowToEvaluateTriggerAlert) that reaches back
 * further than this, within the run, is refused */
#define OR_TRIG_RING 256

struct or_engine {
    char err[256];
    cs_params P;
    const double* klookup; /* owned copies */
    const double* khires;
    double* kl;
    double* kh;
    int N, NH, NC, NCT, NM, NA, NPT;
    int32_t *hh, *mcell, *pc; /* pc: populated-cell index */
    uint8_t *age, *flags;
    int32_t* plinks;
    uint16_t* pglinks;
    int32_t* hfirst;
    uint16_t *hnh, *hnhr;
    float* hloc;
    int32_t *cidx, *cn, *cstart, *cmem;
    int32_t *madunit, *mstart, *mmem;
    int np[CS_MAX_PLACE_TYPES];
    int32_t *pl_n[CS_MAX_PLACE_TYPES], *pl_mcell[CS_MAX_PLACE_TYPES], *pl_ms[CS_MAX_PLACE_TYPES],
        *pl_mem[CS_MAX_PLACE_TYPES], *pl_gs0[CS_MAX_PLACE_TYPES], *pl_gst[CS_MAX_PLACE_TYPES],
        *pl_gsz[CS_MAX_PLACE_TYPES];
    float* max_trans;
    int32_t *mpl_start, *mpl; /* places by microcell (Microcell.places): type << 27 | place */
    int nreal;
    or_real* R;
    /* scratch */
    or_event* ev;
    size_t nev, cap_ev;
    float* cell_inf;
    int32_t* samp;
    int32_t* ts_buf;
    int ts_n, ts_cap;
    int ts_fresh;           /* no sample since the last alloc/reset/load: the series restarts */
    int64_t* ts_last;       /* [nreal][3] cumI, cumC, D at the previous sample */
    /* the samples by age group and by administrative unit: rows of ext_w
     * words, differencing the cumulative counters kept in ext_last */
    int32_t* ts_ext;
    int64_t* ext_last;      /* [nreal][CS_TSA_CUM * CS_NUM_AGE_GROUPS + CS_TSU_FIELDS * NAU] */
    int NAU, ext_w;         /* units counted (0: none), row width */
    cs_work_counts wc;
    int error;
};

const char* or_last_error(const or_engine* e) { return e ? e->err : "null engine"; }

static int fail(or_engine* e, int code, const char* msg)
{
    snprintf(e->err, sizeof(e->err), "%s", msg);
    e->error = code;
    return code;
}

static void* dupmem(const void* src, size_t bytes)
{
    void* d = malloc(bytes ? bytes : 1);
    if (src && bytes) memcpy(d, src, bytes);
    else memset(d, 0, bytes ? bytes : 1);
    return d;
}

/* ---- kernel lookup (src/Kernels.cpp:108-129) ---- */
static double k_num(or_engine* e, double r2)
{
    const cs_params* P = &e->P;
    double t = r2 / P->kernel_delta;
    if (t > (double)P->kernel_size) {
        e->error = CS_ERANGE;
        snprintf(e->err, sizeof(e->err), "r too large in NumKernel (r2=%g)", r2);
        return 0.0;
    }
    double s = t * (double)P->kernel_expansion;
    int sz = P->kernel_size;
    if (s < (double)sz) {
        t = s - floor(s);
        int i0 = (int)s, i1 = (int)(s + 1.0);
        if (i1 > sz) i1 = sz;
        t = (1.0 - t) * e->kh[i0] + t * e->kh[i1];
    } else {
        s = t - floor(t);
        int i0 = (int)t, i1 = (int)(t + 1.0);
        if (i1 > sz) i1 = sz;
        t = (1.0 - s) * e->kl[i0] + s * e->kl[i1];
    }
    return t;
}

/* ---- distances (src/Dist.cpp:14-168) ---- */
static double periodic_xy(const cs_params* P, double x, double y)
{
    if (P->do_periodic_boundaries) {
        if (x > P->degrees_width * 0.5) x = P->degrees_width - x;
        if (y > P->degrees_height * 0.5) y = P->degrees_height - y;
    }
    return x * x + y * y;
}

static double dist2_utm(const cs_params* P, double x1, double y1, double x2, double y2)
{
    double x, y, cy1, cy2, yt, xi, yi;
    x = fabs(x1 - x2) / 2;
    y = fabs(y1 - y2) / 2;
    xi = floor(x);
    yi = floor(y);
    x -= xi;
    y -= yi;
    x = (1 - x) * P->sinx[(int)xi] + x * P->sinx[((int)xi) + 1];
    y = (1 - y) * P->sinx[(int)yi] + y * P->sinx[((int)yi) + 1];
    yt = fabs(y1 + P->bbox_bottom_left_y);
    yi = floor(yt);
    cy1 = yt - yi;
    cy1 = (1 - cy1) * P->cosx[((int)yi)] + cy1 * P->cosx[((int)yi) + 1];
    yt = fabs(y2 + P->bbox_bottom_left_y);
    yi = floor(yt);
    cy2 = yt - yi;
    cy2 = (1 - cy2) * P->cosx[((int)yi)] + cy2 * P->cosx[((int)yi) + 1];
    x = fabs(1000 * (y * y + x * x * cy1 * cy2));
    xi = floor(x);
    x -= xi;
    y = (1 - x) * P->asin2sqx[((int)xi)] + x * P->asin2sqx[((int)xi) + 1];
    /* 4 * EARTHRADIUS^2, EARTHRADIUS = (1852*60*360/3.1415926535)/2 (src/Constants.h) */
    const double er = ((double)(1852 * 60 * 360) / 3.1415926535) / 2;
    return 4 * er * er * y;
}

static double dist2_pp(or_engine* e, int a, int b)
{
    const cs_params* P = &e->P;
    const float* la = e->hloc + 2 * (size_t)e->hh[a];
    const float* lb = e->hloc + 2 * (size_t)e->hh[b];
    if (P->do_utm_coords) return dist2_utm(P, la[0], la[1], lb[0], lb[1]);
    double x = fabs((double)la[0] - (double)lb[0]);
    double y = fabs((double)la[1] - (double)lb[1]);
    return periodic_xy(P, x, y);
}

static double dist2_cc_min(const cs_params* P, int l, int m)
{
    int nch = P->nch, i = l, j = m;
    if (P->do_utm_coords) {
        const double PI_ = 3.1415926535;
        if (P->cell_width * ((double)abs(m / nch - l / nch)) > PI_) {
            if (m / nch > l / nch) j += nch;
            else if (m / nch < l / nch) i += nch;
        } else {
            if (m / nch > l / nch) i += nch;
            else if (m / nch < l / nch) j += nch;
        }
        if (m % nch > l % nch) i++;
        else if (m % nch < l % nch) j++;
        return dist2_utm(P, P->cell_width * fabs((double)(i / nch)), P->cell_height * fabs((double)(i % nch)),
                         P->cell_width * fabs((double)(j / nch)), P->cell_height * fabs((double)(j % nch)));
    }
    if (P->do_periodic_boundaries && (P->cell_width * ((double)abs(m / nch - l / nch)) > P->degrees_width * 0.5)) {
        if (m / nch > l / nch) j += nch;
        else if (m / nch < l / nch) i += nch;
    } else {
        if (m / nch > l / nch) i += nch;
        else if (m / nch < l / nch) j += nch;
    }
    if (P->do_periodic_boundaries && (P->degrees_height * ((double)abs(m % nch - l % nch)) > P->degrees_height * 0.5)) {
        if (m % nch > l % nch) j++;
        else if (m % nch < l % nch) i++;
    } else {
        if (m % nch > l % nch) i++;
        else if (m % nch < l % nch) j++;
    }
    double x = P->cell_width * fabs((double)(i / nch - j / nch));
    double y = P->cell_height * fabs((double)(i % nch - j % nch));
    return periodic_xy(P, x, y);
}

/* ---- InverseCdf::choose with a given uniform (src/InverseCdf.cpp:60-74) ---- */
static uint16_t icdf_choose(const double* c, double mean, double u, double tspd)
{
    double q;
    int i = (int)floor(q = u * CS_CDF_RES);
    q -= (double)i;
    double ti = -mean * or_log(q * c[i + 1] + (1.0 - q) * c[i]);
    return (uint16_t)floor(0.5 + (ti * tspd));
}

/* ---- state predicates (src/ModelMacros.h:6-23, src/Person.cpp) ---- */
#define AGE_GROUP(e, x) ((int)(e)->age[x] / CS_AGE_GROUP_WIDTH)
static int is_susceptible(const or_real* r, int x) { return r->inf[x] == CS_SUSCEPTIBLE; }
static int is_dead(const or_real* r, int x) { return r->inf[x] == CS_DEAD_ASYMP || r->inf[x] == CS_DEAD_SYMP; }
static int is_case(const or_real* r, int x) { return r->inf[x] == CS_CASE; }
static int host_absent(const or_real* r, int x, unsigned ts) { return r->t_as[x] <= ts && r->t_ae[x] > ts; }
static int place_closed(const or_real* r, int t, int k, unsigned ts) { return r->pl_cs[t][k] <= ts && r->pl_ce[t][k] > ts; }
static int host_treated(const or_engine* e, const or_real* r, int x, unsigned ts)
{
    return e->P.interventions_active && r->t_te[x] > ts && r->t_ts[x] <= ts;
}
static int host_vacced(const or_engine* e, const or_real* r, int x, unsigned ts)
{
    return e->P.interventions_active && (int)r->t_vs[x] + (int)e->P.us_vacc_time_to_efficacy <= (int)ts;
}
static int host_vacced_switch(const or_engine* e, const or_real* r, int x)
{
    return r->t_vs[x] >= e->P.us_vacc_time_efficacy_switch;
}
static int host_quarantined(const or_engine* e, const or_real* r, int x, unsigned ts)
{
    return e->P.interventions_active && r->qc[x] == 1 &&
           (int)r->t_qs[x] + (int)e->P.us_hquarantine_house_duration > (int)ts && r->t_qs[x] <= ts;
}
static int host_isolated(const or_engine* e, const or_real* r, int x, unsigned ts)
{
    return e->P.interventions_active &&
           (int)r->t_is[x] + (int)e->P.us_case_isolation_delay <= (int)ts &&
           (int)r->t_is[x] + (int)e->P.us_case_isolation_delay + (int)e->P.us_case_isolation_duration > (int)ts;
}
static int dct_traced(const or_engine* e, const or_real* r, int x) { return e->P.interventions_active && r->dct[x] == 1; }
static int socdist_treated(const or_engine* e, const or_real* r, int x)
{
    return e->P.interventions_active && r->m_socdist[e->mcell[x]] == 2;
}
#define CARE(e, x) (((e)->flags[x] & 1) != 0)
#define KEYW(e, x) (((e)->flags[x] >> 1) & 1)
#define ESOC(e, x) (((e)->flags[x] & 4) != 0)

/* ---- hazards (src/CalcInfSusc.cpp:12-80) ---- */
static double inf_profile(const or_engine* e, const or_real* r, int x, unsigned ts)
{
    int idx = (int)ts - (int)r->t_lat[x] - 1;
    if (idx < 0) idx = 0;
    if (idx >= CS_MAX_INFECTIOUS_STEPS) idx = CS_MAX_INFECTIOUS_STEPS - 1;
    return e->P.infectiousness[idx];
}

static double calc_person_inf(const or_engine* e, const or_real* r, int x, unsigned ts)
{
    const cs_params* P = &e->P;
    return (host_treated(e, r, x, ts) ? P->treat_inf_drop : 1.0)
         * (host_vacced(e, r, x, ts) ? P->vacc_inf_drop : 1.0)
         * fabs((double)r->infness[x])
         * inf_profile(e, r, x, ts);
}

static double calc_house_inf(const or_engine* e, const or_real* r, int x, unsigned ts)
{
    const cs_params* P = &e->P;
    int iso = host_isolated(e, r, x, ts), dct = dct_traced(e, r, x), qua = host_quarantined(e, r, x, ts);
    return ((iso && !dct) ? P->efficacies[0][CS_HOUSE] : 1.0)
         * (dct ? P->efficacies[5][CS_HOUSE] : 1.0)
         * ((qua && !dct && !iso) ? P->efficacies[1][CS_HOUSE] : 1.0)
         * P->household_denom_lookup[e->hnhr[e->hh[x]] - 1]
         * (CARE(e, x) ? P->care_home_resident_household_scaling : 1.0)
         * (host_treated(e, r, x, ts) ? P->treat_inf_drop : 1.0)
         * (host_vacced(e, r, x, ts) ? P->vacc_inf_drop : 1.0)
         * (P->no_infectiousness_sd_in_hh ? ((r->infness[x] < 0) ? P->sympt_infectiousness : P->asympt_infectiousness)
                                          : fabs((double)r->infness[x]))
         * inf_profile(e, r, x, ts);
}

static double calc_place_inf(const or_engine* e, const or_real* r, int x, int pt, unsigned ts)
{
    const cs_params* P = &e->P;
    int iso = host_isolated(e, r, x, ts), dct = dct_traced(e, r, x), qua = host_quarantined(e, r, x, ts);
    return ((iso && !dct) ? P->efficacies[0][pt] : 1.0)
         * (dct ? P->efficacies[5][pt] : 1.0)
         * ((qua && !CARE(e, x) && !dct && !iso) ? P->efficacies[1][pt] : 1.0)
         * ((is_case(r, x) && !CARE(e, x)) ? P->sympt_place_type_contact_rate[pt] : 1.0)
         * P->place_type_trans[pt] / P->place_type_group_size_param1[pt] * calc_person_inf(e, r, x, ts);
}

static double calc_spatial_inf(const or_engine* e, const or_real* r, int x, unsigned ts)
{
    const cs_params* P = &e->P;
    int iso = host_isolated(e, r, x, ts), dct = dct_traced(e, r, x), qua = host_quarantined(e, r, x, ts);
    return ((iso && !dct) ? P->efficacies[0][CS_SPATIAL] : 1.0)
         * (dct ? P->efficacies[5][CS_SPATIAL] : 1.0)
         * ((qua && !CARE(e, x) && !dct && !iso) ? P->efficacies[1][CS_SPATIAL] : 1.0)
         * (is_case(r, x) ? P->sympt_spatial_contact_rate : 1.0)
         * P->relative_spatial_contact[AGE_GROUP(e, x)]
         * calc_person_inf(e, r, x, ts);
}

static double calc_person_susc(const or_engine* e, const or_real* r, int x, unsigned ts, int infector)
{
    const cs_params* P = &e->P;
    return P->waifw[AGE_GROUP(e, x)][AGE_GROUP(e, infector)]
         * P->age_susceptibility[AGE_GROUP(e, x)] * (double)r->susc[x]
         * (host_treated(e, r, x, ts) ? P->treat_susc_drop : 1.0)
         * (host_vacced(e, r, x, ts) ? (host_vacced_switch(e, r, x) ? P->vacc_susc_drop2 : P->vacc_susc_drop) : 1.0);
}

static double calc_house_susc(const or_engine* e, const or_real* r, int x, unsigned ts, int infector)
{
    const cs_params* P = &e->P;
    return calc_person_susc(e, r, x, ts, infector)
         * (socdist_treated(e, r, x) ? (ESOC(e, x) ? P->efficacies[4][CS_HOUSE] : P->efficacies[3][CS_HOUSE]) : 1.0)
         * (dct_traced(e, r, x) ? P->efficacies[5][CS_HOUSE] : 1.0)
         * (CARE(e, x) ? P->care_home_resident_household_scaling : 1.0);
}

static double calc_place_susc(const or_engine* e, const or_real* r, int x, int pt, unsigned ts)
{
    const cs_params* P = &e->P;
    int dct = dct_traced(e, r, x);
    return ((host_quarantined(e, r, x, ts) && !CARE(e, x) && !dct) ? P->efficacies[1][pt] : 1.0)
         * (socdist_treated(e, r, x) ? (ESOC(e, x) ? P->efficacies[4][pt] : P->efficacies[3][pt]) : 1.0)
         * (dct ? P->efficacies[5][pt] : 1.0);
}

static double calc_spatial_susc(const or_engine* e, const or_real* r, int x, unsigned ts)
{
    const cs_params* P = &e->P;
    int dct = dct_traced(e, r, x);
    return ((host_quarantined(e, r, x, ts) && !CARE(e, x) && !dct) ? P->efficacies[1][CS_SPATIAL] : 1.0)
         * (socdist_treated(e, r, x) ? (ESOC(e, x) ? P->efficacies[4][CS_SPATIAL] : P->efficacies[3][CS_SPATIAL]) : 1.0)
         * (dct ? P->efficacies[5][CS_SPATIAL] : 1.0)
         * P->relative_spatial_contact_susc[AGE_GROUP(e, x)];
}

/* ---------------------------------------------------------- set-up */

/* KernelStruct families (src/Kernels.cpp:68-106), evaluated with libm as
 * the reference does, so the tables below match KernelLookup::init
 * (src/Kernels.cpp:17-45, norm = 1.0 as SetupModel calls it) bit for bit. */
static double kernel_fn(const cs_params* P, double r2)
{
    const double sh = P->kernel_shape, sc = P->kernel_scale, p3 = P->kernel_p3, p4 = P->kernel_p4;
    double t, d;
    switch (P->kernel_type) {
    case 1: return exp(-sqrt(r2) / sc);
    case 2: t = -sh * log(sqrt(r2) / sc + 1.0); return (t < -690.0) ? 0.0 : exp(t);
    case 3: return exp(-r2 / (sc * sc));
    case 4: return (r2 > sc * sc) ? 0.0 : 1.0;
    case 5: t = 0.5 * sh * log(r2 / (sc * sc)); return (t > 690.0) ? 0.0 : (1.0 / (exp(t) + 1.0));
    case 6: t = log(sqrt(r2) / sc + 1.0); return (t < -690.0) ? 0.0 : (exp(-sh * t) + p3 * exp(-p4 * t)) / (1.0 + p3);
    case 7: d = sqrt(r2); t = -sh * log(d / sc + 1.0); return (t < -690.0) ? 0.0 : exp(t - pow(d / p3, p4));
    default: return NAN;
    }
}

int or_create(or_engine** out, const cs_params* p, const double* kl, const double* kh, const cs_population* pop)
{
    or_engine* e = (or_engine*)calloc(1, sizeof(or_engine));
    *out = e;
    e->P = *p;
    const cs_params* P = &e->P;
    size_t ksz = (size_t)P->kernel_size + 1;
    if (P->kernel_type < 1 || P->kernel_type > 7) return fail(e, CS_EINVAL, "unknown kernel type");
    e->kl = (double*)dupmem(kl, ksz * 8);
    e->kh = (double*)dupmem(kh, ksz * 8);
    if (!kl || !kh)
        for (int i = 0; i <= P->kernel_size; i++) {
            if (!kl) e->kl[i] = kernel_fn(P, i * P->kernel_delta) / 1.0;
            if (!kh) e->kh[i] = kernel_fn(P, i * P->kernel_delta / P->kernel_expansion) / 1.0;
        }
    e->N = pop->n_people; e->NH = pop->n_households; e->NC = pop->n_cells; e->NCT = pop->n_cells_total;
    e->NM = pop->n_mcells; e->NA = pop->n_adunits; e->NPT = P->num_place_types;
    int N = e->N, NC = e->NC;
    e->hh = (int32_t*)dupmem(pop->hh, N * 4);
    e->mcell = (int32_t*)dupmem(pop->mcell, N * 4);
    e->age = (uint8_t*)dupmem(pop->age, N);
    e->flags = (uint8_t*)dupmem(pop->flags, N);
    e->plinks = (int32_t*)dupmem(pop->place_links, (size_t)N * CS_MAX_PLACE_TYPES * 4);
    e->pglinks = (uint16_t*)dupmem(pop->place_group_links, (size_t)N * CS_MAX_PLACE_TYPES * 2);
    e->hfirst = (int32_t*)dupmem(pop->hh_first, e->NH * 4);
    e->hnh = (uint16_t*)dupmem(pop->hh_nh, e->NH * 2);
    e->hnhr = (uint16_t*)dupmem(pop->hh_nhr, e->NH * 2);
    e->hloc = (float*)dupmem(pop->hh_loc, (size_t)e->NH * 8);
    e->cidx = (int32_t*)dupmem(pop->cell_index, NC * 4);
    e->cn = (int32_t*)dupmem(pop->cell_n, NC * 4);
    e->cstart = (int32_t*)dupmem(pop->cell_member_start, (NC + 1) * 4);
    e->cmem = (int32_t*)dupmem(pop->cell_members, (size_t)N * 4);
    e->madunit = (int32_t*)dupmem(pop->mcell_adunit, e->NM * 4);
    /* populated-cell index of every person (Person.pcell is a Cells[] index) */
    int32_t* map = (int32_t*)malloc(sizeof(int32_t) * (size_t)(e->NCT > 0 ? e->NCT : 1));
    for (int i = 0; i < e->NCT; i++) map[i] = -1;
    for (int j = 0; j < NC; j++) map[e->cidx[j]] = j;
    e->pc = (int32_t*)malloc((size_t)N * 4);
    for (int i = 0; i < N; i++) e->pc[i] = map[pop->pcell[i]];
    free(map);
    /* microcell member lists (ascending person id) */
    e->mstart = (int32_t*)calloc((size_t)e->NM + 1, 4);
    e->mmem = (int32_t*)malloc((size_t)N * 4);
    for (int i = 0; i < N; i++) e->mstart[e->mcell[i] + 1]++;
    for (int m = 0; m < e->NM; m++) e->mstart[m + 1] += e->mstart[m];
    {
        int32_t* cur = (int32_t*)malloc((size_t)e->NM * 4 + 4);
        memcpy(cur, e->mstart, (size_t)e->NM * 4);
        for (int i = 0; i < N; i++) e->mmem[cur[e->mcell[i]]++] = i;
        free(cur);
    }
    int maxplace = 1;
    for (int t = 0; t < e->NPT; t++) {
        int np = pop->n_places[t];
        e->np[t] = np;
        e->pl_n[t] = (int32_t*)dupmem(pop->place_n[t], np * 4);
        e->pl_mcell[t] = (int32_t*)dupmem(pop->place_mcell[t], np * 4);
        e->pl_ms[t] = (int32_t*)dupmem(pop->place_member_start[t], (np + 1) * 4);
        e->pl_mem[t] = (int32_t*)dupmem(pop->place_members[t], (size_t)e->pl_ms[t][np] * 4);
        e->pl_gs0[t] = (int32_t*)dupmem(pop->place_groups_start[t], (np + 1) * 4);
        e->pl_gst[t] = (int32_t*)dupmem(pop->place_group_start[t], (size_t)e->pl_gs0[t][np] * 4);
        e->pl_gsz[t] = (int32_t*)dupmem(pop->place_group_size[t], (size_t)e->pl_gs0[t][np] * 4);
        for (int k = 0; k < np; k++) if (e->pl_n[t][k] > maxplace) maxplace = e->pl_n[t][k];
    }
    e->samp = (int32_t*)malloc(sizeof(int32_t) * (size_t)maxplace);
    /* the places of each microcell (Microcell.places), by type then index */
    e->mpl_start = (int32_t*)calloc((size_t)e->NM + 1, 4);
    {
        int tot = 0;
        for (int t = 0; t < e->NPT; t++)
            for (int k = 0; k < e->np[t]; k++) { e->mpl_start[e->pl_mcell[t][k] + 1]++; tot++; }
        for (int m = 0; m < e->NM; m++) e->mpl_start[m + 1] += e->mpl_start[m];
        e->mpl = (int32_t*)malloc(sizeof(int32_t) * (size_t)(tot > 0 ? tot : 1));
        int32_t* cur = (int32_t*)malloc((size_t)e->NM * 4 + 4);
        memcpy(cur, e->mpl_start, (size_t)e->NM * 4);
        for (int t = 0; t < e->NPT; t++)
            for (int k = 0; k < e->np[t]; k++) e->mpl[cur[e->pl_mcell[t][k]]++] = (t << 27) | k;
        free(cur);
    }
    int maxcell = 1;
    for (int j = 0; j < NC; j++) if (e->cn[j] > maxcell) maxcell = e->cn[j];
    e->cell_inf = (float*)malloc(sizeof(float) * (size_t)maxcell);
    /* cell max_trans (src/Kernels.cpp:48-63): (float)num(dist2_cc_min(l,m)) */
    e->max_trans = (float*)malloc(sizeof(float) * (size_t)NC * NC);
    for (int l = 0; l < NC; l++)
        for (int m = 0; m < NC; m++)
            e->max_trans[(size_t)l * NC + m] = (float)k_num(e, dist2_cc_min(P, e->cidx[l], e->cidx[m]));
    if (e->error) return e->error;
    return 0;
}

static void free_real(or_engine* e, or_real* r)
{
    void* ptrs[] = {r->inf, r->sev_cur, r->sev_final, r->bits, r->qc, r->dct, r->hh_done, r->listpos, r->infector, r->itype,
        r->infness, r->susc, r->prob_care, r->t_inf, r->t_lat, r->t_rod, r->t_sari, r->t_crit, r->t_step, r->t_det,
        r->t_as, r->t_ae, r->t_ts, r->t_te, r->t_vs, r->t_is, r->t_qs, r->cS, r->cL, r->cI, r->cR, r->cD, r->cumTC,
        r->S0, r->tot_prob, r->cum_trans, r->invcdf, r->perm, r->m_socdist, r->m_moverest, r->m_trig, r->ad_cnt};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    for (int t = 0; t < CS_MAX_PLACE_TYPES; t++) { free(r->pl_cs[t]); free(r->pl_ce[t]); free(r->pl_once[t]); }
    free(r->hist); free(r->mc_state); free(r->mc_end);
    (void)e;
}

void or_destroy(or_engine* e)
{
    if (!e) return;
    for (int i = 0; i < e->nreal; i++) free_real(e, &e->R[i]);
    free(e->R);
    free(e->kl); free(e->kh); free(e->hh); free(e->mcell); free(e->pc); free(e->age); free(e->flags);
    free(e->plinks); free(e->pglinks); free(e->hfirst); free(e->hnh); free(e->hnhr); free(e->hloc);
    free(e->cidx); free(e->cn); free(e->cstart); free(e->cmem); free(e->madunit); free(e->mstart); free(e->mmem);
    for (int t = 0; t < e->NPT; t++) {
        free(e->pl_n[t]); free(e->pl_mcell[t]); free(e->pl_ms[t]); free(e->pl_mem[t]);
        free(e->pl_gs0[t]); free(e->pl_gst[t]); free(e->pl_gsz[t]);
    }
    free(e->samp); free(e->cell_inf); free(e->max_trans); free(e->ev); free(e->ts_buf); free(e->ts_last);
    free(e->ts_ext); free(e->ext_last);
    free(e->mpl_start); free(e->mpl);
    free(e);
}

int or_alloc_realisations(or_engine* e, int n_real, const uint64_t* seeds)
{
    for (int i = 0; i < e->nreal; i++) free_real(e, &e->R[i]);
    free(e->R);
    free(e->ts_buf); free(e->ts_last); free(e->ts_ext); free(e->ext_last);
    e->ts_buf = NULL; e->ts_last = NULL; e->ts_ext = NULL; e->ext_last = NULL;
    e->ts_cap = e->ts_n = 0; e->ts_fresh = 1;
    {   /* units are counted when known for every person */
        int ok = e->P.do_ad_units && e->NA > 0 && e->madunit != NULL;
        for (int i = 0; i < e->N && ok; i++) {
            const int u = e->madunit[e->mcell[i]];
            if (u < 0 || u >= e->NA) ok = 0;
        }
        e->NAU = ok ? e->NA : 0;
        e->ext_w = CS_TSA_FIELDS * CS_NUM_AGE_GROUPS + CS_TSU_FIELDS * e->NAU;
    }
    e->nreal = n_real;
    e->R = (or_real*)calloc((size_t)n_real, sizeof(or_real));
    int N = e->N, NC = e->NC;
    for (int i = 0; i < n_real; i++) {
        or_real* r = &e->R[i];
        r->seed = seeds[i];
        r->n_seed = -1;
#define AL(f, n, T) r->f = (T*)calloc((size_t)(n), sizeof(T))
        AL(inf, N, int8_t); AL(sev_cur, N, uint8_t); AL(sev_final, N, uint8_t); AL(bits, N, uint8_t);
        AL(qc, N, uint8_t); AL(dct, N, uint8_t); AL(hh_done, N, uint8_t); AL(listpos, N, int32_t); AL(infector, N, int32_t);
        AL(itype, N, int16_t); AL(infness, N, float); AL(susc, N, float); AL(prob_care, N, float);
        AL(t_inf, N, uint16_t); AL(t_lat, N, uint16_t); AL(t_rod, N, uint16_t); AL(t_sari, N, uint16_t);
        AL(t_crit, N, uint16_t); AL(t_step, N, uint16_t); AL(t_det, N, uint16_t); AL(t_as, N, uint16_t);
        AL(t_ae, N, uint16_t); AL(t_ts, N, uint16_t); AL(t_te, N, uint16_t); AL(t_vs, N, uint16_t);
        AL(t_is, N, uint16_t); AL(t_qs, N, uint16_t);
        AL(cS, NC, int32_t); AL(cL, NC, int32_t); AL(cI, NC, int32_t); AL(cR, NC, int32_t); AL(cD, NC, int32_t);
        AL(cumTC, NC, int32_t); AL(S0, NC, int32_t); AL(tot_prob, NC, float);
        AL(cum_trans, (size_t)NC * NC, float); AL(invcdf, (size_t)NC * 1025, int32_t); AL(perm, N, int32_t);
        for (int t = 0; t < e->NPT; t++) { AL(pl_cs[t], e->np[t], uint16_t); AL(pl_ce[t], e->np[t], uint16_t); AL(pl_once[t], e->np[t], uint8_t); }
        AL(hist, (size_t)OR_TRIG_RING * 3, int32_t); AL(mc_state, e->NM, uint8_t); AL(mc_end, e->NM, uint16_t);
        AL(m_socdist, e->NM, uint8_t); AL(m_moverest, e->NM, uint8_t); AL(m_trig, (size_t)e->NM * 6, uint32_t);
        AL(ad_cnt, (size_t)(e->NAU ? e->NAU : 1) * CS_TSU_FIELDS, int64_t);
#undef AL
    }
    return 0;
}

/* UpdateProbs(0) for one realisation (src/CovidSim.cpp:2752-2802) */
static void update_probs_real(or_engine* e, or_real* r)
{
    const int NC = e->NC;
    for (int j = 0; j < NC; j++) {
        r->tot_prob[j] = 0;
        r->S0[j] = r->cS[j] + r->cL[j] + r->cI[j];
        if (e->P.do_death) {
            r->S0[j] += e->cn[j] / 5;
            if ((e->cn[j] < 100) || (r->S0[j] > e->cn[j])) r->S0[j] = e->cn[j];
        }
    }
    for (int j = 0; j < NC; j++) {
        float* ct = r->cum_trans + (size_t)j * NC;
        const float* mt = e->max_trans + (size_t)j * NC;
        float t;
        ct[0] = ((float)r->S0[0]) * mt[0];
        t = ((float)e->cn[0]) * mt[0];
        for (int m = 1; m < NC; m++) {
            ct[m] = ct[m - 1] + ((float)r->S0[m]) * mt[m];
            t += ((float)e->cn[m]) * mt[m];
        }
        r->tot_prob[j] = ct[NC - 1];
        for (int m = 0; m < NC; m++) ct[m] /= r->tot_prob[j];
        r->tot_prob[j] /= t;
        int32_t* inv = r->invcdf + (size_t)j * 1025;
        int m = 0;
        for (int k = 0; k <= 1024; k++) {
            while (ct[m] * 1024 < ((float)k)) m++;
            inv[k] = m;
        }
    }
}

int or_update_probs(or_engine* e)
{
    for (int i = 0; i < e->nreal; i++) update_probs_real(e, &e->R[i]);
    return 0;
}

static int64_t sum_S(const or_engine* e, const or_real* r)
{
    int64_t s = 0;
    for (int j = 0; j < e->NC; j++) s += r->cS[j];
    return s;
}

/* ---- the batch list move that replaces UpdateCell's one-at-a-time swaps ----
 * (src/Update.cpp:1391-1578).  X (n persons, all in the cell-local segment
 * [lo,hi) of the cell whose permutation starts at perm[base]) moves to the
 * tail of the segment.  listpos is cell-local, as Person.listpos.
 * Deterministic for a given set X. */
static int cmp_i32(const void* a, const void* b)
{
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

static void move_to_tail(or_real* r, int base, int lo, int hi, int32_t* X, int n, int32_t* pos)
{
    (void)lo;
    if (n <= 0) return;
    for (int i = 0; i < n; i++) pos[i] = r->listpos[X[i]];
    qsort(pos, (size_t)n, sizeof(int32_t), cmp_i32);
    int bnd = hi - n;
    int nh = 0;
    while (nh < n && pos[nh] < bnd) nh++;
    /* fillers: positions in [bnd,hi) not in pos[nh..n) */
    int pi = nh, hole = 0;
    for (int q = bnd; q < hi && hole < nh; q++) {
        if (pi < n && pos[pi] == q) { pi++; continue; }
        int a = pos[hole++];
        int pa = r->perm[base + a], pb = r->perm[base + q];
        r->perm[base + a] = pb; r->listpos[pb] = a;
        r->perm[base + q] = pa; r->listpos[pa] = q;
    }
}

/* ------------------------------------------------------ transitions */

static void sev_to(or_real* r, int kind, int ag) { r->cnt.sev_prev[kind]++; r->cnt.sev_prev_age[kind][ag]++; r->cnt.sev_cum[kind]++; r->cnt.sev_cum_age[kind][ag]++; }
static void sev_from(or_real* r, int kind, int ag) { r->cnt.sev_prev[kind]--; r->cnt.sev_prev_age[kind][ag]--; }
static void sev_cum_only(or_real* r, int kind, int ag) { r->cnt.sev_cum[kind]++; r->cnt.sev_cum_age[kind][ag]++; }

/* DoInfect (src/Update.cpp:100-195) without the cell-list swap, which the
 * caller batches.  infector/itype hold the first event's values. */
static void do_infect(or_engine* e, or_real* r, int p, double t, unsigned ts)
{
    const cs_params* P = &e->P;
    if (!is_susceptible(r, p)) return;
    r->inf[p] = CS_LATENT;
    r->t_inf[p] = (uint16_t)ts;
    const float* loc = e->hloc + 2 * (size_t)e->hh[p];
    double x = (loc[0] - r->loc_init[0]);
    double y = (loc[1] - r->loc_init[1]);
    double rad2 = x * x + y * y;
    int ag = AGE_GROUP(e, p);
    r->cnt.cumI++;
    r->cnt.cumItype[r->itype[p] % CS_INFECT_TYPE_MASK]++;
    r->cnt.cumIa[ag]++;
    if (e->NAU) r->ad_cnt[(size_t)e->madunit[e->mcell[p]] * CS_TSU_FIELDS + CS_TSU_INC_I]++; /* cumI_adunit (src/Update.cpp:159) */
    r->cnt.sumRad2 += rad2;
    if (rad2 > r->cnt.maxRad2) r->cnt.maxRad2 = rad2;
    r->cnt.cumI_keyworker[KEYW(e, p)]++;
    if (P->do_latent) {
        rs s;
        rs_init(&s, r->seed, cw0(ts, PS_INFECT), (uint32_t)p, 0);
        double q;
        int i = (int)floor((q = rs_next(&s) * CS_CDF_RES));
        q -= ((double)i);
        const double* c = P->icdf[CS_ICDF_LATENT];
        r->t_lat[p] = (uint16_t)floor(0.5 + (t - P->latent_period * or_log(q * c[i + 1] + (1.0 - q) * c[i])) * P->time_steps_per_day);
    } else
        r->t_lat[p] = (uint16_t)(t * P->time_steps_per_day);
    if (r->infector[p] >= 0) {
        r->cnt.cumTG += ((int)r->t_inf[p]) - ((int)r->t_inf[r->infector[p]]);
        r->cnt.cumSI += ((int)r->t_lat[p]) - ((int)r->t_lat[r->infector[p]]);
        r->cnt.nTG++;
    }
}

/* batch S -> L moves for the persons in list[0..n) (any order) */
static void batch_move(or_engine* e, or_real* r, int32_t* list, int n, int seg)
{
    /* group by cell with a counting pass */
    int NC = e->NC;
    int32_t* cnt = (int32_t*)calloc((size_t)NC + 1, 4);
    for (int i = 0; i < n; i++) cnt[e->pc[list[i]] + 1]++;
    for (int c = 0; c < NC; c++) cnt[c + 1] += cnt[c];
    int32_t* by = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)NC + 4);
    memcpy(cur, cnt, sizeof(int32_t) * (size_t)NC);
    for (int i = 0; i < n; i++) by[cur[e->pc[list[i]]]++] = list[i];
    int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    for (int c = 0; c < NC; c++) {
        int k = cnt[c + 1] - cnt[c];
        if (!k) continue;
        int base = e->cstart[c];
        if (seg == 0) { /* S -> L */
            move_to_tail(r, base, 0, r->cS[c], by + cnt[c], k, pos);
            r->cS[c] -= k; r->cL[c] += k;
        } else if (seg == 1) { /* L -> I */
            move_to_tail(r, base, r->cS[c], r->cS[c] + r->cL[c], by + cnt[c], k, pos);
            r->cL[c] -= k; r->cI[c] += k;
        } else { /* I -> R or D (state already set) */
            int lo = r->cS[c] + r->cL[c];
            move_to_tail(r, base, lo, lo + r->cI[c], by + cnt[c], k, pos);
            r->cI[c] -= k;
            for (int i = 0; i < k; i++) {
                if (is_dead(r, by[cnt[c] + i])) r->cD[c]++; else r->cR[c]++;
            }
        }
    }
    free(cnt); free(by); free(cur); free(pos);
}

static int cmp_event(const void* a, const void* b)
{
    const or_event* x = (const or_event*)a;
    const or_event* y = (const or_event*)b;
    if (x->infectee != y->infectee) return (x->infectee > y->infectee) - (x->infectee < y->infectee);
    return (x->key > y->key) - (x->key < y->key);
}

static void add_event(or_engine* e, uint64_t key, int infectee, int infector, int type)
{
    if (e->nev == e->cap_ev) {
        e->cap_ev = e->cap_ev ? 2 * e->cap_ev : 4096;
        e->ev = (or_event*)realloc(e->ev, e->cap_ev * sizeof(or_event));
    }
    or_event* v = &e->ev[e->nev++];
    v->key = key; v->infectee = infectee; v->infector = infector; v->type = (int16_t)type;
}

static uint64_t ev_key(int infector, int pass, uint32_t sub)
{
    return ((uint64_t)(uint32_t)infector << 32) | ((uint64_t)pass << 27) | (uint64_t)(sub & 0x7ffffff);
}

/* the reference's infector binary search (src/Sweep.cpp:716-740) */
static int select_infector_n(const float* ci, int n, double u, int64_t* probes)
{
    if (n == 1) return 0;
    int idx, step;
    idx = step = n / 2;
    int keep = 1;
    do {
        if (probes) (*probes)++;
        if (step > 1) step /= 2;
        if ((idx > 0) && (fabs(ci[idx - 1]) >= u)) {
            idx -= step;
            if (idx == 0) keep = 0;
        } else if ((idx < n - 1) && (fabs(ci[idx]) < u)) {
            idx += step;
            if (idx == n - 1) keep = 0;
        } else keep = 0;
    } while (keep);
    return idx;
}
static int select_infector(const float* ci, int n, double u) { return select_infector_n(ci, n, u, NULL); }

/* InvCDF bucket, then the linear scan of cum_trans (src/Sweep.cpp:757-762);
 * *scanned counts the cum_trans entries read */
static int select_dest_n(const or_engine* e, const or_real* r, int c, double u, int64_t* scanned)
{
    int l = r->invcdf[(size_t)c * 1025 + (int)floor(u * 1024)];
    const float* ct = r->cum_trans + (size_t)c * e->NC;
    const int l0 = l;
    while (ct[l] < u) l++;
    if (scanned) *scanned += l - l0 + 1;
    return l;
}
static int select_dest(const or_engine* e, const or_real* r, int c, double u) { return select_dest_n(e, r, c, u, NULL); }

/* resolve this step's events of realisation r (src/Sweep.cpp:860-880) */
static void resolve(or_engine* e, or_real* r, double t, unsigned ts)
{
    qsort(e->ev, e->nev, sizeof(or_event), cmp_event);
    int32_t* newl = (int32_t*)malloc(sizeof(int32_t) * (e->nev ? e->nev : 1));
    int nn = 0;
    size_t g0 = 0;
    while (g0 < e->nev) {
        size_t g1 = g0 + 1;
        while (g1 < e->nev && e->ev[g1].infectee == e->ev[g0].infectee) g1++;
        int p = e->ev[g0].infectee;
        if (is_susceptible(r, p)) {
            r->infector[p] = e->ev[g0].infector;
            r->itype[p] = e->ev[g0].type;
            do_infect(e, r, p, t, ts);
            newl[nn++] = p;
        }
        r->infector[p] = e->ev[g1 - 1].infector;
        r->itype[p] = e->ev[g1 - 1].type;
        g0 = g1;
    }
    e->wc.n_new += nn;
    batch_move(e, r, newl, nn, 0);
    free(newl);
}

static int infect_sweep(or_engine* e, double t)
{
    const cs_params* P = &e->P;
    if (P->false_positive_rate != 0.0) return fail(e, CS_EUNSUPPORTED, "false positives are outside this path");
    unsigned ts = (unsigned)(uint16_t)(P->time_steps_per_day * t);
    double fp = P->model_time_step / (1 - P->false_positive_rate);
    double seas = P->do_seasonality ? P->seasonality[((int)t) % CS_DAYS_PER_YEAR] : 1.0;
    double beta_s_all = seas * fp * P->local_beta;
    double beta_h = P->do_households ? seas * fp * P->household_trans : 0;
    const int NPT = e->NPT;
    for (int ri = 0; ri < e->nreal; ri++) {
        or_real* r = &e->R[ri];
        const double beta_s = r->has_beta ? seas * fp * r->local_beta : beta_s_all;
        e->nev = 0;
        for (int c = 0; c < e->NC; c++) {
            const int base = e->cstart[c], S = r->cS[c], L = r->cL[c], I = r->cI[c];
            if (I == 0) continue;
            double sis = 0;
            fx128 sacc = {0, 0};
            e->wc.n_inf += I;
            for (int j = 0; j < I; j++) {
                const int i = r->perm[base + S + L + j];
                const int gen = r->itype[i] / CS_INFECT_TYPE_MASK;
                /* household pass (src/Sweep.cpp:331-377) */
                if (beta_h > 0) {
                    const int h = e->hh[i], nh = e->hnh[h];
                    if (nh > 1) {
                        const int f0 = e->hfirst[h];
                        double HI = beta_h * calc_house_inf(e, r, i, ts);
                        int absent = 0;
                        for (int m = f0; m < f0 + nh && !absent; m++)
                            for (int pt = 0; pt < NPT && !absent; pt++) {
                                int lk = e->plinks[(size_t)m * CS_MAX_PLACE_TYPES + pt];
                                if (lk >= 0) absent = place_closed(r, pt, lk, ts) && host_absent(r, m, ts);
                            }
                        if (absent) HI *= P->efficacies[2][CS_HOUSE];
                        e->wc.n_hh_inf++;
                        e->wc.n_hh_member += nh;
                        int any_susc = 0;
                        for (int m = f0; m < f0 + nh; m++) any_susc |= is_susceptible(r, m);
                        if (!r->hh_done[i]) e->wc.n_hh_scan += nh;
                        if (!any_susc) r->hh_done[i] = 1;
                        for (int m = f0; m < f0 + nh; m++) {
                            if (is_susceptible(r, m)) {
                                double foi = HI * calc_house_susc(e, r, m, ts, i);
                                rs s;
                                rs_init(&s, r->seed, cw0(ts, PS_HH), (uint32_t)i, (uint32_t)(m - f0));
                                if (rs_next(&s) < foi)
                                    add_event(e, ev_key(i, KP_HH, (uint32_t)(m - f0)), m, i, 1 + CS_INFECT_TYPE_MASK * (1 + gen));
                            }
                        }
                    }
                }
                /* place passes (src/Sweep.cpp:381-640) */
                if (P->do_places && !host_absent(r, i, ts)) {
                    for (int pt = 0; pt < NPT; pt++) {
                        const int lk = e->plinks[(size_t)i * CS_MAX_PLACE_TYPES + pt];
                        if (lk < 0) continue;
                        e->wc.n_place_link++;
                        double PI = fp * seas * calc_place_inf(e, r, i, pt, ts);
                        const int mstart = e->pl_ms[pt][lk];
                        if (pt != P->hotel_place_type) {
                            const int g = e->pglinks[(size_t)i * CS_MAX_PLACE_TYPES + pt];
                            double PIc = PI, PIs;
                            if (PIc > 1) PIc = 1;
                            PIs = PIc;
                            const int gi = e->pl_gs0[pt][lk] + g;
                            const int gstart = e->pl_gst[pt][gi], gsize = e->pl_gsz[pt][gi];
                            rs s;
                            rs_init(&s, r->seed, cw0(ts, PS_PW_SAMPLE), (uint32_t)i, (uint32_t)pt);
                            int n = (PIs >= 1) ? gsize : smp_binomial(&s, gsize, PIs);
                            if (n > 0) n = smp_sample_wor(&s, n, gsize, e->samp);
                            for (int c2 = 0; c2 < n; c2++) {
                                const int m = e->pl_mem[pt][mstart + gstart + e->samp[c2]];
                                e->wc.n_place_cand++;
                                double PS = calc_place_susc(e, r, m, pt, ts);
                                if ((pt == P->care_home_place_type) && ((!CARE(e, i)) || (!CARE(e, m)))) PS *= P->care_home_worker_group_scaling;
                                if (is_susceptible(r, m) && !host_absent(r, m, ts)) {
                                    PS *= calc_person_susc(e, r, m, ts, i) * PIc / PIs;
                                    rs s2;
                                    rs_init(&s2, r->seed, cw0(ts, PS_PW_ACCEPT), (uint32_t)i, ((uint32_t)pt << 24) | (uint32_t)c2);
                                    if ((PS == 1) || (rs_next(&s2) < PS))
                                        add_event(e, ev_key(i, KP_PW + pt, (uint32_t)c2), m, i, 2 + pt + CS_INFECT_TYPE_MASK * (1 + gen));
                                }
                            }
                        }
                        {
                            PI *= P->place_type_prop_between_group_links[pt] * P->place_type_group_size_param1[pt] / ((double)e->pl_n[pt][lk]);
                            if (PI > 1) PI = 1;
                            double PIs = PI;
                            const int pn = e->pl_n[pt][lk];
                            rs s;
                            rs_init(&s, r->seed, cw0(ts, PS_PB_SAMPLE), (uint32_t)i, (uint32_t)pt);
                            int n = (PIs >= 1) ? pn : smp_binomial(&s, pn, PIs);
                            if (n > 0) n = smp_sample_wor(&s, n, pn, e->samp);
                            for (int c2 = 0; c2 < n; c2++) {
                                const int m = e->pl_mem[pt][mstart + e->samp[c2]];
                                e->wc.n_place_cand++;
                                double PS = calc_place_susc(e, r, m, pt, ts);
                                if (CARE(e, i) && CARE(e, m) &&
                                    e->pglinks[(size_t)i * CS_MAX_PLACE_TYPES + pt] != e->pglinks[(size_t)m * CS_MAX_PLACE_TYPES + pt])
                                    PS *= P->care_home_resident_place_scaling;
                                if ((pt == P->care_home_place_type) && ((!CARE(e, i)) || (!CARE(e, m)))) PS *= P->care_home_worker_group_scaling;
                                if (is_susceptible(r, m) && !host_absent(r, m, ts)) {
                                    PS *= calc_person_susc(e, r, m, ts, i) * PI / PIs;
                                    rs s2;
                                    rs_init(&s2, r->seed, cw0(ts, PS_PB_ACCEPT), (uint32_t)i, ((uint32_t)pt << 24) | (uint32_t)c2);
                                    if ((PS == 1) || (rs_next(&s2) < PS))
                                        add_event(e, ev_key(i, KP_PB + pt, (uint32_t)c2), m, i,
                                                  2 + pt + CS_MAX_PLACE_TYPES + CS_INFECT_TYPE_MASK * (1 + gen));
                                }
                            }
                        }
                    }
                }
                /* spatial infectiousness (src/Sweep.cpp:649-690) */
                if (beta_s > 0) {
                    double SI = calc_spatial_inf(e, r, i, ts);
                    int flag = 0;
                    if (P->do_places)
                        for (int pt = 0; pt < NPT && !flag; pt++) {
                            int lk = e->plinks[(size_t)i * CS_MAX_PLACE_TYPES + pt];
                            if (lk >= 0) flag = place_closed(r, pt, lk, ts);
                        }
                    const int neg = flag && host_absent(r, i, ts);
                    if (neg) SI *= P->efficacies[2][CS_SPATIAL];
                    sacc = fx_add(sacc, fx_from(SI));
                    sis = fx_to_double(sacc);
                    e->cell_inf[j] = neg ? (float)-sis : (float)sis;
                }
            }
            /* cell-to-cell doling (src/Sweep.cpp:694-856) */
            if (sis > 0) {
                rs s;
                rs_init(&s, r->seed, cw0(ts, PS_SP_COUNT), (uint32_t)c, 0);
                int K = smp_poisson(&s, sis * beta_s * ((double)r->tot_prob[c]));
                if (K > 0) {
                    for (int j = 0; j < I - 1; j++) e->cell_inf[j] /= ((float)sis);
                    e->cell_inf[I - 1] = (e->cell_inf[I - 1] < 0) ? -1.0f : 1.0f;
                }
                for (int k = 0; k < K; k++) {
                    rs d;
                    rs_init(&d, r->seed, cw0(ts, PS_SP_DRAW), (uint32_t)c, (uint32_t)k);
                    int jj = (I == 1) ? 0 : select_infector_n(e->cell_inf, I, rs_next(&d), &e->wc.n_sp_probe);
                    int closed = (e->cell_inf[jj] < 0);
                    const int i = r->perm[base + S + L + jj];
                    for (int att = 0;; att++) {
                        if (att >= OR_MAX_ATTEMPTS) { e->error = CS_ERANGE; snprintf(e->err, sizeof(e->err), "spatial draw found no live candidate"); break; }
                        e->wc.n_sp_draw++;
                        int l = select_dest_n(e, r, c, rs_next(&d), &e->wc.n_sp_scan);
                        int slot = (int)(rs_next(&d) * ((double)r->S0[l]));
                        int m = r->perm[e->cstart[l] + slot];
                        double d2 = dist2_pp(e, m, i);
                        double acc = k_num(e, d2) / e->max_trans[(size_t)c * e->NC + l];
                        if ((rs_next(&d) >= acc) || is_dead(r, m)) continue;
                        if ((c != l) || (e->hh[m] != e->hh[i])) {
                            double SS = calc_spatial_susc(e, r, m, ts);
                            if (CARE(e, m) || CARE(e, i)) SS *= P->care_home_resident_spatial_scaling;
                            if (slot < r->cS[l]) {
                                e->wc.n_sp_susc++;
                                SS *= calc_person_susc(e, r, m, ts, i);
                                SS *= P->waifw_spatial[AGE_GROUP(e, m)][AGE_GROUP(e, i)];
                                if (!closed && host_absent(r, m, ts)) {
                                    int f2 = 0;
                                    for (int pt = 0; pt < NPT && !f2; pt++) {
                                        int lk = e->plinks[(size_t)m * CS_MAX_PLACE_TYPES + pt];
                                        if (lk >= 0) f2 = place_closed(r, pt, lk, ts);
                                    }
                                    if (f2) SS *= P->efficacies[2][CS_SPATIAL];
                                }
                                if ((SS == 1) || (rs_next(&d) < SS))
                                    if (is_susceptible(r, m))
                                        add_event(e, ev_key(i, KP_SP, (uint32_t)k), m, i,
                                                  2 + 2 * CS_MAX_PLACE_TYPES + CS_INFECT_TYPE_MASK * (1 + r->itype[i] / CS_INFECT_TYPE_MASK));
                            }
                        }
                        break;
                    }
                }
            }
        }
        if (e->error) return e->error;
        resolve(e, r, t, ts);
    }
    return e->error;
}

int or_infect_sweep(or_engine* e, double t)
{
    if (e->P.interventions_active) return fail(e, CS_EUNSUPPORTED, "interventions run only in the day loop (or_run_steps)");
    return infect_sweep(e, t);
}

/* ChooseFinalDiseaseSeverity (src/Update.cpp:1377-1389) */
static int choose_final_severity(const cs_params* P, int ag, double x)
{
    if (x < P->prop_ili[ag]) return CS_SEV_ILI;
    if (x < P->prop_ili[ag] + P->prop_sari[ag]) return CS_SEV_SARI;
    if (x < P->prop_ili[ag] + P->prop_sari[ag] + P->prop_critical[ag]) return CS_SEV_CRITICAL;
    return CS_SEV_MILD;
}

#define TO_DIE(r, x) ((r)->bits[x] & 1)
#define DETECTED(r, x) (((r)->bits[x] >> 1) & 1)

/* DoIncub (src/Update.cpp:408-523) without the list swap */
static void do_incub(or_engine* e, or_real* r, int p, unsigned ts)
{
    const cs_params* P = &e->P;
    if (r->inf[p] != CS_LATENT) return;
    int ag = AGE_GROUP(e, p);
    if (ag >= CS_NUM_AGE_GROUPS) ag = CS_NUM_AGE_GROUPS - 1;
    rs s;
    rs_init(&s, r->seed, cw0(ts, PS_INCUB), (uint32_t)p, 0);
    const double tspd = P->time_steps_per_day;
    r->infness[p] = (float)P->age_infectiousness[ag];
    if (P->infectiousness_sd > 0) {
        double k = 1 / (P->infectiousness_sd * P->infectiousness_sd);
        r->infness[p] *= (float)smp_gamma(&s, k, k);
    }
    double ps = P->proportion_symptomatic[ag] * (host_treated(e, r, p, ts) ? (1 - P->treat_sympt_drop) : 1)
              * (host_vacced(e, r, p, ts) ? (1 - P->vacc_sympt_drop) : 1);
    if (rs_next(&s) < ps) {
        r->inf[p] = CS_INF_ALMOST_SYMPT;
        r->infness[p] *= (float)(-P->sympt_infectiousness);
    } else {
        r->inf[p] = CS_INF_ASYMPT;
        r->infness[p] *= (float)P->asympt_infectiousness;
    }
    if (!P->do_severity || r->inf[p] == CS_INF_ASYMPT) {
        if (P->do_infectiousness_profile)
            r->t_rod[p] = (uint16_t)(r->t_lat[p] + (uint16_t)(P->infectious_period * tspd));
        else
            r->t_rod[p] = (uint16_t)(r->t_lat[p] + icdf_choose(P->icdf[CS_ICDF_INFECTIOUS], P->infectious_period, rs_next(&s), tspd));
    } else {
        int ct = r->t_lat[p] + ((int)(P->latent_to_sympt_delay / P->model_time_step));
        int fin = choose_final_severity(P, ag, rs_next(&s));
        r->sev_final[p] = (uint8_t)fin;
        /* DecideIfPersonDies (src/Update.cpp:388-406) */
        if (fin == CS_SEV_CRITICAL) { if (rs_next(&s) < P->cfr_critical[ag] * P->cfr_critical_scale) r->bits[p] |= 1; }
        else if (fin == CS_SEV_SARI) { if (rs_next(&s) < P->cfr_sari[ag] * P->cfr_sari_scale) r->bits[p] |= 1; }
        else if (fin == CS_SEV_ILI) { if (rs_next(&s) < P->cfr_ili[ag] * P->cfr_ili_scale) r->bits[p] |= 1; }
        if (CARE(e, p) && (fin == CS_SEV_CRITICAL || fin == CS_SEV_SARI) && (rs_next(&s) > P->care_home_rel_prob_hosp)) {
            r->bits[p] |= 1;
            r->sev_final[p] = CS_SEV_ILI;
            fin = CS_SEV_ILI;
        }
        const double(*mn)[CS_NUM_AGE_GROUPS] = P->mean_by_age;
#define CH(tab, mean) icdf_choose(P->icdf[tab], mn[mean][ag], rs_next(&s), tspd)
        if (fin == CS_SEV_MILD) {
            r->t_rod[p] = (uint16_t)(ct + CH(CS_ICDF_MILD_RECOV, CS_MEAN_MILD_RECOV));
        } else if (fin == CS_SEV_CRITICAL) {
            r->t_sari[p] = (uint16_t)(ct + CH(CS_ICDF_ILI_SARI, CS_MEAN_ILI_SARI));
            r->t_crit[p] = (uint16_t)(r->t_sari[p] + CH(CS_ICDF_SARI_CRIT, CS_MEAN_SARI_CRIT));
            if (TO_DIE(r, p)) {
                if (P->include_stepdown_to_death == 1) {
                    r->t_step[p] = (uint16_t)(r->t_crit[p] + CH(CS_ICDF_CRIT_CRITRECOV, CS_MEAN_CRIT_CRITRECOV));
                    r->t_rod[p] = (uint16_t)(r->t_step[p] + CH(CS_ICDF_STEPDOWN_DEATH, CS_MEAN_STEPDOWN_DEATH));
                } else
                    r->t_rod[p] = (uint16_t)(r->t_crit[p] + CH(CS_ICDF_CRIT_DEATH, CS_MEAN_CRIT_DEATH));
            } else {
                r->t_step[p] = (uint16_t)(r->t_crit[p] + CH(CS_ICDF_CRIT_CRITRECOV, CS_MEAN_CRIT_CRITRECOV));
                r->t_rod[p] = (uint16_t)(r->t_step[p] + CH(CS_ICDF_CRITRECOV_RECOV, CS_MEAN_CRITRECOV_RECOV));
            }
        } else if (fin == CS_SEV_SARI) {
            r->t_sari[p] = (uint16_t)(ct + CH(CS_ICDF_ILI_SARI, CS_MEAN_ILI_SARI));
            if (TO_DIE(r, p)) r->t_rod[p] = (uint16_t)(r->t_sari[p] + CH(CS_ICDF_SARI_DEATH, CS_MEAN_SARI_DEATH));
            else r->t_rod[p] = (uint16_t)(r->t_sari[p] + CH(CS_ICDF_SARI_RECOV, CS_MEAN_SARI_RECOV));
        } else {
            if (TO_DIE(r, p)) r->t_rod[p] = (uint16_t)(ct + CH(CS_ICDF_ILI_DEATH, CS_MEAN_ILI_DEATH));
            else r->t_rod[p] = (uint16_t)(ct + CH(CS_ICDF_ILI_RECOV, CS_MEAN_ILI_RECOV));
        }
#undef CH
    }
    /* P.ControlPropCasesId: the realisation's own, switched at the alert, when interventions run */
    const double ctrl = P->interventions_active ? r->ctrl : P->control_prop_cases_id;
    if (r->inf[p] == CS_INF_ALMOST_SYMPT && ((ctrl == 1) || (rs_next(&s) < ctrl))) {
        r->bits[p] |= 2;
        r->t_det[p] = (uint16_t)(ts + (uint16_t)(P->latent_to_sympt_delay * tspd));
    }
}

typedef struct { int32_t child; } or_care_req;

/* DoCase (src/Update.cpp:794-894) for the no-intervention path.  The
 * child-care absence of another household member is returned as a request
 * and applied after the phase (see or_incub_recovery_sweep). */
static int do_case(or_engine* e, or_real* r, int p, unsigned ts)
{
    const cs_params* P = &e->P;
    int req = 0;
    if (r->inf[p] != CS_INF_ALMOST_SYMPT) return 0;
    int ag = AGE_GROUP(e, p);
    if (ag >= CS_NUM_AGE_GROUPS) ag = CS_NUM_AGE_GROUPS - 1;
    r->inf[p] = CS_CASE;
    const int adel = P->us_case_absenteeism_delay, adur = P->us_case_absenteeism_duration;
    if (host_absent(r, p, ts)) {
        if ((int)r->t_ae[p] < (int)ts + adel + adur) r->t_ae[p] = (uint16_t)(ts + adel + adur);
    } else if (P->do_real_sympt_withdrawal && P->do_places) {
        r->t_as[p] = CS_UNSET_TIME;
        rs s;
        rs_init(&s, r->seed, cw0(ts, PS_CASE), (uint32_t)p, 0);
        for (int j = 0; j < e->NPT; j++) {
            if ((e->plinks[(size_t)p * CS_MAX_PLACE_TYPES + j] >= 0) && (j != P->hotel_place_type) && (!host_absent(r, p, ts)) &&
                (P->sympt_place_type_withdrawal_prop[j] > 0)) {
                if ((!CARE(e, p)) && ((P->sympt_place_type_withdrawal_prop[j] == 1) || (rs_next(&s) < P->sympt_place_type_withdrawal_prop[j]))) {
                    r->t_as[p] = (uint16_t)(ts + adel);
                    r->t_ae[p] = (uint16_t)(ts + adel + adur);
                    if ((!host_quarantined(e, r, p, ts)) && (e->plinks[(size_t)p * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] >= 0) &&
                        ((int)e->age[p] >= P->case_absent_child_age_cutoff))
                        r->cnt.cumAC++;
                    if (P->do_households && ((int)e->age[p] < P->case_absent_child_age_cutoff)) {
                        if (!host_quarantined(e, r, p, ts)) r->cnt.cumACS++;
                        if (r->prob_care[p] < P->case_absent_child_prop_adult_carers) req = 1;
                    }
                }
            }
        }
    }
    if (DETECTED(r, p)) {
        r->cnt.cumDC++;
        if (e->NAU) r->ad_cnt[(size_t)e->madunit[e->mcell[p]] * CS_TSU_FIELDS + CS_TSU_INC_DC]++; /* cumDC_adunit (src/Update.cpp:872) */
        /* DoDetectedCase (src/Update.cpp:525-560): microcell trigger counts */
        uint32_t* tg = r->m_trig + (size_t)e->mcell[p] * 6;
        for (int q = 0; q < 5; q++) tg[q]++;
        if (!P->absenteeism_place_closure && P->place_close_round_household) tg[5]++;
    }
    if (host_treated(e, r, p, ts)) r->cumTC[e->pc[p]]++;
    r->cnt.cumC++;
    r->cnt.cumCa[ag]++;
    if (e->NAU) r->ad_cnt[(size_t)e->madunit[e->mcell[p]] * CS_TSU_FIELDS + CS_TSU_INC_C]++; /* cumC_adunit (src/Update.cpp:892) */
    r->cnt.cumC_keyworker[KEYW(e, p)]++;
    if (P->do_severity) {
        if (r->sev_final[p] == CS_SEV_MILD) {
            if (r->sev_cur[p] == CS_SEV_ASYMPTOMATIC) { r->sev_cur[p] = CS_SEV_MILD; sev_to(r, CS_SK_MILD, ag); }
        } else {
            if (r->sev_cur[p] == CS_SEV_ASYMPTOMATIC) { r->sev_cur[p] = CS_SEV_ILI; sev_to(r, CS_SK_ILI, ag); }
        }
    }
    return req;
}

/* one infectious person's timers (src/Sweep.cpp:929-977); returns 1 if the
 * person leaves the infectious list */
static int infectious_timers(or_engine* e, or_real* r, int p, unsigned ts, int* care_req)
{
    const cs_params* P = &e->P;
    int ag = AGE_GROUP(e, p);
    if (ag >= CS_NUM_AGE_GROUPS) ag = CS_NUM_AGE_GROUPS - 1;
    int leave = 0;
    uint16_t ct = (uint16_t)(r->t_lat[p] + ((int)(P->latent_to_sympt_delay / P->model_time_step)));
    if (P->do_symptoms && ts == ct) {
        *care_req = do_case(e, r, p, ts);
        e->wc.n_transition++;
    }
    if (P->do_severity) {
        if (ts == r->t_sari[p]) { /* DoSARI */
            if (r->sev_cur[p] == CS_SEV_ILI) { r->sev_cur[p] = CS_SEV_SARI; sev_from(r, CS_SK_ILI, ag); sev_to(r, CS_SK_SARI, ag); }
        }
        if (ts == r->t_crit[p]) { /* DoCritical */
            if (r->sev_cur[p] == CS_SEV_SARI) { r->sev_cur[p] = CS_SEV_CRITICAL; sev_from(r, CS_SK_SARI, ag); sev_to(r, CS_SK_CRITICAL, ag); }
        }
        if (ts == r->t_step[p]) { /* DoRecoveringFromCritical */
            if (r->sev_cur[p] == CS_SEV_CRITICAL && !TO_DIE(r, p)) { r->sev_cur[p] = CS_SEV_STEPDOWN; sev_from(r, CS_SK_CRITICAL, ag); sev_to(r, CS_SK_CRITRECOV, ag); }
        }
        if (ts == r->t_rod[p]) {
            if (TO_DIE(r, p)) { /* DoDeath_FromCriticalorSARIorILI */
                switch (r->sev_cur[p]) {
                case CS_SEV_CRITICAL: sev_from(r, CS_SK_CRITICAL, ag); sev_cum_only(r, CS_SK_DEATH_CRITICAL, ag); r->sev_cur[p] = CS_SEV_DEAD; break;
                case CS_SEV_SARI: sev_from(r, CS_SK_SARI, ag); sev_cum_only(r, CS_SK_DEATH_SARI, ag); r->sev_cur[p] = CS_SEV_DEAD; break;
                case CS_SEV_ILI: sev_from(r, CS_SK_ILI, ag); sev_cum_only(r, CS_SK_DEATH_ILI, ag); r->sev_cur[p] = CS_SEV_DEAD; break;
                default: break;
                }
            } else { /* DoRecover_FromSeverity */
                if (r->inf[p] == CS_INF_ASYMPT || r->inf[p] == CS_CASE) {
                    switch (r->sev_cur[p]) {
                    case CS_SEV_MILD: sev_from(r, CS_SK_MILD, ag); r->sev_cur[p] = CS_SEV_RECOVERED; break;
                    case CS_SEV_ILI: sev_from(r, CS_SK_ILI, ag); r->sev_cur[p] = CS_SEV_RECOVERED; break;
                    case CS_SEV_SARI: sev_from(r, CS_SK_SARI, ag); r->sev_cur[p] = CS_SEV_RECOVERED; break;
                    case CS_SEV_STEPDOWN: sev_from(r, CS_SK_CRITRECOV, ag); r->sev_cur[p] = CS_SEV_RECOVERED; break;
                    default: break;
                    }
                }
            }
        }
    }
    if (ts == r->t_rod[p]) {
        int recover;
        if (!TO_DIE(r, p)) recover = 1;
        else {
            recover = 0;
            if (host_treated(e, r, p, ts)) {
                rs s;
                rs_init(&s, r->seed, cw0(ts, PS_RECOVER), (uint32_t)p, 0);
                recover = rs_next(&s) < P->treat_death_drop;
            }
        }
        if (r->inf[p] == CS_INF_ASYMPT || r->inf[p] == CS_CASE) {
            if (recover) { /* DoRecover (src/Update.cpp:907-951) */
                r->inf[p] = (r->inf[p] == CS_CASE) ? CS_RECOVERED_SYMP : CS_RECOVERED_ASYMP;
            } else { /* DoDeath (src/Update.cpp:953-998) */
                r->inf[p] = (r->inf[p] == CS_CASE) ? CS_DEAD_SYMP : CS_DEAD_ASYMP;
                r->cnt.cumDa[ag]++;
                if (e->NAU) r->ad_cnt[(size_t)e->madunit[e->mcell[p]] * CS_TSU_FIELDS + CS_TSU_INC_DEATH]++; /* cumD_adunit (src/Update.cpp:976) */
            }
            leave = 1;
            e->wc.n_transition++;
        }
    }
    return leave;
}


/* ------------------------------------------------------ interventions */

static void totals(const or_engine* e, const or_real* r, int64_t* S, int64_t* L, int64_t* I, int64_t* R, int64_t* D);

/* The intervention subset served: case isolation and household quarantine
 * of detected cases, and place closure by microcell, all on global
 * (detected-case or ICU) triggers, with the alert logic of
 * CalibrationThresholdCheck when no calibration is asked for.  NULL when p
 * lies inside it (or configures no intervention), else the reason. */
const char* or_interventions_unsupported(const cs_params* p)
{
    const double never = 10000.0;
    if (p->false_positive_rate != 0.0) return "false-positive infections are outside this path";
    if (p->do_digital_contact_tracing) return "digital contact tracing is outside this path";
    if (p->do_blanket_move_restr) return "movement restrictions are outside this path";
    if (p->absenteeism_place_closure) return "absenteeism place closure is outside this path";
    if (!p->interventions_active) return NULL;
    if (!p->do_global_triggers || p->do_admin_triggers) return "interventions need global triggers (admin-unit and microcell triggers are outside this path)";
    if (p->do_per_capita_triggers) return "per-capita triggers are outside this path";
    if (p->do_no_calibration || p->do_alert_trigger_after_interv || p->date_trigger_reached_cal_time >= 0)
        return "calibration to a trigger date is outside this path";
    if (p->place_close_by_admin_unit || p->do_intervention_delays_by_adunit) return "admin-unit place closure and delays are outside this path";
    if (p->vary_efficacies_over_time) return "efficacies varying over time are outside this path";
    if (p->do_mass_vacc || p->treat_time_start_base < never || p->vacc_time_start_base < never || p->vacc_time_start_geo < never ||
        p->soc_dist_time_start_base < never || p->move_restr_time_start_base < never || p->key_worker_proph_time_start_base < never ||
        p->airport_close_time_start_base < never)
        return "treatment, vaccination, social distancing, movement restrictions, key-worker prophylaxis and airport closure are outside this path";
    if (p->place_close_time_start_base2 < never) return "a second place closure is outside this path";
    if (p->place_close_indep_thresh > 0) return "the independent place-closure threshold is outside this path";
    if (p->place_close_cell_inc_stop_thresh != 0) return "reopening places on a stop threshold is outside this path";
    if (p->triggers_sampling_interval < 1 || p->window_to_evaluate_trigger_alert < 1) return "trigger windows must be at least one sample";
    return NULL;
}

/* RecordSample's trigger value (src/CovidSim.cpp:3150-3160) and
 * CalibrationThresholdCheck's alert and start times
 * (src/CovidSim.cpp:3485-3689) for the realisation's sample m at time t */
/* a trigger threshold of realisation r: its own, else the run's value v */
#define THR(r, v) ((r)->has_thr ? (r)->thr : (v))

static void itv_triggers(or_engine* e, or_real* r, int m, double t)
{
    const cs_params* P = &e->P;
    int64_t S, L, I, R, D;
    totals(e, r, &S, &L, &I, &R, &D);
    int32_t* h = r->hist + (size_t)(m % OR_TRIG_RING) * 3;
    h[0] = (int32_t)r->cnt.cumDC;
    h[1] = (int32_t)r->cnt.sev_cum[CS_SK_CRITICAL];
    h[2] = (int32_t)D;
    const int tsi = P->triggers_sampling_interval, win = P->window_to_evaluate_trigger_alert;
    if ((m >= tsi && tsi >= OR_TRIG_RING) || (m >= win && win >= OR_TRIG_RING)) {
        e->error = CS_EUNSUPPORTED;
        snprintf(e->err, sizeof(e->err), "a trigger window reaches back more than %d samples", OR_TRIG_RING - 1);
        return;
    }
    const int iv = P->do_icu_triggers ? 1 : 0;
    int32_t trig = h[iv];
    if (m >= tsi) trig -= r->hist[(size_t)((m - tsi) % OR_TRIG_RING) * 3 + iv];
    r->trig = trig;
    int32_t alert_cases = h[0];
    if (m >= win) alert_cases -= r->hist[(size_t)((m - win) % OR_TRIG_RING) * 3];
    int32_t alert = alert_cases;
    if (P->trigger_alert_on_deaths) {
        alert = h[2];
        if (m >= win) alert -= r->hist[(size_t)((m - win) % OR_TRIG_RING) * 3 + 2];
    }
    if (alert >= P->case_or_death_threshold_before_alert) {
        r->alert = 1;
        r->ctrl = P->post_alert_control_prop_cases_id;
        /* DO_OR_DONT_AMEND_START_TIME: a start time is set once */
        if (trig > THR(r, P->case_isolation_cell_inc_thresh) && r->ci_start >= 1e10) r->ci_start = t + P->case_isolation_time_start_base;
        if (trig > THR(r, P->hhquar_cell_inc_thresh) && r->hq_start >= 1e10) r->hq_start = t + P->hquarantine_time_start_base;
        if (trig >= THR(r, P->place_close_cell_inc_thresh) && r->pc_start >= 1e10) r->pc_start = t + P->place_close_time_start_base;
    }
    /* the first closure ends after its duration when nothing re-triggers it */
    if (t == r->pc_start + P->place_close_duration) {
        r->pc_prev = r->pc_start;
        if (P->place_close_inc_trig == 0 && P->place_close_frac_inc_trig == 0 && THR(r, P->place_close_cell_inc_thresh) == 0) r->pc_start = 9e9;
    }
}

static int host_to_be_quarantined(const or_engine* e, const or_real* r, int x, unsigned ts)
{
    return ((int)r->t_qs[x] + (int)e->P.us_hquarantine_house_duration > (int)ts) && (r->qc[x] < 2);
}

/* the household-member loop of a carer search (src/Update.cpp:836-868,
 * :752-776, :1227-1246): is some member alive, adult and free (no link to the
 * place type PlaceTypeNoAirNum, or absent, or quarantined)?  `with_absent`
 * selects whether absence counts, as the three call sites differ */
static int itv_household_has_free_adult(const or_engine* e, const or_real* r, int p, unsigned ts, int with_absent)
{
    const cs_params* P = &e->P;
    const int h = e->hh[p], j1 = e->hfirst[h], j2 = j1 + e->hnh[h];
    for (int j = j1; j < j2; j++)
        if (!is_dead(r, j) && ((int)e->age[j] >= P->case_absent_child_age_cutoff) &&
            ((e->plinks[(size_t)j * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] < 0) ||
             (with_absent && host_absent(r, j, ts)) || host_quarantined(e, r, j, ts)))
            return 1;
    return 0;
}
static int itv_first_adult(const or_engine* e, const or_real* r, int p)
{
    const int h = e->hh[p], j1 = e->hfirst[h], j2 = j1 + e->hnh[h];
    for (int j = j1; j < j2; j++)
        if (((int)e->age[j] >= e->P.case_absent_child_age_cutoff) && !is_dead(r, j)) return j;
    return -1;
}

/* the detected-case onset test: p became a detected case at this step */
static int itv_detected_now(const or_engine* e, const or_real* r, int p, unsigned ts)
{
    const uint16_t ct = (uint16_t)(r->t_lat[p] + ((int)(e->P.latent_to_sympt_delay / e->P.model_time_step)));
    return e->P.do_symptoms && ct == ts && DETECTED(r, p);
}

/* household quarantine of DoDetectedCase (src/Update.cpp:612-639) for
 * detected case p.  One application per household per step: the draws are
 * keyed by the household (its first member), so the household's first
 * detected case of the step (lowest index) applies it, with the result any
 * other order would give. */
static void itv_quarantine(or_engine* e, or_real* r, int p, double t, unsigned ts)
{
    const cs_params* P = &e->P;
    if (!P->do_households) return;
    if (!(t >= r->hq_start && t < r->hq_start + P->hquarantine_policy_duration)) return;
    const int h = e->hh[p], j1 = e->hfirst[h], j2 = j1 + e->hnh[h];
    for (int j = j1; j < p; j++)
        if (itv_detected_now(e, r, j, ts)) return; /* an earlier member applies it */
    if (host_to_be_quarantined(e, r, j1, ts) && !P->do_hq_retrigger) return;
    const uint16_t start = (uint16_t)(ts + (uint16_t)(P->time_steps_per_day * P->hquarantine_delay));
    rs s;
    rs_init(&s, r->seed, cw0(ts, PS_HQ), (uint32_t)j1, 0);
    const int k = (rs_next(&s) < P->hquarantine_prop_house_compliant) ? 1 : 0;
    if (k) r->cnt.cumHQ++;
    for (int j = j1; j < j2; j++) {
        r->t_qs[j] = start;
        r->qc[j] = (uint8_t)((k == 0) ? 0 : ((rs_next(&s) < P->hquarantine_prop_indiv_compliant) ? 1 : 0));
        if (r->qc[j] && !host_absent(r, j, ts)) {
            if ((int)e->age[j] >= P->case_absent_child_age_cutoff) {
                if (e->plinks[(size_t)j * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] >= 0) r->cnt.cumAH++;
            } else
                r->cnt.cumACS++;
        }
    }
}

/* case isolation of DoDetectedCase (src/Update.cpp:641-692) for detected
 * case p; returns 1 when an adult carer is requested (applied with the
 * child-care requests of the step, isolation form) */
static int itv_isolate(or_engine* e, or_real* r, int p, double t, unsigned ts)
{
    const cs_params* P = &e->P;
    if (!(t >= r->ci_start && t < r->ci_start + P->case_isolation_policy_duration)) return 0;
    rs s;
    rs_init(&s, r->seed, cw0(ts, PS_ISO), (uint32_t)p, 0);
    if (!((P->case_isolation_prop == 1) || (rs_next(&s) < P->case_isolation_prop))) return 0;
    r->t_is[p] = (uint16_t)ts;
    if (host_absent(r, p, ts)) {
        const int stop = (int)ts + P->us_case_absenteeism_delay + P->us_case_isolation_duration;
        if ((int)r->t_ae[p] < stop) r->t_ae[p] = (uint16_t)stop;
        return 0;
    }
    if (!P->do_real_sympt_withdrawal) return 0;
    r->t_as[p] = (uint16_t)(ts + P->us_case_isolation_delay);
    r->t_ae[p] = (uint16_t)(ts + P->us_case_isolation_delay + P->us_case_isolation_duration);
    const int child = (int)e->age[p] < P->case_absent_child_age_cutoff;
    if (P->do_places && !host_quarantined(e, r, p, ts) && e->plinks[(size_t)p * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] >= 0 && !child)
        r->cnt.cumAC++;
    if (P->do_households && P->do_places && child) {
        if (!host_quarantined(e, r, p, ts)) r->cnt.cumACS++;
        if (r->prob_care[p] < P->case_absent_child_prop_adult_carers) return 1;
    }
    return 0;
}

/* Place.ProbClose (src/CovidSim.cpp:1120) and Person.ProbAbsent
 * (src/CovidSim.cpp:1017), drawn at the realisation's reset from their own
 * streams and recomputed where read */
static float itv_prob_close(const or_real* r, int pt, int k)
{
    rs s;
    rs_init(&s, r->seed, cw0(0, PS_PLACE), (uint32_t)k, (uint32_t)pt);
    return (float)rs_next(&s);
}
static float itv_prob_absent(const or_real* r, int p)
{
    rs s;
    rs_init(&s, r->seed, cw0(0, PS_INIT), (uint32_t)p, 0);
    return (float)rs_next(&s);
}

/* DoPlaceClose with DoAnyway (src/Update.cpp:1095-1190): the place's
 * closure window; 1 when the place closes (its members' absences follow) */
static int itv_close_place(or_engine* e, or_real* r, int pt, int k, unsigned ts, uint16_t t_start, uint16_t t_stop)
{
    const cs_params* P = &e->P;
    if (!(r->pl_ce[pt][k] < t_stop) || r->pl_once[pt][k]) return 0;
    if (P->do_place_close_once_only) r->pl_once[pt][k] = 1;
    (void)ts;
    if ((double)itv_prob_close(r, pt, k) >= P->efficacies[2][pt]) {
        if (r->pl_cs[pt][k] > t_start) r->pl_cs[pt][k] = t_start;
        r->pl_ce[pt][k] = t_stop;
        return 1;
    }
    r->pl_cs[pt][k] = r->pl_ce[pt][k] = t_stop;
    return 0;
}

/* TreatSweep's place-closure block on global triggers
 * (src/Sweep.cpp:1422-1437, :1651-1740) after the step's IncubRecoverySweep,
 * then the members' absences (src/Update.cpp:1192-1260) and the carers'
 * (UpdateHostClosure, src/Update.cpp:1263-1279).  Every member test of
 * every closing place reads the state before any absence of this sweep is
 * written. */
static void itv_treat_sweep(or_engine* e, or_real* r, double t, unsigned ts)
{
    const cs_params* P = &e->P;
    if (!(t >= r->pc_start)) return;
    const uint16_t t_end = (uint16_t)ceil(P->time_steps_per_day * (t + P->place_close_delay_mean + P->place_close_duration));
    const uint16_t t_start = (uint16_t)(ts + (uint16_t)(P->time_steps_per_day * P->place_close_delay_mean));
    const uint16_t t_stop = (uint16_t)(ts + (uint16_t)(P->time_steps_per_day * (P->place_close_delay_mean + P->place_close_duration)));
    int32_t* closed = NULL;
    int nclosed = 0, cap = 0;
    for (int m = 0; m < e->NM; m++) {
        if (e->mstart[m + 1] == e->mstart[m]) continue; /* unpopulated */
        if (r->mc_state[m] == 1 && ts >= r->mc_end[m]) {
            r->mc_state[m] = P->do_place_close_once_only ? 2 : 0;
            r->mc_end[m] = (uint16_t)ts;
        }
        if (P->do_places && t >= r->pc_start && r->mc_state[m] == 0 && r->trig >= THR(r, P->place_close_cell_inc_thresh)) {
            r->mc_end[m] = t_end;
            r->mc_state[m] = 1;
            for (int q = e->mpl_start[m]; q < e->mpl_start[m + 1]; q++) {
                const int pt = e->mpl[q] >> 27, k = e->mpl[q] & ((1 << 27) - 1);
                if (pt == P->hotel_place_type) continue;
                if (itv_close_place(e, r, pt, k, ts, t_start, t_stop)) {
                    if (nclosed == cap) { cap = cap ? 2 * cap : 64; closed = (int32_t*)realloc(closed, sizeof(int32_t) * (size_t)cap); }
                    closed[nclosed++] = e->mpl[q];
                }
            }
        }
    }
    if (!nclosed) { free(closed); return; }
    if (P->do_real_sympt_withdrawal) {
        int32_t* carers = (int32_t*)malloc(sizeof(int32_t) * (size_t)e->N);
        int ncarer = 0;
        for (int pass = 0; pass < 2; pass++) /* 0: tests and counters, 1: absences */
            for (int a = 0; a < nclosed; a++) {
                const int pt = closed[a] >> 27, k = closed[a] & ((1 << 27) - 1);
                const double att = P->place_close_prop_attending[pt];
                for (int q = e->pl_ms[pt][k]; q < e->pl_ms[pt][k + 1]; q++) {
                    const int w = e->pl_mem[pt][q];
                    if (!((att == 0) || ((double)itv_prob_absent(r, w) >= att))) continue;
                    const int child = (int)e->age[w] < P->case_absent_child_age_cutoff;
                    if (pass == 0) {
                        if (!host_absent(r, w, ts) && !host_quarantined(e, r, w, ts) && child) {
                            r->cnt.cumAPCS++;
                            if (r->prob_care[w] < P->case_absent_child_prop_adult_carers && !itv_household_has_free_adult(e, r, w, ts, 0)) {
                                const int c = itv_first_adult(e, r, w);
                                if (c >= 0 && ncarer < e->N) { carers[ncarer++] = c; r->cnt.cumAPA++; }
                            }
                        }
                        if (!child && e->plinks[(size_t)w * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] >= 0) r->cnt.cumAPC++;
                    } else {
                        if (r->t_as[w] > t_start) r->t_as[w] = t_start;
                        if (r->t_ae[w] < t_stop) r->t_ae[w] = t_stop;
                    }
                }
            }
        for (int a = 0; a < ncarer; a++) {
            const int c = carers[a];
            if (r->t_as[c] > t_start) r->t_as[c] = t_start;
            if (r->t_ae[c] < t_stop) r->t_ae[c] = t_stop;
        }
        free(carers);
    }
    free(closed);
}

int or_get_intervention_state(or_engine* e, int ri, cs_intervention_state* out)
{
    if (ri < 0 || ri >= e->nreal) return fail(e, CS_EINVAL, "bad realisation");
    const or_real* r = &e->R[ri];
    memset(out, 0, sizeof(*out));
    out->alert = r->alert;
    out->trigger = r->trig;
    out->control_prop_cases_id = e->P.interventions_active ? r->ctrl : e->P.control_prop_cases_id;
    out->case_isolation_start = r->ci_start;
    out->hquarantine_start = r->hq_start;
    out->place_close_start = r->pc_start;
    for (int m = 0; m < e->NM; m++) out->mcells_closed += r->mc_state[m] == 1;
    return 0;
}

static int incub_recovery_sweep(or_engine* e, double t)
{
    const cs_params* P = &e->P;
    unsigned ts = (unsigned)(uint16_t)(P->time_steps_per_day * t);
    for (int ri = 0; ri < e->nreal; ri++) {
        or_real* r = &e->R[ri];
        /* holidays (src/Sweep.cpp:889-918) */
        if (P->do_places)
            for (int h = 0; h < P->num_holidays; h++) {
                double ht = P->holiday_start[h] + P->holidays_start_day_simtime;
                if ((t + P->model_time_step >= ht) && (t < ht)) {
                    for (int pt = 0; pt < e->NPT; pt++)
                        for (int k = 0; k < e->np[pt]; k++) {
                            int close = 0;
                            if (P->holiday_effect[pt] < 1) {
                                if (P->holiday_effect[pt] == 0) close = 1;
                                else {
                                    rs s;
                                    rs_init(&s, r->seed, cw0((unsigned)h, PS_HOLIDAY), (uint32_t)k, (uint32_t)pt);
                                    close = rs_next(&s) >= P->holiday_effect[pt];
                                }
                            }
                            if (!close) continue;
                            int hs = (int)(ht * P->time_steps_per_day);
                            int he = (int)((ht + P->holiday_duration[h]) * P->time_steps_per_day);
                            if (r->pl_cs[pt][k] > hs) r->pl_cs[pt][k] = (uint16_t)hs;
                            if (r->pl_ce[pt][k] < he) r->pl_ce[pt][k] = (uint16_t)he;
                            for (int q = e->pl_ms[pt][k]; q < e->pl_ms[pt][k + 1]; q++) {
                                int m = e->pl_mem[pt][q];
                                if (r->t_as[m] > hs) r->t_as[m] = (uint16_t)hs;
                                if (r->t_ae[m] < he) r->t_ae[m] = (uint16_t)he;
                            }
                        }
                }
            }
        /* latent -> infectious (src/Sweep.cpp:925-927) */
        int32_t* lst = (int32_t*)malloc(sizeof(int32_t) * (size_t)e->N);
        int nl = 0;
        for (int c = 0; c < e->NC; c++) {
            int base = e->cstart[c];
            e->wc.n_active += r->cL[c] + r->cI[c];
            for (int q = r->cS[c]; q < r->cS[c] + r->cL[c]; q++) {
                int p = r->perm[base + q];
                if (ts == r->t_lat[p]) { do_incub(e, r, p, ts); lst[nl++] = p; }
            }
        }
        e->wc.n_transition += nl;
        batch_move(e, r, lst, nl, 1);
        /* infectious timers (src/Sweep.cpp:929-977); a child-care request is
         * p (DoCase's form) or ~p (case isolation's form) */
        int32_t* care = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)e->N);
        int32_t* det = (int32_t*)malloc(sizeof(int32_t) * (size_t)e->N);
        int nc = 0, nd = 0;
        nl = 0;
        for (int c = 0; c < e->NC; c++) {
            int base = e->cstart[c], lo = r->cS[c] + r->cL[c];
            for (int q = lo; q < lo + r->cI[c]; q++) {
                int p = r->perm[base + q];
                int req = 0;
                if (infectious_timers(e, r, p, ts, &req)) lst[nl++] = p;
                if (req) care[nc++] = p;
                if (P->interventions_active && itv_detected_now(e, r, p, ts)) det[nd++] = p;
            }
        }
        /* DoDetectedCase's interventions (src/Update.cpp:525-792) for the
         * step's detected cases: household quarantine for all of them, then
         * case isolation, so that isolation sees every quarantine of the step */
        for (int a = 0; a < nd; a++) itv_quarantine(e, r, det[a], t, ts);
        for (int a = 0; a < nd; a++)
            if (itv_isolate(e, r, det[a], t, ts)) care[nc++] = ~det[a];
        free(det);
        /* child-care absence (src/Update.cpp:836-868): evaluated against the
         * state after every timer of the step, written after all requests */
        if (nc) {
            int32_t* kk = (int32_t*)malloc(sizeof(int32_t) * (size_t)nc);
            for (int a = 0; a < nc; a++) {
                int p = care[a] >= 0 ? care[a] : ~care[a], h = e->hh[p], j1 = e->hfirst[h], j2 = j1 + e->hnh[h], f = 0;
                kk[a] = -1;
                for (int j3 = j1; (j3 < j2) && !f; j3++)
                    f = (!is_dead(r, j3) && ((int)e->age[j3] >= P->case_absent_child_age_cutoff) &&
                         ((e->plinks[(size_t)j3 * CS_MAX_PLACE_TYPES + P->place_type_no_air_num - 1] < 0) ||
                          host_absent(r, j3, ts) || host_quarantined(e, r, j3, ts)));
                if (!f) {
                    int k = -1;
                    for (int j3 = j1; (j3 < j2) && !f; j3++)
                        if (((int)e->age[j3] >= P->case_absent_child_age_cutoff) && !is_dead(r, j3)) { k = j3; f = 1; }
                    if (f) { kk[a] = k; r->cnt.cumAA++; }
                }
            }
            uint8_t* was_absent = (uint8_t*)malloc((size_t)nc);
            for (int a = 0; a < nc; a++) was_absent[a] = kk[a] >= 0 ? (uint8_t)host_absent(r, kk[a], ts) : 0;
            for (int a = 0; a < nc; a++) {
                int k = kk[a];
                if (k < 0) continue;
                /* DoCase's form keeps an absence already running; isolation's sets it anew */
                if (!was_absent[a] || care[a] < 0) r->t_as[k] = (uint16_t)(ts + P->us_case_isolation_delay);
                r->t_ae[k] = (uint16_t)(ts + P->us_case_isolation_delay + P->us_case_isolation_duration);
            }
            free(was_absent);
            free(kk);
        }
        batch_move(e, r, lst, nl, 2);
        free(lst);
        free(care);
    }
    return e->error;
}

int or_incub_recovery_sweep(or_engine* e, double t)
{
    if (e->P.interventions_active) return fail(e, CS_EUNSUPPORTED, "interventions run only in the day loop (or_run_steps)");
    return incub_recovery_sweep(e, t);
}

/* ------------------------------------------------- realisation start */

/* InitModel's per-run reset (src/CovidSim.cpp:1006-1098) and the
 * random-location branch of SeedInfection (src/CovidSim.cpp:1338-1377). */
int or_set_realisation_params(or_engine* e, const double* local_beta, const int32_t* n_seed)
{
    for (int ri = 0; ri < e->nreal; ri++) {
        or_real* r = &e->R[ri];
        r->has_beta = local_beta != NULL;
        r->local_beta = local_beta ? local_beta[ri] : 0.0;
        r->n_seed = n_seed ? n_seed[ri] : -1;
        if (local_beta && !(local_beta[ri] >= 0)) return fail(e, CS_EINVAL, "local_beta must be >= 0");
        if (n_seed && n_seed[ri] < 0) return fail(e, CS_EINVAL, "n_seed must be >= 0");
    }
    return 0;
}

/* per-realisation trigger thresholds (the three cell-incidence thresholds
 * at once); NULL clears them */
int or_set_realisation_triggers(or_engine* e, const double* threshold)
{
    for (int ri = 0; ri < e->nreal; ri++) {
        if (threshold && !(threshold[ri] >= 0)) return fail(e, CS_EINVAL, "threshold must be >= 0");
        e->R[ri].has_thr = threshold != NULL;
        e->R[ri].thr = threshold ? threshold[ri] : 0.0;
    }
    return 0;
}

int or_reset_realisations(or_engine* e, int n_seed)
{
    const cs_params* P = &e->P;
    const char* why = or_interventions_unsupported(P);
    if (why) return fail(e, CS_EUNSUPPORTED, why);
    e->ts_fresh = 1;
    for (int ri = 0; ri < e->nreal; ri++) {
        or_real* r = &e->R[ri];
        for (int k = 0; k < e->N; k++) {
            r->t_as[k] = CS_UNSET_TIME; r->t_ae[k] = 0;
            r->t_vs[k] = r->t_ts[k] = r->t_is[k] = CS_UNSET_TIME;
            r->t_te[k] = 0;
            r->bits[k] = 0; r->t_det[k] = 0; r->dct[k] = 0;
            r->inf[k] = CS_SUSCEPTIBLE;
            r->t_lat[k] = r->t_rod[k] = 0; r->t_inf[k] = 0;
            r->infector[k] = -1; r->itype[k] = 0;
            r->qc[k] = 2; r->t_qs[k] = CS_UNSET_TIME;
            r->hh_done[k] = 0;
            rs s;
            rs_init(&s, r->seed, cw0(0, PS_INIT), (uint32_t)k, 0);
            (void)rs_next(&s); /* ProbAbsent (unused on this path) */
            r->prob_care[k] = (float)rs_next(&s);
            r->susc[k] = 1.0f;
            if (P->susceptibility_sd > 0) {
                double kk = 1 / (P->susceptibility_sd * P->susceptibility_sd);
                r->susc[k] *= (float)smp_gamma(&s, kk, kk);
            }
            r->infness[k] = 0.0f;
            r->t_sari[k] = r->t_crit[k] = r->t_step[k] = CS_UNSET_TIME;
            r->sev_cur[k] = r->sev_final[k] = CS_SEV_ASYMPTOMATIC;
        }
        for (int c = 0; c < e->NC; c++) {
            int base = e->cstart[c];
            for (int j = 0; j < e->cn[c]; j++) {
                int k = e->cmem[base + j];
                r->perm[base + j] = k;
                r->listpos[k] = j;
            }
            r->cS[c] = e->cn[c];
            r->cL[c] = r->cI[c] = r->cR[c] = r->cD[c] = r->cumTC[c] = 0;
        }
        for (int t = 0; t < e->NPT; t++)
            for (int k = 0; k < e->np[t]; k++) { r->pl_cs[t][k] = CS_UNSET_TIME; r->pl_ce[t][k] = 0; }
        memset(r->m_socdist, 0, (size_t)e->NM);
        memset(r->m_moverest, 0, (size_t)e->NM);
        memset(r->m_trig, 0, (size_t)e->NM * 6 * 4);
        memset(&r->cnt, 0, sizeof(r->cnt));
        memset(r->ad_cnt, 0, (size_t)(e->NAU ? e->NAU : 1) * CS_TSU_FIELDS * 8);
        /* intervention state (InitModel, src/CovidSim.cpp:1104, :1116-1122, :1216-1232) */
        r->alert = 0;
        r->trig = 0;
        r->ctrl = P->pre_alert_control_prop_cases_id;
        r->ci_start = r->hq_start = r->pc_start = r->pc_prev = 1e10;
        memset(r->hist, 0, sizeof(int32_t) * OR_TRIG_RING * 3);
        memset(r->mc_state, 0, (size_t)e->NM);
        memset(r->mc_end, 0, (size_t)e->NM * 2);
        for (int t = 0; t < e->NPT; t++) memset(r->pl_once[t], 0, (size_t)e->np[t]);
        update_probs_real(e, r);
        r->prev_prop_susc = 1.0;
        r->loc_init[0] = P->location_initial_infection[0];
        r->loc_init[1] = P->location_initial_infection[1];
        /* seeding */
        rs s;
        rs_init(&s, r->seed, cw0(0, PS_SEED), 0, 0);
        int choices = 0;
        const int nsd = (r->n_seed >= 0) ? r->n_seed : n_seed;
        int32_t* seeds = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nsd > 0 ? nsd : 1));
        int ns = 0;
        for (int inf = 0; (inf < nsd) && (choices < 10000); inf++) {
            int p, m, nm;
            do {
                p = (int)(rs_next(&s) * (double)e->N);
                m = e->mcell[p];
                nm = e->mstart[m + 1] - e->mstart[m];
            } while (nm == 0);
            int q = e->mmem[e->mstart[m] + (int)(rs_next(&s) * (double)nm)];
            if (is_susceptible(r, q) && (calc_person_susc(e, r, q, 0, 0) > 0) &&
                (P->care_home_place_type < 0 || e->plinks[(size_t)q * CS_MAX_PLACE_TYPES + P->care_home_place_type] < 0)) {
                const float* loc = e->hloc + 2 * (size_t)e->hh[q];
                r->loc_init[0] = loc[0];
                r->loc_init[1] = loc[1];
                r->infector[q] = -2;
                r->itype[q] = CS_INFECT_TYPE_MASK - 1;
                do_infect(e, r, q, 0.0, 0);
                seeds[ns++] = q;
                choices = 0;
            } else {
                inf--;
                choices++;
            }
        }
        batch_move(e, r, seeds, ns, 0);
        free(seeds);
    }
    return e->error;
}

/* --------------------------------------------------------- state I/O */

int or_load_state(or_engine* e, int ri, const cs_state* s)
{
    if (ri < 0 || ri >= e->nreal) return fail(e, CS_EINVAL, "bad realisation");
    or_real* r = &e->R[ri];
    size_t N = (size_t)e->N, NC = (size_t)e->NC;
    e->ts_fresh = 1;
#define CP(dst, src, n) if (src) memcpy(r->dst, src, (n) * sizeof(*r->dst))
    CP(inf, s->inf, N); CP(sev_cur, s->severity_current, N); CP(sev_final, s->severity_final, N);
    CP(bits, s->bits, N); CP(listpos, s->listpos, N); CP(infector, s->infector, N); CP(itype, s->infect_type, N);
    CP(infness, s->infectiousness, N); CP(susc, s->susc, N); CP(prob_care, s->prob_care, N);
    CP(t_inf, s->infection_time, N); CP(t_lat, s->latent_time, N); CP(t_rod, s->recovery_or_death_time, N);
    CP(t_sari, s->sari_time, N); CP(t_crit, s->critical_time, N); CP(t_step, s->stepdown_time, N);
    CP(t_det, s->detected_time, N); CP(t_as, s->absent_start_time, N); CP(t_ae, s->absent_stop_time, N);
    CP(t_ts, s->treat_start_time, N); CP(t_te, s->treat_stop_time, N); CP(t_vs, s->vacc_start_time, N);
    CP(t_is, s->isolation_start_time, N); CP(t_qs, s->quar_start_time, N); CP(qc, s->quar_comply, N);
    CP(dct, s->dct_traced, N);
    CP(cS, s->cell_S, NC); CP(cL, s->cell_L, NC); CP(cI, s->cell_I, NC); CP(cR, s->cell_R, NC); CP(cD, s->cell_D, NC);
    CP(cumTC, s->cell_cumTC, NC); CP(perm, s->cell_perm, N);
    for (int t = 0; t < e->NPT; t++) {
        if (s->place_close_start[t]) memcpy(r->pl_cs[t], s->place_close_start[t], (size_t)e->np[t] * 2);
        if (s->place_close_end[t]) memcpy(r->pl_ce[t], s->place_close_end[t], (size_t)e->np[t] * 2);
    }
    CP(m_socdist, s->mcell_socdist, (size_t)e->NM); CP(m_moverest, s->mcell_moverest, (size_t)e->NM);
#undef CP
    memset(r->hh_done, 0, N);
    r->loc_init[0] = e->P.location_initial_infection[0];
    r->loc_init[1] = e->P.location_initial_infection[1];
    int64_t S = sum_S(e, r);
    r->prev_prop_susc = (double)S / (double)e->N;
    update_probs_real(e, r);
    return 0;
}

int or_store_state(or_engine* e, int ri, cs_state* s)
{
    if (ri < 0 || ri >= e->nreal) return fail(e, CS_EINVAL, "bad realisation");
    or_real* r = &e->R[ri];
    size_t N = (size_t)e->N, NC = (size_t)e->NC;
#define CP(src, dst, n) if (dst) memcpy(dst, r->src, (n) * sizeof(*r->src))
    CP(inf, s->inf, N); CP(sev_cur, s->severity_current, N); CP(sev_final, s->severity_final, N);
    CP(bits, s->bits, N); CP(listpos, s->listpos, N); CP(infector, s->infector, N); CP(itype, s->infect_type, N);
    CP(infness, s->infectiousness, N); CP(susc, s->susc, N); CP(prob_care, s->prob_care, N);
    CP(t_inf, s->infection_time, N); CP(t_lat, s->latent_time, N); CP(t_rod, s->recovery_or_death_time, N);
    CP(t_sari, s->sari_time, N); CP(t_crit, s->critical_time, N); CP(t_step, s->stepdown_time, N);
    CP(t_det, s->detected_time, N); CP(t_as, s->absent_start_time, N); CP(t_ae, s->absent_stop_time, N);
    CP(t_ts, s->treat_start_time, N); CP(t_te, s->treat_stop_time, N); CP(t_vs, s->vacc_start_time, N);
    CP(t_is, s->isolation_start_time, N); CP(t_qs, s->quar_start_time, N); CP(qc, s->quar_comply, N);
    CP(dct, s->dct_traced, N);
    CP(cS, s->cell_S, NC); CP(cL, s->cell_L, NC); CP(cI, s->cell_I, NC); CP(cR, s->cell_R, NC); CP(cD, s->cell_D, NC);
    CP(cumTC, s->cell_cumTC, NC); CP(perm, s->cell_perm, N);
    for (int t = 0; t < e->NPT; t++) {
        if (s->place_close_start[t]) memcpy(s->place_close_start[t], r->pl_cs[t], (size_t)e->np[t] * 2);
        if (s->place_close_end[t]) memcpy(s->place_close_end[t], r->pl_ce[t], (size_t)e->np[t] * 2);
    }
    CP(m_socdist, s->mcell_socdist, (size_t)e->NM); CP(m_moverest, s->mcell_moverest, (size_t)e->NM);
#undef CP
    return 0;
}

/* ----------------------------------------------------------- run loop */

static void totals(const or_engine* e, const or_real* r, int64_t* S, int64_t* L, int64_t* I, int64_t* R, int64_t* D)
{
    *S = *L = *I = *R = *D = 0;
    for (int c = 0; c < e->NC; c++) { *S += r->cS[c]; *L += r->cL[c]; *I += r->cI[c]; *R += r->cR[c]; *D += r->cD[c]; }
}

/* RunModel's timestep loop (src/CovidSim.cpp:1426-1522) for the two sweeps,
 * with the S-drop UpdateProbs check; sps = timesteps per output step (the
 * sampling interval, NumModelTimeStepsPerOutputTimeStep) */
/* the cumulative counters a sample differences by age group (infections,
 * cases, deaths, the 8 severity kinds) and by unit (cumI, cumC, cumDC,
 * cumD), field-major: [field][age] then [field][unit] */
static void ext_now(const or_engine* e, const or_real* r, int64_t* out)
{
    for (int a = 0; a < CS_NUM_AGE_GROUPS; a++) {
        out[CS_TSA_INC_I * CS_NUM_AGE_GROUPS + a] = r->cnt.cumIa[a];
        out[CS_TSA_INC_C * CS_NUM_AGE_GROUPS + a] = r->cnt.cumCa[a];
        out[CS_TSA_INC_DEATH * CS_NUM_AGE_GROUPS + a] = r->cnt.cumDa[a];
        for (int k = 0; k < CS_NUM_SEV_KINDS; k++) out[(CS_TSA_INC_SEV + k) * CS_NUM_AGE_GROUPS + a] = r->cnt.sev_cum_age[k][a];
    }
    for (int f = 0; f < CS_TSU_FIELDS; f++)
        for (int u = 0; u < e->NAU; u++)
            out[CS_TSA_CUM * CS_NUM_AGE_GROUPS + f * e->NAU + u] = r->ad_cnt[(size_t)u * CS_TSU_FIELDS + f];
}

int or_run_steps(or_engine* e, double t0, int n_steps, int sps)
{
    if (sps <= 0) return fail(e, CS_EINVAL, "steps per sample must be positive");
    double t = t0;
    const int ns = n_steps / sps;
    if (e->ts_fresh || !e->ts_last) {
        /* a new series: RunModel samples before its first timestep (src/CovidSim.cpp:1413) */
        free(e->ts_last);
        e->ts_last = (int64_t*)calloc((size_t)e->nreal * CS_TS_LAST, 8);
        free(e->ext_last);
        e->ext_last = (int64_t*)calloc((size_t)e->nreal * (CS_TSA_CUM * CS_NUM_AGE_GROUPS + CS_TSU_FIELDS * e->NAU), 8);
        e->ts_n = 0;
        for (int ri = 0; ri < e->nreal; ri++) {
            int64_t S, L, I, R, D;
            totals(e, &e->R[ri], &S, &L, &I, &R, &D);
            int64_t* lt = e->ts_last + (size_t)CS_TS_LAST * ri;
            lt[0] = e->R[ri].cnt.cumI; lt[1] = e->R[ri].cnt.cumC; lt[2] = D;
            for (int k = 0; k < CS_NUM_SEV_KINDS; k++) lt[3 + k] = e->R[ri].cnt.sev_cum[k];
            ext_now(e, &e->R[ri], e->ext_last + (size_t)ri * (CS_TSA_CUM * CS_NUM_AGE_GROUPS + CS_TSU_FIELDS * e->NAU));
            e->R[ri].samp_prop_susc = (double)S / (double)e->N;
            /* the first sample's CalibrationThresholdCheck (src/CovidSim.cpp:1414) */
            if (e->P.interventions_active) itv_triggers(e, &e->R[ri], 0, t0);
        }
        e->ts_fresh = 0;
        if (e->error) return e->error;
    }
    {
        const char* why = or_interventions_unsupported(&e->P);
        if (why) return fail(e, CS_EUNSUPPORTED, why);
    }
    if (e->ts_n + ns > e->ts_cap || !e->ts_buf) { /* grow, keeping the samples taken so far */
        int cap = e->ts_n + ns > 2 * e->ts_cap ? e->ts_n + ns : 2 * e->ts_cap;
        if (cap < 1) cap = 1;
        int32_t* nb = (int32_t*)calloc((size_t)e->nreal * (size_t)cap * CS_TS_FIELDS, 4);
        int32_t* nx = (int32_t*)calloc((size_t)e->nreal * (size_t)cap * e->ext_w, 4);
        for (int ri = 0; ri < e->nreal && e->ts_buf; ri++) {
            memcpy(nb + (size_t)ri * cap * CS_TS_FIELDS, e->ts_buf + (size_t)ri * e->ts_cap * CS_TS_FIELDS,
                   (size_t)e->ts_n * CS_TS_FIELDS * 4);
            memcpy(nx + (size_t)ri * cap * e->ext_w, e->ts_ext + (size_t)ri * e->ts_cap * e->ext_w,
                   (size_t)e->ts_n * e->ext_w * 4);
        }
        free(e->ts_buf);
        free(e->ts_ext);
        e->ts_buf = nb;
        e->ts_ext = nx;
        e->ts_cap = cap;
    }
    for (int st = 0; st < n_steps; st++) {
        int rc = infect_sweep(e, t);
        if (rc) return rc;
        rc = incub_recovery_sweep(e, t);
        if (rc) return rc;
        /* TreatSweep and UpdateHostClosure (src/CovidSim.cpp:1498-1503) */
        if (e->P.interventions_active) {
            const unsigned ts = (unsigned)(uint16_t)(e->P.time_steps_per_day * t);
            for (int ri = 0; ri < e->nreal; ri++) itv_treat_sweep(e, &e->R[ri], t, ts);
        }
        t += e->P.model_time_step;
        /* the S-drop test reads State.S, which only RecordSample refreshes
         * (src/CovidSim.cpp:1515, :3163), i.e. S as of the last sample */
        for (int ri = 0; ri < e->nreal; ri++) {
            or_real* r = &e->R[ri];
            double prop = r->samp_prop_susc;
            if ((r->prev_prop_susc - prop) > 0.2) {
                r->prev_prop_susc = prop;
                update_probs_real(e, r);
            }
        }
        if ((st + 1) % sps == 0) {
            for (int ri = 0; ri < e->nreal; ri++) {
                or_real* r = &e->R[ri];
                int64_t S, L, I, R, D;
                totals(e, r, &S, &L, &I, &R, &D);
                r->samp_prop_susc = (double)S / (double)e->N;
                int64_t* lt = e->ts_last + (size_t)CS_TS_LAST * ri;
                int32_t* o = e->ts_buf + ((size_t)ri * e->ts_cap + e->ts_n) * CS_TS_FIELDS;
                int64_t now[CS_TS_LAST];
                now[0] = r->cnt.cumI; now[1] = r->cnt.cumC; now[2] = D;
                for (int k = 0; k < CS_NUM_SEV_KINDS; k++) now[3 + k] = r->cnt.sev_cum[k];
                o[0] = (int32_t)S; o[1] = (int32_t)L; o[2] = (int32_t)I; o[3] = (int32_t)R; o[4] = (int32_t)D;
                o[CS_TS_INC_I] = (int32_t)(now[0] - lt[0]);
                o[CS_TS_INC_C] = (int32_t)(now[1] - lt[1]);
                o[CS_TS_INC_DEATH] = (int32_t)(now[2] - lt[2]);
                /* incMild .. incDeath_Critical and the running severity
                 * totals (src/CovidSim.cpp RecordSample, from StateT) */
                for (int k = 0; k < CS_NUM_SEV_KINDS; k++) o[CS_TS_INC_SEV + k] = (int32_t)(now[3 + k] - lt[3 + k]);
                for (int k = 0; k <= CS_SK_CRITRECOV; k++) o[CS_TS_PREV_SEV + k] = (int32_t)r->cnt.sev_prev[k];
                memcpy(lt, now, sizeof(now));
                /* by age group and by administrative unit (RecordSample's
                 * incIa/incCa/incDa, severity by age and *_adunit fields) */
                {
                    const int nc = CS_TSA_CUM * CS_NUM_AGE_GROUPS + CS_TSU_FIELDS * e->NAU;
                    int64_t* le = e->ext_last + (size_t)ri * nc;
                    int64_t* cur = (int64_t*)malloc((size_t)nc * 8);
                    ext_now(e, r, cur);
                    int32_t* ox = e->ts_ext + ((size_t)ri * e->ts_cap + e->ts_n) * e->ext_w;
                    for (int q = 0; q < CS_TSA_CUM * CS_NUM_AGE_GROUPS; q++) ox[q] = (int32_t)(cur[q] - le[q]);
                    for (int k = 0; k <= CS_SK_CRITRECOV; k++)
                        for (int a = 0; a < CS_NUM_AGE_GROUPS; a++)
                            ox[(CS_TSA_PREV_SEV + k) * CS_NUM_AGE_GROUPS + a] = (int32_t)r->cnt.sev_prev_age[k][a];
                    for (int q = 0; q < CS_TSU_FIELDS * e->NAU; q++)
                        ox[CS_TSA_FIELDS * CS_NUM_AGE_GROUPS + q] = (int32_t)(cur[CS_TSA_CUM * CS_NUM_AGE_GROUPS + q] - le[CS_TSA_CUM * CS_NUM_AGE_GROUPS + q]);
                    memcpy(le, cur, (size_t)nc * 8);
                    free(cur);
                }
            }
            e->ts_n++;
            /* the next output step's RecordSample and CalibrationThresholdCheck
             * (src/CovidSim.cpp:1413-1414): sample number ts_n of the series */
            if (e->P.interventions_active)
                for (int ri = 0; ri < e->nreal; ri++) itv_triggers(e, &e->R[ri], e->ts_n, t);
            if (e->error) return e->error;
        }
    }
    return e->error;
}

int or_get_timeseries(or_engine* e, int32_t* out, int max_samples, int* n_samples)
{
    int n = e->ts_n < max_samples ? e->ts_n : max_samples;
    for (int ri = 0; ri < e->nreal; ri++)
        memcpy(out + (size_t)ri * max_samples * CS_TS_FIELDS, e->ts_buf + (size_t)ri * e->ts_cap * CS_TS_FIELDS,
               (size_t)n * CS_TS_FIELDS * 4);
    *n_samples = n;
    return 0;
}

int or_get_timeseries_age(or_engine* e, int32_t* out, int max_samples, int* n_samples)
{
    const int n = e->ts_n < max_samples ? e->ts_n : max_samples;
    const size_t W = CS_TSA_FIELDS * CS_NUM_AGE_GROUPS;
    for (int ri = 0; ri < e->nreal; ri++)
        for (int k = 0; k < n; k++)
            memcpy(out + ((size_t)ri * max_samples + k) * W, e->ts_ext + ((size_t)ri * e->ts_cap + k) * e->ext_w, W * 4);
    *n_samples = n;
    return 0;
}

int or_get_timeseries_adunit(or_engine* e, int32_t* out, int max_samples, int* n_samples)
{
    if (!e->NAU) return fail(e, CS_ESTATE, "no series by administrative unit");
    const int n = e->ts_n < max_samples ? e->ts_n : max_samples;
    const size_t W = (size_t)CS_TSU_FIELDS * e->NAU;
    for (int ri = 0; ri < e->nreal; ri++)
        for (int k = 0; k < n; k++)
            memcpy(out + ((size_t)ri * max_samples + k) * W,
                   e->ts_ext + ((size_t)ri * e->ts_cap + k) * e->ext_w + CS_TSA_FIELDS * CS_NUM_AGE_GROUPS, W * 4);
    *n_samples = n;
    return 0;
}

int or_get_adunit_counters(or_engine* e, int ri, int64_t* out, int reset)
{
    if (!e->NAU) return fail(e, CS_ESTATE, "no counters by administrative unit");
    if (ri < 0 || ri >= e->nreal) return fail(e, CS_EINVAL, "bad realisation");
    memcpy(out, e->R[ri].ad_cnt, (size_t)e->NAU * CS_TSU_FIELDS * 8);
    if (reset) memset(e->R[ri].ad_cnt, 0, (size_t)e->NAU * CS_TSU_FIELDS * 8);
    return 0;
}

int or_get_counters(or_engine* e, int ri, cs_counters* out, int reset)
{
    if (ri < 0 || ri >= e->nreal) return fail(e, CS_EINVAL, "bad realisation");
    or_real* r = &e->R[ri];
    *out = r->cnt;
    totals(e, r, &out->S, &out->L, &out->I, &out->R, &out->D);
    if (reset) {
        double mr = r->cnt.maxRad2;
        memset(&r->cnt, 0, sizeof(r->cnt));
        r->cnt.maxRad2 = mr;
    }
    return 0;
}

int or_get_work_counts(or_engine* e, cs_work_counts* out, int reset)
{
    *out = e->wc;
    if (reset) memset(&e->wc, 0, sizeof(e->wc));
    return 0;
}

/* --------------------------------------------------- component access */

/* ---- InitTransmissionCoeffs (src/SetupModel.cpp:640-846) ----
 * Sums accumulate in the reference's order with one thread (person order,
 * day order inside a person), so from the same draws the results equal the
 * reference's bit for bit. */

/* survival product over infectious steps [d0, d1) (src/SetupModel.cpp:690-694, :780-792) */
static double itc_survive(const cs_params* P, double ps, double rate, int d0, int d1)
{
    for (int d = d0; d < d1 && d < CS_MAX_INFECTIOUS_STEPS; d++) {
        double today = 1.0 - rate * P->infectiousness[d];
        ps *= (today < 0) ? 0 : today;
    }
    return ps;
}

/* per-person draws of src/SetupModel.cpp:667-690 from the stream (seed, ITC, person) */
static void itc_draw(const or_engine* e, uint64_t seed, int p, float* susc, float* infness, uint16_t* rec)
{
    const cs_params* P = &e->P;
    const int a = AGE_GROUP(e, p);
    rs s;
    rs_init(&s, seed, cw0(0, PS_ITC), (uint32_t)p, 0);
    const double imm = P->do_partial_immunity ? (1.0 - P->initial_immunity[a]) : 1.0;
    if (P->susceptibility_sd == 0) *susc = (float)imm;
    else {
        double k = 1 / (P->susceptibility_sd * P->susceptibility_sd);
        *susc = (float)(imm * smp_gamma(&s, k, k));
    }
    float in;
    if (P->infectiousness_sd == 0) in = (float)P->age_infectiousness[a];
    else {
        double k = 1 / (P->infectiousness_sd * P->infectiousness_sd);
        in = (float)(P->age_infectiousness[a] * smp_gamma(&s, k, k));
    }
    if (rs_next(&s) < P->proportion_symptomatic[a]) in *= (float)(-P->sympt_infectiousness);
    else in *= (float)P->asympt_infectiousness;
    *infness = in;
    double q = rs_next(&s) * CS_CDF_RES;
    int j = (int)floor(q);
    q -= (double)j;
    const double* ic = P->icdf[CS_ICDF_INFECTIOUS];
    *rec = (uint16_t)floor(0.5 - (P->infectious_period * or_log(q * ic[j + 1] + (1.0 - q) * ic[j]) / P->model_time_step));
}

static int itc_sums(or_engine* e, const double* prop_age_group, const float* infness, const uint16_t* rec,
                    const int8_t* infstat, cs_transmission_coeffs* out)
{
    const cs_params* P = &e->P;
    const double dt = P->model_time_step;
    if (P->got_waifw_spatial && !prop_age_group) return fail(e, CS_EINVAL, "prop_age_group is needed with a spatial WAIFW matrix");
    double hh_inf = 0, sp_inf = 0, hh_denom = 0;
    for (int p = 0; p < e->N; p++) {
        const int a = AGE_GROUP(e, p);
        const double inf = (double)infness[p];
        if (P->do_households) {
            double hi = P->no_infectiousness_sd_in_hh ? ((infness[p] < 0) ? P->sympt_infectiousness : P->asympt_infectiousness)
                                                       : fabs(inf);
            if (CARE(e, p)) hi *= P->care_home_resident_household_scaling;
            const int h = e->hh[p];
            hi *= dt * P->household_trans * P->household_denom_lookup[e->hnhr[h] - 1];
            const double ps = itc_survive(P, 1.0, hi, 0, (int)rec[p]);
            for (int m = e->hfirst[h]; m < e->hfirst[h] + e->hnh[h]; m++)
                if ((!infstat || infstat[m] == CS_SUSCEPTIBLE) && m != p)
                    hh_inf += (1 - ps) * P->age_susceptibility[a] * (CARE(e, m) ? P->care_home_resident_household_scaling : 1.0);
            hh_denom += (double)(e->hnhr[h] - 1);
        }
        const double lsd = (P->latent_to_sympt_delay > rec[p] * dt) ? rec[p] * dt : P->latent_to_sympt_delay;
        double si = fabs(inf) * P->relative_spatial_contact[a] * (CARE(e, p) ? P->care_home_resident_spatial_scaling : 1.0) * dt;
        if (P->got_waifw_spatial) {
            double av = 0;
            const double* pa = prop_age_group + (size_t)e->madunit[e->mcell[p]] * CS_NUM_AGE_GROUPS;
            for (int b = 0; b < CS_NUM_AGE_GROUPS; b++) av += pa[b] * P->waifw_spatial[b][a];
            si *= av;
        }
        const int nds = (int)(lsd / dt);
        int d;
        for (d = 0; d < nds && d < CS_MAX_INFECTIOUS_STEPS; d++) sp_inf += si * P->infectiousness[d];
        si *= (infness[p] < 0) ? P->sympt_spatial_contact_rate : 1;
        for (; d < (int)rec[p] && d < CS_MAX_INFECTIOUS_STEPS; d++) sp_inf += si * P->infectiousness[d];
    }
    memset(out, 0, sizeof(*out));
    const double spatial_r0 = sp_inf / (double)e->N;
    out->household_sar = hh_inf / hh_denom;
    out->r0_household = hh_inf / (double)e->N;
    double place_inf = 0;
    if (P->do_places)
        for (int pt = 0; pt < e->NPT; pt++) {
            if (pt == P->hotel_place_type) continue;
            /* the reduction's private sum starts at 0 and joins the total
             * after the loop (src/SetupModel.cpp:759) */
            double type_inf = 0;
            for (int p = 0; p < e->N; p++) {
                const int k = e->plinks[(size_t)p * CS_MAX_PLACE_TYPES + pt];
                if (k < 0) continue;
                const double inf = (double)infness[p];
                const double lsd = (P->latent_to_sympt_delay > rec[p] * dt) ? rec[p] * dt : P->latent_to_sympt_delay;
                const double pinf = fabs(inf) * dt * P->place_type_trans[pt];
                double scaled = pinf / P->place_type_group_size_param1[pt];
                const int nds = (int)(lsd / dt);
                double ps = itc_survive(P, 1.0, scaled, 0, nds);
                const int g = e->pglinks[(size_t)p * CS_MAX_PLACE_TYPES + pt];
                const double ngroup = (double)((int64_t)e->pl_gsz[pt][e->pl_gs0[pt][k] + g] - 1);
                const double sympt_rate = ((infness[p] < 0) && !CARE(e, p))
                    ? (P->sympt_place_type_contact_rate[pt] * (1 - P->sympt_place_type_withdrawal_prop[pt])) : 1;
                scaled *= sympt_rate;
                ps = itc_survive(P, ps, scaled, nds, (int)rec[p]);
                const double ps_group = ps;
                const double n = (double)e->pl_n[pt][k];
                scaled = P->place_type_prop_between_group_links[pt] * pinf / n;
                if (CARE(e, p))
                    scaled *= (1.0 - P->care_home_prop_residents) + P->care_home_prop_residents *
                              (P->care_home_worker_group_scaling * ((n - 1) - ngroup) + ngroup) / (n - 1);
                ps = itc_survive(P, 1.0, scaled, 0, nds);
                scaled *= sympt_rate;
                ps = itc_survive(P, ps, scaled, nds, (int)rec[p]);
                type_inf += (1 - ps_group * ps) * ngroup + (1 - ps) * (((double)((int64_t)e->pl_n[pt][k] - 1)) - ngroup);
            }
            place_inf += type_inf;
            out->r0_places_by_type[pt] = type_inf / (double)e->N;
        }
    double rec_days = 0, rec_steps = 0;
    for (int p = 0; p < e->N; p++) { rec_days += rec[p] * dt; rec_steps += rec[p]; }
    out->r0_places = place_inf / (double)e->N;
    out->mean_infectious_days = rec_days / (double)e->N;
    out->mean_infectious_steps = rec_steps / (double)e->N;
    double beta = P->local_beta;
    if (!P->fix_local_beta) {
        beta = P->do_si ? (P->r0 / spatial_r0 - out->r0_household - out->r0_places)
                        : (P->r0 - out->r0_household - out->r0_places) / spatial_r0;
        if (beta < 0) beta = 0;
    }
    out->spatial_r0 = spatial_r0;
    out->local_beta = beta;
    out->r0_spatial = spatial_r0 * beta;
    out->r0 = out->r0_household + out->r0_places + out->r0_spatial;
    return 0;
}

int or_transmission_coeffs_from_draws(or_engine* e, const double* prop_age_group, const float* infness,
                                      const uint16_t* rec, const int8_t* infstat, cs_transmission_coeffs* out)
{
    return itc_sums(e, prop_age_group, infness, rec, infstat, out);
}

int or_init_transmission_coeffs(or_engine* e, const double* prop_age_group, uint64_t seed,
                                cs_transmission_coeffs* out, float* susc_out, float* infness_out, uint16_t* rec_out)
{
    float* su = (float*)malloc(sizeof(float) * (size_t)e->N);
    float* in = (float*)malloc(sizeof(float) * (size_t)e->N);
    uint16_t* rc = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)e->N);
    for (int p = 0; p < e->N; p++) itc_draw(e, seed, p, su + p, in + p, rc + p);
    int ret = itc_sums(e, prop_age_group, in, rc, NULL, out);
    if (!ret && !e->P.fix_local_beta) e->P.local_beta = out->local_beta;
    if (susc_out) memcpy(susc_out, su, sizeof(float) * (size_t)e->N);
    if (infness_out) memcpy(infness_out, in, sizeof(float) * (size_t)e->N);
    if (rec_out) memcpy(rec_out, rc, sizeof(uint16_t) * (size_t)e->N);
    free(su); free(in); free(rc);
    return ret;
}

int or_kernel_num(or_engine* e, const double* r2, double* out, int n)
{
    for (int i = 0; i < n; i++) out[i] = k_num(e, r2[i]);
    return e->error;
}

int or_dist2_persons(or_engine* e, const int32_t* a, const int32_t* b, double* out, int n)
{
    for (int i = 0; i < n; i++) out[i] = dist2_pp(e, a[i], b[i]);
    return 0;
}

int or_get_cell_tables(or_engine* e, int ri, float* mt, float* ct, int32_t* inv, float* tp, int32_t* S0)
{
    size_t NC = (size_t)e->NC;
    if (mt) memcpy(mt, e->max_trans, NC * NC * 4);
    if (ri >= 0 && ri < e->nreal) {
        or_real* r = &e->R[ri];
        if (ct) memcpy(ct, r->cum_trans, NC * NC * 4);
        if (inv) memcpy(inv, r->invcdf, NC * 1025 * 4);
        if (tp) memcpy(tp, r->tot_prob, NC * 4);
        if (S0) memcpy(S0, r->S0, NC * 4);
    }
    return 0;
}

int or_eval_hazards(or_engine* e, int ri, int kind, uint16_t ts, const int32_t* person, const int32_t* other, double* out, int n)
{
    if (ri < 0 || ri >= e->nreal) return fail(e, CS_EINVAL, "bad realisation");
    or_real* r = &e->R[ri];
    for (int i = 0; i < n; i++) {
        int x = person[i], o = other ? other[i] : 0;
        switch (kind) {
        case 0: out[i] = calc_house_inf(e, r, x, ts); break;
        case 1: out[i] = calc_place_inf(e, r, x, o, ts); break;
        case 2: out[i] = calc_spatial_inf(e, r, x, ts); break;
        case 3: out[i] = calc_person_inf(e, r, x, ts); break;
        case 4: out[i] = calc_house_susc(e, r, x, ts, o); break;
        case 5: out[i] = calc_place_susc(e, r, x, o, ts); break;
        case 6: out[i] = calc_spatial_susc(e, r, x, ts); break;
        case 7: out[i] = calc_person_susc(e, r, x, ts, o); break;
        default: return fail(e, CS_EINVAL, "bad hazard kind");
        }
    }
    return 0;
}

int or_icdf_choose(or_engine* e, int table, const double* u, const double* mean, uint16_t* out, int n)
{
    if (table < 0 || table >= CS_NUM_ICDF) return fail(e, CS_EINVAL, "bad table");
    for (int i = 0; i < n; i++) out[i] = icdf_choose(e->P.icdf[table], mean[i], u[i], e->P.time_steps_per_day);
    return 0;
}

int or_select_infector(const float* ci, int n_inf, const double* u, int32_t* out, int n)
{
    for (int i = 0; i < n; i++) out[i] = select_infector(ci, n_inf, u[i]);
    return 0;
}

int or_select_dest_cell(or_engine* e, int ri, int cell, const double* u, int32_t* out, int n)
{
    if (ri < 0 || ri >= e->nreal) return fail(e, CS_EINVAL, "bad realisation");
    for (int i = 0; i < n; i++) out[i] = select_dest(e, &e->R[ri], cell, u[i]);
    return 0;
}

int or_sample(int which, uint64_t key, uint32_t c0, uint32_t c1, uint32_t c2, double a, double b, int n, double* out)
{
    rs s;
    rs_init(&s, key, c0, c1, c2);
    int32_t* tmp = NULL;
    for (int i = 0; i < n; i++) {
        switch (which) {
        case 0: out[i] = rs_next(&s); break;
        case 1: out[i] = (double)smp_binomial(&s, (int32_t)a, b); break;
        case 2: out[i] = (double)smp_poisson(&s, a); break;
        case 3: out[i] = smp_normal(&s); break;
        case 4: out[i] = smp_gamma(&s, a, b); break;
        case 5: { /* one k-subset of [0,n) per call: a = k, b = n; returns indices */
            int k = (int)a, nn = (int)b;
            tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nn > 0 ? nn : 1));
            int m = smp_sample_wor(&s, k, nn, tmp);
            for (int j = 0; j < m && j < n; j++) out[j] = tmp[j];
            free(tmp);
            return m;
        }
        default: return -1;
        }
    }
    return n;
}

/* ---- population set-up: AssignHouseholdAges -------------------------------
 * A restatement of src/SetupModel.cpp:1820-1988 on the library's streams:
 * household k of a population generated with `seed` draws from stream
 * (seed; pass PG_HHAGE of PS_POPGEN, k, 0) after its two location uniforms.
 * Where the reference draws two uniforms in one expression
 * (:1950, InvAgeDist[...] - a[(int)(ranf * nc)]) the left one comes first.
 * Every rejection loop is bounded (the reference's are not); a household
 * whose loop runs out is an error. */
#define OR_PG_HHAGE 3
#define OR_PG_AGES 85
#define OR_PG_MAX_TRIES (1 << 20)

static void or_inv_age(const cs_popgen_params* p, int32_t inv[1000])
{
    /* State.InvAgeDist[0] (src/SetupModel.cpp:1101-1111) */
    double cum[CS_NUM_AGE_GROUPS + 1];
    cum[0] = 0;
    for (int i = 1; i <= CS_NUM_AGE_GROUPS; i++) cum[i] = cum[i - 1] + p->prop_age_group[i - 1];
    for (int i = 0, j = 0; i < 1000; i++) {
        double t = ((double)i) / 1000;
        if (t >= cum[j + 1]) j++;
        t = 5 * (((double)j) + (t - cum[j]) / (cum[j + 1] - cum[j]));
        inv[i] = (int)t;
    }
    inv[999] = CS_NUM_AGE_GROUPS * 5 - 1;
}

static int pg_age(const int32_t* inv, rs* s) { return inv[(int)(1000.0 * rs_next(s))]; }

static int pg_household_ages(const cs_popgen_params* P, const int32_t* inv, rs* s, int n, int* a)
{
    int tries = 0;
#define OR_TRY if (++tries > OR_PG_MAX_TRIES) return -1
    if (!P->do_households) {
        for (int i = 0; i < n; i++) a[i] = pg_age(inv, s);
        return 0;
    }
    const double span_old = (double)((long long)P->old_pers_age - P->no_child_pers_age + 1);
    const double span_young = (double)((long long)P->young_and_single - P->min_adult_age);
    if (n == 1) {
        if (rs_next(s) < P->one_pers_house_prob_old) {
            do { OR_TRY; a[0] = pg_age(inv, s); }
            while ((a[0] < P->no_child_pers_age) || (rs_next(s) > (((double)a[0]) - P->no_child_pers_age + 1) / span_old));
        } else if ((P->one_pers_house_prob_young > 0) && (rs_next(s) < P->one_pers_house_prob_young / (1 - P->one_pers_house_prob_old))) {
            do { OR_TRY; a[0] = pg_age(inv, s); }
            while ((a[0] > P->young_and_single) || (a[0] < P->min_adult_age) ||
                   (rs_next(s) > 1 - P->young_and_single_slope * (((double)a[0]) - P->min_adult_age) / span_young));
        } else {
            do { OR_TRY; } while ((a[0] = pg_age(inv, s)) < P->min_adult_age);
        }
    } else if (n == 2) {
        if (rs_next(s) < P->two_pers_house_prob_old) {
            do { OR_TRY; a[0] = pg_age(inv, s); }
            while ((a[0] < P->no_child_pers_age) || (rs_next(s) > (((double)a[0]) - P->no_child_pers_age + 1) / span_old));
            do { OR_TRY; a[1] = pg_age(inv, s); }
            while ((a[1] > a[0] + P->max_mf_partner_age_gap) || (a[1] < a[0] - P->max_fm_partner_age_gap) ||
                   (a[1] < P->no_child_pers_age) || (rs_next(s) > (((double)a[1]) - P->no_child_pers_age + 1) / span_old));
        } else if (rs_next(s) < P->one_child_two_pers_prob / (1 - P->two_pers_house_prob_old)) {
            do { OR_TRY; } while ((a[0] = pg_age(inv, s)) > P->max_child_age);
            do { OR_TRY; a[1] = pg_age(inv, s); }
            while ((a[1] > a[0] + P->max_parent_age_gap) || (a[1] < a[0] + P->min_parent_age_gap) || (a[1] < P->min_adult_age));
        } else if ((P->two_pers_house_prob_young > 0) &&
                   (rs_next(s) < P->two_pers_house_prob_young / (1 - P->two_pers_house_prob_old - P->one_child_two_pers_prob))) {
            do { OR_TRY; a[0] = pg_age(inv, s); }
            while ((a[0] < P->min_adult_age) || (a[0] > P->young_and_single) ||
                   (rs_next(s) > 1 - P->young_and_single_slope * (((double)a[0]) - P->min_adult_age) / span_young));
            do { OR_TRY; a[1] = pg_age(inv, s); }
            while ((a[1] > a[0] + P->max_mf_partner_age_gap) || (a[1] < a[0] - P->max_fm_partner_age_gap) || (a[1] < P->min_adult_age));
        } else {
            do { OR_TRY; a[0] = pg_age(inv, s); } while (a[0] < P->min_adult_age);
            do { OR_TRY; a[1] = pg_age(inv, s); }
            while ((a[1] > a[0] + P->max_mf_partner_age_gap) || (a[1] < a[0] - P->max_fm_partner_age_gap) || (a[1] < P->min_adult_age));
        }
    } else {
        int nc;
        if (n == 3) {
            if ((P->zero_child_three_pers_prob > 0) || (P->two_child_three_pers_prob > 0))
                nc = (rs_next(s) < P->zero_child_three_pers_prob) ? 0 : ((rs_next(s) < P->two_child_three_pers_prob) ? 2 : 1);
            else
                nc = 1;
        } else if (n == 4) {
            nc = (rs_next(s) < P->one_child_four_pers_prob) ? 1 : 2;
        } else if (n == 5) {
            nc = (rs_next(s) < P->three_child_five_pers_prob) ? 3 : 2;
        } else {
            nc = n - 2 - (int)(3 * rs_next(s));
        }
        if (nc <= 0) {
            do { OR_TRY; a[0] = pg_age(inv, s); a[1] = pg_age(inv, s); }
            while ((a[1] < P->min_adult_age) || (a[0] < P->min_adult_age));
            do { OR_TRY; a[2] = pg_age(inv, s); }
            while ((a[2] >= a[1] + P->max_mf_partner_age_gap) || (a[2] < a[1] - P->max_fm_partner_age_gap));
        } else {
            int k, j;
            do {
                OR_TRY;
                a[0] = 0;
                for (int i = 1; i < nc; i++)
                    a[i] = a[i - 1] + 1 + (int)smp_poisson(s, (double)((long long)P->mean_child_age_gap - 1));
                int base = pg_age(inv, s);
                a[0] = base - a[(int)(rs_next(s) * ((double)nc))];
                for (int i = 1; i < nc; i++) a[i] += a[0];
                k = (((nc == 1) && (rs_next(s) < P->one_child_prob_youngest_under_five)) ||
                     ((nc == 2) && (rs_next(s) < P->two_children_prob_youngest_under_five)) ||
                     ((nc > 2) && (rs_next(s) < P->prob_youngest_child_under_five))) ? 5 : P->max_child_age;
            } while ((a[0] < 0) || (a[0] > k) || (a[nc - 1] > P->max_child_age));
            j = a[nc - 1] - a[0] - (P->max_parent_age_gap - P->min_parent_age_gap);
            if (j > 0) j += P->max_parent_age_gap;
            else j = P->max_parent_age_gap;
            do { OR_TRY; a[nc] = pg_age(inv, s); k = a[nc - 1]; }
            while ((a[nc] > a[0] + j) || (a[nc] < k + P->min_parent_age_gap) || (a[nc] < P->min_adult_age));
            if ((n > nc + 1) && (rs_next(s) > 0.0)) {  /* PROP_OTHER_PARENT_AWAY = 0 (src/SetupModel.cpp:1818) */
                do { OR_TRY; a[nc + 1] = pg_age(inv, s); }
                while ((a[nc + 1] > a[nc] + P->max_mf_partner_age_gap) || (a[nc + 1] < a[nc] - P->max_fm_partner_age_gap) ||
                       (a[nc + 1] > a[0] + j) || (a[nc + 1] < k + P->min_parent_age_gap) || (a[nc + 1] < P->min_adult_age));
            }
            if (n > nc + 2) {
                j = ((a[nc + 1] > a[nc]) ? a[nc + 1] : a[nc]) + P->older_gen_gap;
                if (j >= OR_PG_AGES) j = OR_PG_AGES - 1;
                if (j < P->no_child_pers_age) j = P->no_child_pers_age;
                for (int i = nc + 2; i < n; i++) {
                    do { OR_TRY; } while ((a[i] = pg_age(inv, s)) < j);
                }
            }
        }
    }
#undef OR_TRY
    return 0;
}

int or_household_ages(const cs_popgen_params* p, uint64_t seed, const int32_t* sizes, int n, uint8_t* ages)
{
    int32_t inv[1000];
    or_inv_age(p, inv);
    size_t off = 0;
    for (int h = 0; h < n; h++) {
        if (sizes[h] < 1 || sizes[h] > CS_MAX_HOUSEHOLD_SIZE) return CS_EINVAL;
        rs s;
        rs_init(&s, seed, (uint32_t)OR_PG_HHAGE | ((uint32_t)PS_POPGEN << 16), (uint32_t)h, 0);
        (void)rs_next(&s);
        (void)rs_next(&s);
        int a[CS_MAX_HOUSEHOLD_SIZE + 2];
        if (pg_household_ages(p, inv, &s, sizes[h], a)) return CS_ERANGE;
        for (int i = 0; i < sizes[h]; i++) ages[off + i] = (uint8_t)a[i];
        off += (size_t)sizes[h];
    }
    return 0;
}

/* ---- population set-up: the nearest-places choice -------------------------
 * A restatement of AssignPeopleToPlaces' choice among the nearest places
 * (src/SetupModel.cpp:2250-2350) for given persons: a square spiral over
 * microcells from the person's own (right, up (y - 1), left, down, the side
 * growing every second turn), while fewer than nn places are held or the
 * side is below 4, and the side below the grid's width; places of other
 * countries are skipped; the nn largest kernel weights are kept (a new
 * weight replaces the first smallest held one when larger); then one place
 * with probability proportional to its weight, from the first uniform of
 * stream (seed; pass PG_NEAR of PS_POPGEN for type tp, person, 0).  The
 * place type's kernel tables are built as KernelLookup::init builds them.
 * out[i] = -1 where no place lies within reach. */
#define OR_PG_NEAR 12
int or_nearest_places(const cs_params* kp, const cs_popgen_params* pg, int tp, uint64_t seed, const int32_t* persons,
                      const float* hloc, const int32_t* pmcell, int n, const int64_t* pl_off, const float* pl_loc,
                      const int32_t* mc_country, int32_t* out)
{
    const int nn = pg->place_nearest[tp];
    if (nn <= 0 || nn > 64) return CS_EINVAL;
    or_engine* e = (or_engine*)calloc(1, sizeof(or_engine));
    e->P = *kp;
    e->P.kernel_type = pg->place_kernel_type[tp];
    e->P.kernel_scale = pg->place_kernel_scale[tp];
    e->P.kernel_shape = pg->place_kernel_shape[tp];
    e->P.kernel_p3 = pg->place_kernel_p3[tp];
    e->P.kernel_p4 = pg->place_kernel_p4[tp];
    const cs_params* P = &e->P;
    const size_t ksz = (size_t)P->kernel_size + 1;
    e->kl = (double*)malloc(ksz * 8);
    e->kh = (double*)malloc(ksz * 8);
    for (int i = 0; i <= P->kernel_size; i++) {
        e->kl[i] = kernel_fn(P, i * P->kernel_delta) / 1.0;
        e->kh[i] = kernel_fn(P, i * P->kernel_delta / P->kernel_expansion) / 1.0;
    }
    const int mcw = pg->mcw, mch = pg->mch;
    for (int q = 0; q < n; q++) {
        const int ic = pmcell[q];
        int px = ic / mch, py = ic % mch;
        const int country = mc_country[ic];
        const double hx = hloc[2 * q], hy = hloc[2 * q + 1];
        int id[64];
        double w[64];
        int dir = 0, side = 1, k = 0, turn = 0, step = 0;
        while (((k < nn) || (side < 4)) && (side < mcw)) {
            if (px >= 0 && py >= 0 && px < mcw && py < mch) {
                const int c = px * mch + py;
                if (mc_country[c] == country)
                    for (int64_t pl = pl_off[c]; pl < pl_off[c + 1]; pl++) {
                        const double lx = pl_loc[2 * pl], ly = pl_loc[2 * pl + 1];
                        double d2;
                        if (P->do_utm_coords) d2 = dist2_utm(P, hx, hy, lx, ly);
                        else d2 = periodic_xy(P, fabs(hx - lx), fabs(hy - ly));
                        const double s = k_num(e, d2);
                        if (s > 0) {
                            if (k < nn) {
                                id[k] = (int)pl;
                                w[k] = s;
                                k++;
                            } else {
                                double t = 1e10;
                                int j2 = 0;
                                for (int i2 = 0; i2 < nn; i2++)
                                    if (w[i2] < t) { t = w[i2]; j2 = i2; }
                                if (s > t) { w[j2] = s; id[j2] = (int)pl; }
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
        out[q] = -1;
        if (k == 0) continue;
        for (int i2 = 1; i2 < k; i2++) w[i2] += w[i2 - 1];
        rs s;
        rs_init(&s, seed, (uint32_t)(OR_PG_NEAR | (tp << 8)) | ((uint32_t)PS_POPGEN << 16), (uint32_t)persons[q], 0);
        const double u = rs_next(&s);
        out[q] = id[k - 1];
        for (int i2 = 0; i2 < k; i2++)
            if (u < w[i2] / w[k - 1]) { out[q] = id[i2]; break; }
    }
    const int err = e->error;
    free(e->kl);
    free(e->kh);
    free(e);
    return err;
}
