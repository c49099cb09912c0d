/*
 * ORACLE — test infrastructure only (see qgate_oracle.h). A C restatement of the reference
 * CPU runtime: qgate/simulator/src/CPUQubitProcessor.cpp, CPUQubitsStatesGetter.cpp,
 * CPUSamplingPool.cpp, CPUQubitStates.cpp, GateMatrix.cpp, with the fork-join work split of
 * Parallel.h:10-117 / Parallel.cpp:21-67 so that FP64 reductions add in the same order.
 * The 8-bit permutation tables of BitPermTable.h:11-75 are replaced by the closed-form bit
 * insertion they tabulate.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, like the reference's -O2 x86-64 build,
 * which has no FMA).
 */
#define _GNU_SOURCE
#include "qgate_oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef long long idx_t;

static __thread char g_error[512];

static void set_error(const char *fmt, ...) {
    va_list va;
    va_start(va, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, va);
    va_end(va);
}

const char *qgo_last_error(void) { return g_error; }
const char *qgo_kind(void) { return "port"; }

/* ---- fork-join (Parallel.h:12-23 distribute, Parallel.cpp:21-67 worker count) -------- */

static int g_default_workers = -1;

static int default_workers(void) {
    if (g_default_workers != -1) return g_default_workers;
    const char *env = getenv("QGATE_NUM_WORKERS");
    if (env != NULL) {
        char *end = NULL;
        long n = strtol(env, &end, 10);
        if (end != env && 0 < n) {
            g_default_workers = (int)n;
            return g_default_workers;
        }
    }
    cpu_set_t set;
    CPU_ZERO(&set);
    sched_getaffinity(0, sizeof(set), &set);
    g_default_workers = CPU_COUNT(&set);
    if (g_default_workers < 1) g_default_workers = 1;
    return g_default_workers;
}

int qgo_num_workers(void) { return default_workers(); }

/* one worker up to 2^16 loop iterations (Parallel.cpp:60-67, dynamic worker mode) */
static int workers_for(idx_t n_loops) { return ((idx_t)1 << 16) < n_loops ? default_workers() : 1; }

typedef void (*span_fn)(void *ctx, int tid, idx_t begin, idx_t end);

typedef struct {
    span_fn fn;
    void *ctx;
    int tid;
    idx_t begin, end;
} span_job;

static void *run_job(void *p) {
    span_job *j = (span_job *)p;
    j->fn(j->ctx, j->tid, j->begin, j->end);
    return NULL;
}

static void distribute(span_fn fn, void *ctx, idx_t begin, idx_t end, idx_t span_base) {
    const int nw = workers_for(end - begin);
    idx_t span = (end - begin + nw - 1) / nw;
    span = ((span + span_base - 1) / span_base) * span_base;
    span_job *jobs = (span_job *)malloc(sizeof(span_job) * (size_t)nw);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nw);
    for (int i = 0; i < nw; ++i) {
        idx_t b = begin + span * i, e = begin + span * (i + 1);
        jobs[i].fn = fn;
        jobs[i].ctx = ctx;
        jobs[i].tid = i;
        jobs[i].begin = b < end ? b : end;
        jobs[i].end = e < end ? e : end;
    }
    for (int i = 1; i < nw; ++i) pthread_create(&th[i], NULL, run_job, &jobs[i]);
    run_job(&jobs[0]);
    for (int i = 1; i < nw; ++i) pthread_join(th[i], NULL);
    free(th);
    free(jobs);
}

/* ---- index helpers (BitPermTable.cpp:5-40, BitPermTable.h:14-70) ------------------------ */

static idx_t insert_bit(idx_t k, int pos) {
    const idx_t low = ((idx_t)1 << pos) - 1;
    return ((k & ~low) << 1) | (k & low);
}

static idx_t insert_bits(idx_t k, const int *sorted_pos, int n) {
    for (int i = 0; i < n; ++i) k = insert_bit(k, sorted_pos[i]);
    return k;
}

static idx_t ext_to_local(uint64_t ext, const int *local_to_ext, int n_lanes) {
    idx_t local = 0;
    for (int l = 0; l < n_lanes; ++l)
        if ((ext >> local_to_ext[l]) & 1) local |= (idx_t)1 << l;
    return local;
}

static void sort_ints(int *v, int n) {
    for (int i = 1; i < n; ++i)
        for (int j = i; 0 < j && v[j] < v[j - 1]; --j) {
            int t = v[j];
            v[j] = v[j - 1];
            v[j - 1] = t;
        }
}

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define COMPLEX_OF(R) R _Complex
#define CMPLX_OF_(R) CMPLX_##R
#define CMPLX_OF(R) CMPLX_OF_(R)
#define CMPLX_double CMPLX
#define CMPLX_float CMPLXF
#define RE_OF_(R) RE_##R
#define RE_OF(R) RE_OF_(R)
#define RE_double creal
#define RE_float crealf
#define IM_OF_(R) IM_##R
#define IM_OF(R) IM_OF_(R)
#define IM_double cimag
#define IM_float cimagf

#define creal(x) RE_OF(REAL)(x)
#define cimag(x) IM_OF(REAL)(x)

#define REAL double
#define SUFFIX _d
#include "qgate_oracle_impl.inc"
#undef REAL
#undef SUFFIX

#define REAL float
#define SUFFIX _f
#include "qgate_oracle_impl.inc"
#undef REAL
#undef SUFFIX

#undef creal
#undef cimag

/* ---- public API ------------------------------------------------------------------------ */

/* the leading members match states_t_d / states_t_f, which views below rely on */
typedef struct {
    int n_lanes;
    idx_t n_states;
    void *a;
    int prec;
} states_t;

void *qgo_states_new(int prec) {
    states_t *qs = (states_t *)calloc(1, sizeof(states_t));
    qs->prec = prec;
    qs->n_lanes = -1;
    return qs;
}

void qgo_states_delete(void *p) {
    states_t *qs = (states_t *)p;
    if (qs == NULL) return;
    free(qs->a);
    free(qs);
}

int qgo_states_n_lanes(const void *p) { return ((const states_t *)p)->n_lanes; }
void *qgo_states_data(void *p) { return ((states_t *)p)->a; }

#define VIEW_D(p) ((states_t_d *)(p))
#define VIEW_F(p) ((states_t_f *)(p))

static size_t amp_size(int prec) { return prec == 0 ? sizeof(double _Complex) : sizeof(float _Complex); }

int qgo_initialize_states(void *p, int n_lanes) {
    states_t *qs = (states_t *)p;
    free(qs->a);
    qs->n_lanes = n_lanes;
    qs->n_states = (idx_t)1 << n_lanes;
    qs->a = malloc(amp_size(qs->prec) * (size_t)qs->n_states);
    if (qs->a == NULL) {
        set_error("out of host memory");
        return -1;
    }
    return 0;
}

int qgo_reset_states(void *p) {
    if (((states_t *)p)->prec == 0)
        reset_states_d(VIEW_D(p));
    else
        reset_states_f(VIEW_F(p));
    return 0;
}

int qgo_calc_probability(const void *p, int lane, double *prob) {
    const states_t *qs = (const states_t *)p;
    *prob = qs->prec == 0 ? calc_prob_d(VIEW_D(p), lane) : (double)calc_prob_f(VIEW_F(p), lane);
    return 0;
}

int qgo_join(void *p, void *const *list, int n, int n_new_lanes) {
    (void)n_new_lanes;
    states_t *qs = (states_t *)p;
    if (qs->prec == 0) {
        states_t_d *src[64];
        for (int i = 0; i < n; ++i) src[i] = VIEW_D(list[i]);
        join_d(VIEW_D(p), src, n);
    } else {
        states_t_f *src[64];
        for (int i = 0; i < n; ++i) src[i] = VIEW_F(list[i]);
        join_f(VIEW_F(p), src, n);
    }
    return 0;
}

int qgo_decohere(int value, double prob, void *p, int lane) {
    if (((states_t *)p)->prec == 0)
        decohere_d(VIEW_D(p), value, prob, lane);
    else
        decohere_f(VIEW_F(p), value, prob, lane);
    return 0;
}

int qgo_decohere_and_separate(int value, double prob, void *p0, void *p1, const void *p, int lane) {
    if (((const states_t *)p)->prec == 0)
        decohere_and_separate_d(VIEW_D(p0), VIEW_D(p1), VIEW_D(p), value, prob, lane);
    else
        decohere_and_separate_f(VIEW_F(p0), VIEW_F(p1), VIEW_F(p), value, prob, lane);
    return 0;
}

int qgo_apply_reset(void *p, int lane) {
    if (((states_t *)p)->prec == 0)
        apply_reset_d(VIEW_D(p), lane);
    else
        apply_reset_f(VIEW_F(p), lane);
    return 0;
}

int qgo_apply_controlled_gate(const double *mat8, void *p, const int *ctrl, int n_ctrl, int target) {
    if (n_ctrl > 62) {
        set_error("too many controls");
        return -1;
    }
    if (((states_t *)p)->prec == 0)
        apply_gate_d(VIEW_D(p), mat8, ctrl, n_ctrl, target);
    else
        apply_gate_f(VIEW_F(p), mat8, ctrl, n_ctrl, target);
    return 0;
}

int qgo_apply_gate(const double *mat8, void *p, int lane) {
    return qgo_apply_controlled_gate(mat8, p, NULL, 0, lane);
}

/* GateMatrix.cpp:13-158 */
int qgo_gate_matrix(const char *name, const double *a, int n_args, int adjoint, double *out) {
    const double _Complex J = CMPLX(0., 1.);
    const double r = sqrt(0.5);
    double _Complex m[4];
    (void)n_args;
    if (strcmp(name, "U") == 0) {
        double c = cos(a[0] * 0.5), s = sin(a[0] * 0.5), l2 = a[2] * 0.5, p2 = a[1] * 0.5;
        m[0] = cexp(J * (-l2 - p2)) * c;
        m[1] = -cexp(J * (l2 - p2)) * s;
        m[2] = cexp(J * (-l2 + p2)) * s;
        m[3] = cexp(J * (l2 + p2)) * c;
    } else if (strcmp(name, "U2") == 0) {
        double l2 = a[1] * 0.5, p2 = a[0] * 0.5;
        m[0] = cexp(J * (-l2 - p2)) * r;
        m[1] = -cexp(J * (l2 - p2)) * r;
        m[2] = cexp(J * (-l2 + p2)) * r;
        m[3] = cexp(J * (l2 + p2)) * r;
    } else if (strcmp(name, "U1") == 0) {
        m[0] = 1.; m[1] = 0.; m[2] = 0.; m[3] = cexp(J * a[0]);
    } else if (strcmp(name, "ID") == 0) {
        m[0] = 1.; m[1] = 0.; m[2] = 0.; m[3] = 1.;
    } else if (strcmp(name, "X") == 0) {
        m[0] = 0.; m[1] = 1.; m[2] = 1.; m[3] = 0.;
    } else if (strcmp(name, "Y") == 0) {
        m[0] = 0.; m[1] = -J; m[2] = J; m[3] = 0.;
    } else if (strcmp(name, "Z") == 0) {
        m[0] = 1.; m[1] = 0.; m[2] = 0.; m[3] = -1.;
    } else if (strcmp(name, "H") == 0) {
        m[0] = 1. * r; m[1] = 1. * r; m[2] = 1. * r; m[3] = -1. * r;
    } else if (strcmp(name, "S") == 0) {
        m[0] = 1.; m[1] = 0.; m[2] = 0.; m[3] = J;
    } else if (strcmp(name, "T") == 0) {
        m[0] = 1.; m[1] = 0.; m[2] = 0.; m[3] = cexp(M_PI * 0.25 * J);
    } else if (strcmp(name, "RX") == 0) {
        double c = cos(a[0] / 2.), s = sin(a[0] / 2.);
        m[0] = c; m[1] = -J * s; m[2] = -J * s; m[3] = c;
    } else if (strcmp(name, "RY") == 0) {
        double c = cos(a[0] / 2.), s = sin(a[0] / 2.);
        m[0] = c; m[1] = -s; m[2] = s; m[3] = c;
    } else if (strcmp(name, "RZ") == 0) {
        double t2 = a[0] / 2.;
        m[0] = cexp(-J * t2); m[1] = 0.; m[2] = 0.; m[3] = cexp(J * t2);
    } else if (strcmp(name, "ExpiI") == 0) {
        double _Complex d = cexp(a[0] * J);
        m[0] = d; m[1] = 0.; m[2] = 0.; m[3] = d;
    } else if (strcmp(name, "ExpiZ") == 0) {
        m[0] = cexp(a[0] * J); m[1] = 0.; m[2] = 0.; m[3] = cexp(-a[0] * J);
    } else if (strcmp(name, "SH") == 0) {
        m[0] = 1. * r; m[1] = 1. * r; m[2] = J * r; m[3] = -J * r;
    } else {
        set_error("Unknown gate type, %s.", name);
        return -1;
    }
    if (adjoint) {
        double _Complex t = m[1];
        m[0] = conj(m[0]);
        m[1] = conj(m[2]);
        m[2] = conj(t);
        m[3] = conj(m[3]);
    }
    for (int i = 0; i < 4; ++i) {
        out[2 * i] = creal(m[i]);
        out[2 * i + 1] = cimag(m[i]);
    }
    return 0;
}

int qgo_get_states(int prec, void *out, int mathop, const int *tables, void *const *qs, int n_qs,
                   uint64_t empty, int64_t n, int64_t start, int64_t step) {
    if (prec == 0) {
        states_t_d *v[64];
        for (int i = 0; i < n_qs; ++i) v[i] = VIEW_D(qs[i]);
        get_states_d(out, mathop, tables, v, n_qs, empty, n, start, step);
    } else {
        states_t_f *v[64];
        for (int i = 0; i < n_qs; ++i) v[i] = VIEW_F(qs[i]);
        get_states_f(out, mathop, tables, v, n_qs, empty, n, start, step);
    }
    return 0;
}

int qgo_prepare_prob_array(int prec, void *prob, const int *tables, void *const *qs, int n_qs,
                           int n_lanes, int n_hidden) {
    if (prec == 0) {
        states_t_d *v[64];
        for (int i = 0; i < n_qs; ++i) v[i] = VIEW_D(qs[i]);
        prepare_prob_d((double *)prob, tables, v, n_qs, n_lanes, n_hidden);
    } else {
        states_t_f *v[64];
        for (int i = 0; i < n_qs; ++i) v[i] = VIEW_F(qs[i]);
        prepare_prob_f((float *)prob, tables, v, n_qs, n_lanes, n_hidden);
    }
    return 0;
}

typedef struct {
    int prec;
    idx_t n;
    void *cdf;
    int empty[64];
    int n_empty;
    idx_t last_nonzero;
} pool_t;

void *qgo_pool_new(int prec, const int *tables, void *const *qs, int n_qs, int n_lanes, int n_hidden,
                   const int *empty, int n_empty) {
    pool_t *pool = (pool_t *)calloc(1, sizeof(pool_t));
    pool->prec = prec;
    pool->n = (idx_t)1 << n_lanes;
    pool->cdf = malloc((prec == 0 ? sizeof(double) : sizeof(float)) * (size_t)pool->n);
    qgo_prepare_prob_array(prec, pool->cdf, tables, qs, n_qs, n_lanes, n_hidden);
    pool->last_nonzero = 0;
    for (idx_t i = pool->n - 1; 0 <= i; --i) {
        double v = prec == 0 ? ((double *)pool->cdf)[i] : ((float *)pool->cdf)[i];
        if (v > 0.) {
            pool->last_nonzero = i;
            break;
        }
    }
    int rc = prec == 0 ? pool_scan_d((double *)pool->cdf, pool->n) : pool_scan_f((float *)pool->cdf, pool->n);
    if (rc != 0) {
        free(pool->cdf);
        free(pool);
        return NULL;
    }
    pool->n_empty = n_empty;
    for (int i = 0; i < n_empty; ++i) pool->empty[i] = empty[i];
    sort_ints(pool->empty, n_empty);
    return pool;
}

int qgo_pool_sample(void *p, int64_t *obs, int n, const double *u) {
    pool_t *pool = (pool_t *)p;
    for (int i = 0; i < n; ++i) {
        idx_t k = pool->prec == 0 ? upper_bound_d((const double *)pool->cdf, pool->n, (double)u[i])
                                  : upper_bound_f((const float *)pool->cdf, pool->n, (float)u[i]);
        if (k >= pool->n) k = pool->last_nonzero;   /* u at or above the normalised total */
        obs[i] = insert_bits(k, pool->empty, pool->n_empty);
    }
    return 0;
}

void qgo_pool_delete(void *p) {
    pool_t *pool = (pool_t *)p;
    if (pool == NULL) return;
    free(pool->cdf);
    free(pool);
}

const void *qgo_pool_cdf(void *p, int64_t *n_states) {
    pool_t *pool = (pool_t *)p;
    *n_states = pool->n;
    return pool->cdf;
}
