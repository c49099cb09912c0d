// ORACLE — test infrastructure only. Binds the reference's own C++ CPU core (compiled from
// its sources under the reference tree by oracle/Makefile, never copied) to the C API of
// qgate_oracle.h, so the tests can run the reference arithmetic itself and compare it with
// the C restatement and with the HIP runtime. The calls mirror the reference's CPython glue
// (qgate/simulator/src/glue.cpp:267-602) without Python.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <exception>
#include <string>
#include <vector>

#include "CPUQubitProcessor.h"
#include "CPUQubitStates.h"
#include "CPUQubitsStatesGetter.h"
#include "CPUSamplingPool.h"
#include "GateMatrix.h"
#include "Parallel.h"

extern "C" {
#include "qgate_oracle.h"
}

using namespace qgate_cpu;

namespace {

thread_local std::string g_error;

template <class F>
int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
    } catch (...) {
        g_error = "unknown error";
    }
    return -1;
}

CPUQubitProcessor<double> &proc_d() {
    static CPUQubitProcessor<double> p;
    return p;
}
CPUQubitProcessor<float> &proc_f() {
    static CPUQubitProcessor<float> p;
    return p;
}

qgate::QubitProcessor &proc(const void *qs) {
    const qgate::QubitStates *s = static_cast<const qgate::QubitStates *>(qs);
    if (s->getPrec() == qgate::precFP64) return proc_d();
    return proc_f();
}

qgate::QubitStates &states(void *qs) { return *static_cast<qgate::QubitStates *>(qs); }

qgate::Matrix2x2C64 matrix(const double *m8) {
    qgate::Matrix2x2C64 m;
    for (int i = 0; i < 4; ++i) m(i / 2, i % 2) = std::complex<double>(m8[2 * i], m8[2 * i + 1]);
    return m;
}

qgate::QubitStatesList states_list(void *const *qs, int n) {
    qgate::QubitStatesList l;
    for (int i = 0; i < n; ++i) l.push_back(static_cast<qgate::QubitStates *>(qs[i]));
    return l;
}

qgate::IdListList lane_tables(const int *tables, void *const *qs, int n) {
    qgate::IdListList out;
    int off = 0;
    for (int i = 0; i < n; ++i) {
        int nl = static_cast<qgate::QubitStates *>(qs[i])->getNLanes();
        out.push_back(qgate::IdList(tables + off, tables + off + nl));
        off += nl;
    }
    return out;
}

struct RefPool {
    qgate::SamplingPool *pool;
    void *cdf;      // the probability buffer the pool scanned in place and owns
    int prec;
    long long n;
    long long last_nonzero;
    std::vector<int> empty;
};

long long insert_bits(long long k, const std::vector<int> &sorted) {
    for (int p : sorted) {
        long long low = (1LL << p) - 1;
        k = ((k & ~low) << 1) | (k & low);
    }
    return k;
}

}  // namespace

extern "C" {

const char *qgo_last_error(void) { return g_error.c_str(); }
const char *qgo_kind(void) { return "reference"; }
int qgo_num_workers(void) { return qgate::Parallel::getDefaultNWorkers(); }

void *qgo_states_new(int prec) {
    if (prec == 0) return new CPUQubitStates<double>();
    return new CPUQubitStates<float>();
}

void qgo_states_delete(void *qs) { delete static_cast<qgate::QubitStates *>(qs); }

int qgo_states_n_lanes(const void *qs) { return static_cast<const qgate::QubitStates *>(qs)->getNLanes(); }

void *qgo_states_data(void *qs) {
    qgate::QubitStates *s = static_cast<qgate::QubitStates *>(qs);
    if (s->getPrec() == qgate::precFP64) return static_cast<CPUQubitStates<double> *>(s)->getPtr();
    return static_cast<CPUQubitStates<float> *>(s)->getPtr();
}

int qgo_initialize_states(void *qs, int n_lanes) {
    return guarded([&] { proc(qs).initializeQubitStates(states(qs), n_lanes); });
}

int qgo_reset_states(void *qs) {
    return guarded([&] { proc(qs).resetQubitStates(states(qs)); });
}

int qgo_calc_probability(const void *qs, int lane, double *prob) {
    return guarded([&] { *prob = proc(qs).calcProbability(*static_cast<const qgate::QubitStates *>(qs), lane); });
}

int qgo_join(void *qs, void *const *list, int n, int n_new_lanes) {
    return guarded([&] { proc(qs).join(states(qs), states_list(list, n), n_new_lanes); });
}

int qgo_decohere(int value, double prob, void *qs, int lane) {
    return guarded([&] { proc(qs).decohere(value, prob, states(qs), lane); });
}

int qgo_decohere_and_separate(int value, double prob, void *qs0, void *qs1, const void *qs, int lane) {
    return guarded([&] {
        proc(qs).decohereAndSeparate(value, prob, states(qs0), states(qs1),
                                     *static_cast<const qgate::QubitStates *>(qs), lane);
    });
}

int qgo_apply_reset(void *qs, int lane) {
    return guarded([&] { proc(qs).applyReset(states(qs), lane); });
}

int qgo_apply_gate(const double *mat8, void *qs, int lane) {
    return guarded([&] { proc(qs).applyGate(matrix(mat8), states(qs), lane); });
}

int qgo_apply_controlled_gate(const double *mat8, void *qs, const int *ctrl, int n_ctrl, int target) {
    return guarded([&] {
        qgate::IdList c(ctrl, ctrl + n_ctrl);
        proc(qs).applyControlledGate(matrix(mat8), states(qs), c, target);
    });
}

int qgo_gate_matrix(const char *name, const double *a, int n_args, int adjoint, double *out) {
    (void)n_args;
    return guarded([&] {
        qgate::Matrix2x2C64 m;
        std::string n(name);
        if (n == "U") qgate::U_mat(m, a[0], a[1], a[2]);
        else if (n == "U2") qgate::U2_mat(m, a[0], a[1]);
        else if (n == "U1") qgate::U1_mat(m, a[0]);
        else if (n == "ID") qgate::ID_mat(m);
        else if (n == "X") qgate::X_mat(m);
        else if (n == "Y") qgate::Y_mat(m);
        else if (n == "Z") qgate::Z_mat(m);
        else if (n == "H") qgate::H_mat(m);
        else if (n == "S") qgate::S_mat(m);
        else if (n == "T") qgate::T_mat(m);
        else if (n == "RX") qgate::RX_mat(m, a[0]);
        else if (n == "RY") qgate::RY_mat(m, a[0]);
        else if (n == "RZ") qgate::RZ_mat(m, a[0]);
        else if (n == "ExpiI") qgate::ExpiI_mat(m, a[0]);
        else if (n == "ExpiZ") qgate::ExpiZ_mat(m, a[0]);
        else if (n == "SH") qgate::SH_mat(m);
        else throw std::runtime_error("Unknown gate type, " + n);
        if (adjoint) qgate::adjoint(&m);
        for (int i = 0; i < 4; ++i) {
            out[2 * i] = m(i / 2, i % 2).real();
            out[2 * i + 1] = m(i / 2, i % 2).imag();
        }
    });
}

int qgo_get_states(int prec, void *out, int mathop, const int *tables, void *const *qs, int n_qs,
                   uint64_t empty, int64_t n, int64_t start, int64_t step) {
    return guarded([&] {
        qgate::IdListList t = lane_tables(tables, qs, n_qs);
        qgate::QubitStatesList l = states_list(qs, n_qs);
        CPUQubitsStatesGetter<double> gd;
        CPUQubitsStatesGetter<float> gf;
        qgate::QubitsStatesGetter &g = prec == 0 ? static_cast<qgate::QubitsStatesGetter &>(gd)
                                                 : static_cast<qgate::QubitsStatesGetter &>(gf);
        g.getStates(out, 0, (qgate::MathOp)mathop, t.data(), (qgate::QstateIdx)empty, l, n, start, step);
    });
}

int qgo_prepare_prob_array(int prec, void *prob, const int *tables, void *const *qs, int n_qs,
                           int n_lanes, int n_hidden) {
    return guarded([&] {
        qgate::IdListList t = lane_tables(tables, qs, n_qs);
        qgate::QubitStatesList l = states_list(qs, n_qs);
        CPUQubitsStatesGetter<double> gd;
        CPUQubitsStatesGetter<float> gf;
        qgate::QubitsStatesGetter &g = prec == 0 ? static_cast<qgate::QubitsStatesGetter &>(gd)
                                                 : static_cast<qgate::QubitsStatesGetter &>(gf);
        g.prepareProbArray(prob, t, l, n_lanes, n_hidden);
    });
}

void *qgo_pool_new(int prec, const int *tables, void *const *qs, int n_qs, int n_lanes, int n_hidden,
                   const int *empty, int n_empty) {
    RefPool *rp = new RefPool();
    rp->prec = prec;
    rp->n = 1LL << n_lanes;
    rp->empty.assign(empty, empty + n_empty);
    std::sort(rp->empty.begin(), rp->empty.end());
    size_t item = prec == 0 ? sizeof(double) : sizeof(float);
    rp->cdf = malloc(item * (size_t)rp->n);
    int rc = qgo_prepare_prob_array(prec, rp->cdf, tables, qs, n_qs, n_lanes, n_hidden);
    if (rc != 0) {
        free(rp->cdf);
        delete rp;
        return nullptr;
    }
    rp->last_nonzero = 0;
    for (long long i = rp->n - 1; 0 <= i; --i) {
        double v = prec == 0 ? static_cast<double *>(rp->cdf)[i] : static_cast<float *>(rp->cdf)[i];
        if (v > 0.) {
            rp->last_nonzero = i;
            break;
        }
    }
    qgate::IdList e(empty, empty + n_empty);
    rc = guarded([&] {
        if (prec == 0)
            rp->pool = new CPUSamplingPool<double>(static_cast<double *>(rp->cdf), n_lanes, e);
        else
            rp->pool = new CPUSamplingPool<float>(static_cast<float *>(rp->cdf), n_lanes, e);
    });
    if (rc != 0) {
        free(rp->cdf);
        delete rp;
        return nullptr;
    }
    return rp;
}

int qgo_pool_sample(void *p, int64_t *obs, int n, const double *u) {
    RefPool *rp = static_cast<RefPool *>(p);
    return guarded([&] {
        // draws at or above the last CDF entry have no bin in the reference (its upper_bound
        // returns N, guarded only by an assert); they map to the last nonzero bin here.
        double last = rp->prec == 0 ? static_cast<double *>(rp->cdf)[rp->n - 1]
                                    : static_cast<float *>(rp->cdf)[rp->n - 1];
        for (int i = 0; i < n; ++i) {
            bool beyond = rp->prec == 0 ? !(u[i] < last) : !((float)u[i] < (float)last);
            if (beyond)
                obs[i] = insert_bits(rp->last_nonzero, rp->empty);
            else
                rp->pool->sample(reinterpret_cast<qgate::QstateIdx *>(obs + i), 1, u + i);
        }
    });
}

void qgo_pool_delete(void *p) {
    RefPool *rp = static_cast<RefPool *>(p);
    if (rp == nullptr) return;
    delete rp->pool;   // frees the CDF buffer it owns
    delete rp;
}

const void *qgo_pool_cdf(void *p, int64_t *n_states) {
    RefPool *rp = static_cast<RefPool *>(p);
    *n_states = rp->n;
    return rp->cdf;
}

}  // extern "C"
