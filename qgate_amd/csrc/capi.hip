// C ABI of the qgate HIP runtime (include/qgate_hip.h) and the objects behind the handles.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qgate_hip.h"
#include "gate_matrix.hpp"
#include "kernels.hpp"
#include "qg_runtime.hpp"

using namespace qg;

struct qghip_qubit_states {
    Prec prec;
    int n_lanes = -1;
    void *d_ptr = nullptr;
    size_t bytes = 0;
    // gates queued by a processor with fusion on, applied in tile passes before anything
    // else reads or writes this state vector
    std::vector<QueuedGate> pending;
    int fuse_mode = 0;              // mode the pending gates were queued under (TILE or DENSE)
    int tile_bits = 0, low_bits = -1;
    // a collapse recorded by decohere and not yet launched (dec_lane < 0: none). It precedes
    // the pending gates; a calc_probability that comes next runs it in the same pass.
    int dec_lane = -1, dec_value = 0;
    double dec_norm = 0.;
    // bounded sampling pools that read this state vector when they draw
    std::vector<qghip_sampling_pool *> readers;

    size_t n_states() const { return size_t(1) << n_lanes; }
};

struct qghip_qubit_processor {
    Prec prec;
    int fusion = 0;             // QGHIP_FUSION_OFF, _TILE or _DENSE
    int tile_bits = 0;          // <= 0: default for the precision (DENSE: the block width, default 6)
    int low_bits = -1;          // < 0: default
};

struct qghip_qubits_states_getter {
    Prec prec;
};

struct qghip_sampling_pool {
    PoolData pool;
    std::vector<int> empty_lanes;
    // bounded pool: the state vector it draws from; stale once that vector is written
    qghip_qubit_states *src = nullptr;
    bool stale = false;
};

namespace {

// Partial sums of qghip_qubit_processor_calc_probability_noalloc: a file-scope device array
// instead of an allocation. 2^22 block partials cover P(lane = 0) of every state of up to 35
// lanes (k_prob_partial sums 4096 units per block); the run sums and the total follow them.
// Each device holds its own copy at its own address for the life of the process. In host code
// `scratch` names a host-side stand-in, never device memory: kernels and copies get the
// address hipGetSymbolAddress returns for the current device (noalloc_scratch()).
constexpr u64 kScratchPartials = u64(1) << 22;
constexpr u64 kScratchDoubles = kScratchPartials + kScratchPartials / 8192 + 1;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxScratchDevices = 64;

// Sampling-pool workspace: QGHIP_POOL_AUTO, _BOUNDED or _FULL (qghip_set_sampling_pool_workspace).
int g_pool_workspace = 0;
double *g_scratch_ptr[kMaxScratchDevices] = {};

struct ValueError : Error {
    using Error::Error;
};

[[noreturn]] void value_error(const char *msg) { throw ValueError(msg); }

template <class F>
int guarded(F &&f) {
    try {
        std::lock_guard<std::recursive_mutex> lk(runtime().mtx);
        f();
        return QGHIP_OK;
    } catch (const ValueError &e) {
        last_error() = e.what();
        return QGHIP_EVALUE;
    } catch (const std::exception &e) {
        last_error() = e.what();
        return QGHIP_ERROR;
    } catch (...) {
        last_error() = "unknown error";
        return QGHIP_ERROR;
    }
}

Prec to_prec(int prec) {
    if (prec == QGHIP_FP64) return precFP64;
    if (prec == QGHIP_FP32) return precFP32;
    value_error("precision must be QGHIP_FP64 or QGHIP_FP32");
}

void check_qs(const qghip_qubit_states *qs, Prec prec) {
    if (qs == nullptr) value_error("null qubit states handle");
    if (qs->d_ptr == nullptr) value_error("qubit states are not allocated");
    if (qs->prec != prec) value_error("qubit states precision does not match");
}

void check_lane(const qghip_qubit_states *qs, int lane) {
    if (lane < 0 || lane >= qs->n_lanes) value_error("lane out of range");
}

void check_pauli_lanes(const qghip_qubit_states *qs, uint64_t x, uint64_t z) {
    if ((x | z) == 0) value_error("the Pauli string is the identity");
    if (qs->n_lanes < 64 && ((x | z) >> qs->n_lanes) != 0) value_error("Pauli lane out of range");
}

// State vectors with queued work (so that synchronize can apply it all).
std::vector<qghip_qubit_states *> &pending_states() {
    static std::vector<qghip_qubit_states *> v;
    return v;
}

void unlist(qghip_qubit_states *qs) {
    auto &v = pending_states();
    v.erase(std::remove(v.begin(), v.end(), qs), v.end());
}

constexpr int kDenseWidth = 6;   // default block width of QGHIP_FUSION_DENSE

bool queued(const qghip_qubit_states *qs) { return !qs->pending.empty() || qs->dec_lane >= 0; }

// qs is about to be written or freed: bounded pools drawing from it can no longer draw.
void touch(qghip_qubit_states *qs) {
    if (qs == nullptr) return;
    for (qghip_sampling_pool *pl : qs->readers) {
        pl->stale = true;
        pl->src = nullptr;
    }
    qs->readers.clear();
}

// Apply the queued work of qs: the recorded collapse, then the queued gates.
void flush(qghip_qubit_states *qs) {
    if (qs == nullptr || !queued(qs)) return;
    std::vector<QueuedGate> gates;
    gates.swap(qs->pending);
    const int dec_lane = qs->dec_lane;
    qs->dec_lane = -1;
    unlist(qs);
    if (qs->d_ptr == nullptr) return;
    if (dec_lane >= 0) {
        if (qs->prec == precFP32)
            launch_decohere<float>(static_cast<Cx<float> *>(qs->d_ptr), qs->n_lanes, dec_lane, qs->dec_value,
                                   (float)qs->dec_norm);
        else
            launch_decohere<double>(static_cast<Cx<double> *>(qs->d_ptr), qs->n_lanes, dec_lane, qs->dec_value,
                                    qs->dec_norm);
    }
    if (gates.empty()) return;
    try {
        if (qs->fuse_mode == QGHIP_FUSION_DENSE) {
            const int width = qs->tile_bits > 0 ? qs->tile_bits : kDenseWidth;
            if (qs->prec == precFP32)
                apply_dense_sequence<float>(static_cast<Cx<float> *>(qs->d_ptr), qs->n_lanes, gates, width);
            else
                apply_dense_sequence<double>(static_cast<Cx<double> *>(qs->d_ptr), qs->n_lanes, gates, width);
            return;
        }
        if (qs->prec == precFP32)
            apply_gate_sequence<float>(static_cast<Cx<float> *>(qs->d_ptr), qs->n_lanes, gates, qs->tile_bits,
                                       qs->low_bits);
        else
            apply_gate_sequence<double>(static_cast<Cx<double> *>(qs->d_ptr), qs->n_lanes, gates, qs->tile_bits,
                                        qs->low_bits);
    } catch (...) {
        // an argument the queue was built under is refused before any launch: keep the gates
        // queued so that the failure is reported again, not lost (a failure after some passes
        // ran cannot be undone; the error then reports it)
        if (qs->pending.empty()) {
            qs->pending.swap(gates);
            pending_states().push_back(qs);
        }
        throw;
    }
}

// Forget the queued work of qs (its contents are about to be overwritten or freed).
void drop(qghip_qubit_states *qs) {
    touch(qs);
    if (qs == nullptr || !queued(qs)) return;
    qs->pending.clear();
    qs->dec_lane = -1;
    unlist(qs);
}

void flush_all() {
    std::vector<qghip_qubit_states *> v(pending_states());
    for (qghip_qubit_states *qs : v) flush(qs);
}

constexpr size_t kMaxQueued = 4096;

void release(qghip_qubit_states *qs) {
    if (qs->d_ptr) dev_free(qs->d_ptr, qs->bytes);
    qs->d_ptr = nullptr;
    qs->bytes = 0;
}

template <class R>
Cx<R> *amps(qghip_qubit_states *qs) {
    return static_cast<Cx<R> *>(qs->d_ptr);
}

template <class R>
const Cx<R> *camps(const qghip_qubit_states *qs) {
    return static_cast<const Cx<R> *>(qs->d_ptr);
}

// Dispatch on precision: f(float()) or f(double()).
template <class F>
void by_prec(Prec p, F &&f) {
    if (p == precFP32)
        f(float());
    else
        f(double());
}

void apply_matrix(const qghip_qubit_processor *p, Prec prec, qghip_qubit_states *qs,
                  const GateMatrix &m, const int *ctrl, int n_ctrl, int target) {
    check_qs(qs, prec);
    check_lane(qs, target);
    if (n_ctrl < 0 || (n_ctrl > 0 && ctrl == nullptr)) value_error("bad control lane list");
    std::vector<int> controls(ctrl, ctrl + n_ctrl);
    for (int c : controls) {
        check_lane(qs, c);
        if (c == target) value_error("control and target lanes overlap");
    }
    std::vector<int> sorted(controls);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        value_error("duplicate control lanes");
    touch(qs);
    if (p != nullptr && (p->fusion == QGHIP_FUSION_TILE || p->fusion == QGHIP_FUSION_DENSE)) {
        // gates queued under other settings are applied under those first
        if (!qs->pending.empty() &&
            (qs->fuse_mode != p->fusion || qs->tile_bits != p->tile_bits || qs->low_bits != p->low_bits))
            flush(qs);
        if (!queued(qs)) pending_states().push_back(qs);
        qs->fuse_mode = p->fusion;
        qs->tile_bits = p->tile_bits;
        qs->low_bits = p->low_bits;
        qs->pending.push_back(QueuedGate{m, controls, target});
        if (qs->pending.size() >= kMaxQueued) flush(qs);
        return;
    }
    flush(qs);
    by_prec(prec, [&](auto r) {
        typedef decltype(r) R;
        launch_gate<R>(amps<R>(qs), qs->n_lanes, m, controls, target);
    });
}

std::vector<StateView> make_views(const int *lane_tables, qghip_qubit_states *const *list, int n,
                                  Prec prec) {
    if (n < 0 || (n > 0 && (list == nullptr || lane_tables == nullptr)))
        value_error("bad qubit states list");
    std::vector<StateView> views;
    int off = 0;
    for (int q = 0; q < n; ++q) {
        check_qs(list[q], prec);
        flush(list[q]);
        StateView v;
        v.ptr = list[q]->d_ptr;
        v.n_lanes = list[q]->n_lanes;
        v.local_to_ext.assign(lane_tables + off, lane_tables + off + v.n_lanes);
        off += v.n_lanes;
        views.push_back(v);
    }
    return views;
}

}  // namespace

extern "C" {

const char *qghip_last_error(void) { return last_error().c_str(); }

namespace {

// Device address of `scratch` on the runtime's device, looked up once per device with that
// device current (runtime_init made it current; the library uses no other).
double *noalloc_scratch() {
    Runtime &rt = runtime();
    stream();   // initialises the runtime on the default device if nothing did yet
    QG_REQUIRE(0 <= rt.device && rt.device < kMaxScratchDevices, "device %d out of range", rt.device);
    double *&ptr = g_scratch_ptr[rt.device];
    if (ptr == nullptr) {
        int cur = -1;
        QG_HIP(hipGetDevice(&cur));
        QG_REQUIRE(cur == rt.device, "device %d is not current", rt.device);
        void *p = nullptr;
        QG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        QG_REQUIRE(p != nullptr, "no device address for the scratch array");
        ptr = static_cast<double *>(p);
    }
    return ptr;
}

}  // namespace

int qghip_devices_initialize(int device_id) {
    return guarded([&] {
        runtime_init(device_id);
        noalloc_scratch();
    });
}

int qghip_devices_clear(void) {
    return guarded([&] {
        flush_all();   // queued gates and recorded collapses first, on the stream about to go
        runtime_clear();
    });
}

int qghip_set_launch_piece(int64_t items) {
    return guarded([&] {
        if (items < (int64_t(1) << 12) || items > (int64_t(1) << 30) || (items & (items - 1)) != 0)
            value_error("launch piece must be a power of two in [2^12, 2^30]");
        set_launch_piece_items((u64)items);
    });
}

int qghip_set_sampling_pool_workspace(int mode) {
    return guarded([&] {
        if (mode != QGHIP_POOL_AUTO && mode != QGHIP_POOL_BOUNDED && mode != QGHIP_POOL_FULL)
            value_error("unknown sampling pool workspace mode");
        g_pool_workspace = mode;
    });
}

int qghip_set_register_tiles(int on) {
    return guarded([&] { set_register_tiles(on != 0); });
}

int qghip_device_info(int *device_id, int *n_cus, uint64_t *total_mem) {
    return guarded([&] {
        stream();
        Runtime &rt = runtime();
        if (device_id) *device_id = rt.device;
        if (n_cus) *n_cus = rt.n_cus;
        if (total_mem) *total_mem = rt.total_mem;
    });
}

int qghip_register_matrix_factory(const char *gate_type, int *factory_id) {
    return guarded([&] {
        if (gate_type == nullptr || factory_id == nullptr) value_error("null argument");
        int id = matrix_factory_id(gate_type);
        if (id < 0) raise_error("Unknown gate type, %s.", gate_type);
        *factory_id = id;
    });
}

int qghip_gate_matrix(int factory_id, const double *args, int n_args, int adjoint, double *mat8) {
    return guarded([&] {
        if (mat8 == nullptr || (n_args > 0 && args == nullptr)) value_error("null argument");
        GateMatrix m = make_gate_matrix(factory_id, args, n_args, adjoint != 0);
        gate_matrix_to_array(m, mat8);
    });
}

int qghip_qubit_states_new(int prec, qghip_qubit_states **out) {
    return guarded([&] {
        if (out == nullptr) value_error("null argument");
        qghip_qubit_states *qs = new qghip_qubit_states;
        qs->prec = to_prec(prec);
        *out = qs;
    });
}

int qghip_qubit_states_delete(qghip_qubit_states *qs) {
    return guarded([&] {
        if (qs == nullptr) return;
        drop(qs);
        release(qs);
        delete qs;
    });
}

int qghip_qubit_states_deallocate(qghip_qubit_states *qs) {
    return guarded([&] {
        if (qs == nullptr) value_error("null qubit states handle");
        drop(qs);
        release(qs);
    });
}

int qghip_qubit_states_get_n_lanes(const qghip_qubit_states *qs, int *n_lanes) {
    return guarded([&] {
        if (qs == nullptr || n_lanes == nullptr) value_error("null argument");
        *n_lanes = qs->n_lanes;
    });
}

int qghip_qubit_processor_new(int prec, qghip_qubit_processor **out) {
    return guarded([&] {
        if (out == nullptr) value_error("null argument");
        qghip_qubit_processor *p = new qghip_qubit_processor;
        p->prec = to_prec(prec);
        *out = p;
    });
}

int qghip_qubit_processor_delete(qghip_qubit_processor *p) {
    return guarded([&] { delete p; });
}

int qghip_qubit_processor_set_fusion(qghip_qubit_processor *p, int mode, int tile_bits, int low_bits) {
    return guarded([&] {
        if (p == nullptr) value_error("null processor handle");
        if (mode != QGHIP_FUSION_OFF && mode != QGHIP_FUSION_TILE && mode != QGHIP_FUSION_DENSE)
            value_error("unknown fusion mode");
        if (tile_bits > 14) value_error("tile_bits out of range");
        if (mode == QGHIP_FUSION_DENSE && (tile_bits == 1 || tile_bits > 6))
            value_error("dense block width must be 2..6");
        if (low_bits > 8) value_error("low_bits out of range");
        // complex64 tiles move amplitudes in pairs, so their lowest bit stays in the tile
        if (mode == QGHIP_FUSION_TILE && p->prec == precFP32 && low_bits == 0)
            value_error("complex64 tile passes need low_bits >= 1");
        p->fusion = mode;
        p->tile_bits = tile_bits;
        p->low_bits = low_bits;
    });
}

int qghip_qubit_processor_synchronize(qghip_qubit_processor *p) {
    return guarded([&] {
        if (p == nullptr) value_error("null processor handle");
        flush_all();
        sync();
    });
}

int qghip_qubit_processor_reset(qghip_qubit_processor *p) {
    return guarded([&] {
        if (p == nullptr) value_error("null processor handle");
    });
}

int qghip_qubit_processor_initialize_qubit_states(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                                  int n_lanes) {
    return guarded([&] {
        if (p == nullptr || qs == nullptr) value_error("null argument");
        if (qs->prec != p->prec) value_error("qubit states precision does not match");
        if (n_lanes < 1 || n_lanes > 40) value_error("n_lanes out of range (1..40)");
        drop(qs);
        release(qs);
        qs->n_lanes = n_lanes;
        qs->bytes = amp_bytes(qs->prec) << n_lanes;
        qs->d_ptr = dev_alloc(qs->bytes);
    });
}

int qghip_qubit_processor_reset_qubit_states(qghip_qubit_processor *p, qghip_qubit_states *qs) {
    return guarded([&] {
        if (p == nullptr) value_error("null processor handle");
        check_qs(qs, p->prec);
        drop(qs);
        by_prec(p->prec, [&](auto r) {
            typedef decltype(r) R;
            launch_set_zero_state<R>(amps<R>(qs), qs->n_lanes);
        });
    });
}

int qghip_qubit_processor_calc_probability(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                           int local_lane, double *prob) {
    return guarded([&] {
        if (p == nullptr || prob == nullptr) value_error("null argument");
        check_qs(qs, p->prec);
        check_lane(qs, local_lane);
        if (qs->dec_lane >= 0 && qs->pending.empty()) {
            // the recorded collapse and this probability in one pass over the state
            const int dec_lane = qs->dec_lane;
            qs->dec_lane = -1;
            unlist(qs);
            by_prec(p->prec, [&](auto r) {
                typedef decltype(r) R;
                *prob = decohere_prob<R>(amps<R>(qs), qs->n_lanes, dec_lane, qs->dec_value, (R)qs->dec_norm,
                                         local_lane);
            });
            return;
        }
        flush(qs);
        by_prec(p->prec, [&](auto r) {
            typedef decltype(r) R;
            *prob = calc_prob<R>(camps<R>(qs), qs->n_lanes, local_lane);
        });
    });
}

int qghip_qubit_processor_calc_probability_noalloc(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                                   int local_lane, double *prob) {
    return guarded([&] {
        if (p == nullptr || prob == nullptr) value_error("null argument");
        check_qs(qs, p->prec);
        check_lane(qs, local_lane);
        // applying queued gates can allocate (dense blocks upload their matrices)
        if (!qs->pending.empty()) value_error("gates are queued on this state vector; synchronize first");
        by_prec(p->prec, [&](auto r) {
            typedef decltype(r) R;
            if (calc_prob_buffer_len<R>(qs->n_lanes, local_lane) > kScratchDoubles)
                value_error("state vector too large for the fixed scratch array (at most 35 lanes)");
        });
        double *buf = noalloc_scratch();
        flush(qs);   // a recorded collapse only: k_decohere, no allocation
        by_prec(p->prec, [&](auto r) {
            typedef decltype(r) R;
            *prob = calc_prob_in<R>(camps<R>(qs), qs->n_lanes, local_lane, buf, kScratchDoubles);
        });
    });
}

int qghip_qubit_processor_join(qghip_qubit_processor *p, qghip_qubit_states *qs,
                               qghip_qubit_states *const *qstates_list, int n_qstates,
                               int n_new_lanes) {
    return guarded([&] {
        if (p == nullptr) value_error("null processor handle");
        check_qs(qs, p->prec);
        if (n_qstates < 1 || qstates_list == nullptr) value_error("empty qubit states list");
        int total = 0;
        std::vector<int> lanes;
        for (int i = 0; i < n_qstates; ++i) {
            check_qs(qstates_list[i], p->prec);
            if (qstates_list[i] == qs) value_error("join destination is also a source");
            lanes.push_back(qstates_list[i]->n_lanes);
            total += qstates_list[i]->n_lanes;
        }
        if (total + n_new_lanes != qs->n_lanes) value_error("joined lane count mismatch");
        for (int i = 0; i < n_qstates; ++i) flush(qstates_list[i]);
        drop(qs);
        by_prec(p->prec, [&](auto r) {
            typedef decltype(r) R;
            std::vector<const Cx<R> *> srcs;
            for (int i = 0; i < n_qstates; ++i) srcs.push_back(camps<R>(qstates_list[i]));
            launch_join<R>(amps<R>(qs), qs->n_lanes, srcs, lanes);
        });
    });
}

int qghip_qubit_processor_decohere(qghip_qubit_processor *p, int value, double prob,
                                   qghip_qubit_states *qs, int local_lane) {
    return guarded([&] {
        if (p == nullptr) value_error("null processor handle");
        check_qs(qs, p->prec);
        check_lane(qs, local_lane);
        if (value != 0 && value != 1) value_error("value must be 0 or 1");
        flush(qs);
        touch(qs);
        // norm in FP64, rounded to the state precision (CPUQubitProcessor.cpp:227, :235). The
        // collapse is recorded and launched by the next call that touches qs (flush), or
        // together with the probability when that call is calc_probability.
        qs->dec_lane = local_lane;
        qs->dec_value = value;
        qs->dec_norm = value == 0 ? 1. / std::sqrt(prob) : 1. / std::sqrt(1. - prob);
        pending_states().push_back(qs);
    });
}

int qghip_qubit_processor_decohere_and_separate(qghip_qubit_processor *p, int value, double prob,
                                                qghip_qubit_states *qs0, qghip_qubit_states *qs1,
                                                qghip_qubit_states *qs, int local_lane) {
    return guarded([&] {
        if (p == nullptr) value_error("null processor handle");
        check_qs(qs, p->prec);
        check_qs(qs0, p->prec);
        check_qs(qs1, p->prec);
        check_lane(qs, local_lane);
        if (value != 0 && value != 1) value_error("value must be 0 or 1");
        if (qs0->n_lanes != qs->n_lanes - 1 || qs1->n_lanes != 1)
            value_error("separated qubit states have wrong sizes");
        flush(qs);
        drop(qs0);
        drop(qs1);
        const double norm = value == 0 ? 1. / std::sqrt(prob) : 1. / std::sqrt(1. - prob);
        by_prec(p->prec, [&](auto r) {
            typedef decltype(r) R;
            launch_decohere_and_separate<R>(amps<R>(qs0), camps<R>(qs), qs->n_lanes, local_lane,
                                            value, (R)norm);
            launch_set_basis1<R>(amps<R>(qs1), value);
        });
    });
}

int qghip_qubit_processor_calc_pauli_probability(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                                 uint64_t x_lanes, uint64_t z_lanes, double *prob) {
    return guarded([&] {
        if (p == nullptr || prob == nullptr) value_error("null argument");
        check_qs(qs, p->prec);
        check_pauli_lanes(qs, x_lanes, z_lanes);
        flush(qs);
        by_prec(p->prec, [&](auto r) {
            typedef decltype(r) R;
            *prob = calc_pauli_prob<R>(camps<R>(qs), qs->n_lanes, x_lanes, z_lanes);
        });
    });
}

int qghip_qubit_processor_pauli_decohere(qghip_qubit_processor *p, int value, double prob,
                                         qghip_qubit_states *qs, uint64_t x_lanes, uint64_t z_lanes) {
    return guarded([&] {
        if (p == nullptr) value_error("null processor handle");
        check_qs(qs, p->prec);
        check_pauli_lanes(qs, x_lanes, z_lanes);
        if (value != 0 && value != 1) value_error("value must be 0 or 1");
        flush(qs);
        touch(qs);
        by_prec(p->prec, [&](auto r) {
            typedef decltype(r) R;
            launch_pauli_decohere<R>(amps<R>(qs), qs->n_lanes, x_lanes, z_lanes, value, prob);
        });
    });
}

int qghip_qubit_processor_apply_reset(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                      int local_lane) {
    return guarded([&] {
        if (p == nullptr) value_error("null processor handle");
        check_qs(qs, p->prec);
        check_lane(qs, local_lane);
        flush(qs);
        touch(qs);
        by_prec(p->prec, [&](auto r) {
            typedef decltype(r) R;
            launch_reset<R>(amps<R>(qs), qs->n_lanes, local_lane);
        });
    });
}

int qghip_qubit_processor_apply_gate(qghip_qubit_processor *p, int factory_id, const double *args,
                                     int n_args, int adjoint, qghip_qubit_states *qs,
                                     int local_lane) {
    return guarded([&] {
        if (p == nullptr) value_error("null processor handle");
        GateMatrix m = make_gate_matrix(factory_id, args, n_args, adjoint != 0);
        apply_matrix(p, p->prec, qs, m, nullptr, 0, local_lane);
    });
}

int qghip_qubit_processor_apply_controlled_gate(qghip_qubit_processor *p, int factory_id,
                                                const double *args, int n_args, int adjoint,
                                                qghip_qubit_states *qs,
                                                const int *local_control_lanes, int n_controls,
                                                int local_target_lane) {
    return guarded([&] {
        if (p == nullptr) value_error("null processor handle");
        GateMatrix m = make_gate_matrix(factory_id, args, n_args, adjoint != 0);
        apply_matrix(p, p->prec, qs, m, local_control_lanes, n_controls, local_target_lane);
    });
}

int qghip_qubit_processor_apply_matrix(qghip_qubit_processor *p, const double *mat8,
                                       qghip_qubit_states *qs, const int *local_control_lanes,
                                       int n_controls, int local_target_lane) {
    return guarded([&] {
        if (p == nullptr || mat8 == nullptr) value_error("null argument");
        apply_matrix(p, p->prec, qs, gate_matrix_from_array(mat8), local_control_lanes, n_controls,
                     local_target_lane);
    });
}

int qghip_qubit_processor_apply_unitary(qghip_qubit_processor *p, qghip_qubit_states *qs,
                                        const int *lanes, int k, const double *mat) {
    return guarded([&] {
        if (p == nullptr || lanes == nullptr || mat == nullptr) value_error("null argument");
        check_qs(qs, p->prec);
        if (k < 1 || k > 6 || k > qs->n_lanes) value_error("unitary width out of range");
        std::vector<int> ls(lanes, lanes + k);
        for (int l : ls) check_lane(qs, l);
        std::vector<int> sorted(ls);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) value_error("duplicate unitary lanes");
        flush(qs);
        touch(qs);
        const int D = 1 << k;
        std::vector<std::complex<double>> m(D * D);
        for (int i = 0; i < D * D; ++i) m[i] = std::complex<double>(mat[2 * i], mat[2 * i + 1]);
        by_prec(p->prec, [&](auto r) {
            typedef decltype(r) R;
            launch_unitary<R>(amps<R>(qs), qs->n_lanes, ls, m);
        });
    });
}

int qghip_qubits_states_getter_new(int prec, qghip_qubits_states_getter **out) {
    return guarded([&] {
        if (out == nullptr) value_error("null argument");
        qghip_qubits_states_getter *g = new qghip_qubits_states_getter;
        g->prec = to_prec(prec);
        *out = g;
    });
}

int qghip_qubits_states_getter_delete(qghip_qubits_states_getter *g) {
    return guarded([&] { delete g; });
}

int qghip_qubits_states_getter_get_states(qghip_qubits_states_getter *g, void *array,
                                          int64_t array_size, int64_t array_offset, int mathop,
                                          const int *lane_tables,
                                          qghip_qubit_states *const *qstates_list, int n_qstates,
                                          uint64_t empty_lane_mask, int n_qregs,
                                          int64_t n_states, int64_t start, int64_t step) {
    return guarded([&] {
        if (g == nullptr || array == nullptr) value_error("null argument");
        if (mathop != QGHIP_MATHOP_NULL && mathop != QGHIP_MATHOP_PROB) value_error("Unknown math op.");
        if (n_states < 0 || array_offset < 0 || array_size - array_offset < n_states)
            value_error("array size too small.");
        if (n_qregs < 0 || n_qregs > 62) value_error("n_qregs out of range");
        if (n_states == 0) return;
        const int64_t n_ext = int64_t(1) << n_qregs;
        const int64_t end = start + step * (n_states - 1);
        if (start < 0 || n_ext <= start || end < 0 || n_ext <= end) value_error("value out of range");
        const size_t real_size = g->prec == precFP64 ? 8 : 4;
        const size_t item = mathop == QGHIP_MATHOP_NULL ? 2 * real_size : real_size;
        char *out = static_cast<char *>(array) + item * (size_t)array_offset;
        if (n_qstates == 0) {
            // no state vector: |0...0> (glue.cpp:480-495)
            std::memset(out, 0, item * (size_t)n_states);
            if (start == 0) {
                if (real_size == 4) {
                    const float one = 1.f;
                    std::memcpy(out, &one, 4);
                } else {
                    const double one = 1.;
                    std::memcpy(out, &one, 8);
                }
            }
            return;
        }
        std::vector<StateView> views = make_views(lane_tables, qstates_list, n_qstates, g->prec);
        by_prec(g->prec, [&](auto r) {
            typedef decltype(r) R;
            get_states<R>(out, mathop, views, empty_lane_mask, n_states, start, step);
        });
    });
}

int qghip_qubits_states_getter_prepare_prob_array(qghip_qubits_states_getter *g, void *prob,
                                                  int64_t array_size, const int *lane_tables,
                                                  qghip_qubit_states *const *qstates_list,
                                                  int n_qstates, int n_lanes, int n_hidden_lanes) {
    return guarded([&] {
        if (g == nullptr || prob == nullptr) value_error("null argument");
        if (n_lanes < 0 || n_hidden_lanes < 0 || n_lanes + n_hidden_lanes > 62)
            value_error("lane counts out of range");
        const int64_t n_states = int64_t(1) << n_lanes;
        if (array_size < n_states) value_error("array size too small.");
        std::vector<StateView> views = make_views(lane_tables, qstates_list, n_qstates, g->prec);
        by_prec(g->prec, [&](auto r) {
            typedef decltype(r) R;
            const size_t bytes = sizeof(R) * (size_t)n_states;
            void *d = dev_alloc(bytes);
            try {
                launch_prob_array<R>(d, false, views, n_lanes, n_hidden_lanes);
                QG_HIP(hipMemcpyAsync(prob, d, bytes, hipMemcpyDeviceToHost, stream()));
                sync();
            } catch (...) {
                dev_free(d, bytes);
                throw;
            }
            dev_free(d, bytes);
        });
    });
}

int qghip_qubits_states_getter_create_sampling_pool(qghip_qubits_states_getter *g,
                                                    const int *lane_tables,
                                                    qghip_qubit_states *const *qstates_list,
                                                    int n_qstates, int n_lanes,
                                                    int n_hidden_lanes, const int *empty_lanes,
                                                    int n_empty_lanes, qghip_sampling_pool **out) {
    return guarded([&] {
        if (g == nullptr || out == nullptr) value_error("null argument");
        if (n_lanes < 1 || n_hidden_lanes < 0 || n_lanes + n_hidden_lanes > 62)
            value_error("lane counts out of range");
        if (n_empty_lanes < 0 || (n_empty_lanes > 0 && empty_lanes == nullptr))
            value_error("bad empty lane list");
        std::vector<StateView> views = make_views(lane_tables, qstates_list, n_qstates, g->prec);
        qghip_sampling_pool *pool = new qghip_sampling_pool;
        pool->empty_lanes.assign(empty_lanes, empty_lanes + n_empty_lanes);
        try {
            bool direct = n_hidden_lanes == 0 && views.size() == 1 && views[0].n_lanes == n_lanes;
            for (int l = 0; direct && l < n_lanes; ++l) direct = views[0].local_to_ext[l] == l;
            bool bounded = false;
            if (direct && g_pool_workspace == QGHIP_POOL_BOUNDED) {
                bounded = true;
            } else if (direct && g_pool_workspace == QGHIP_POOL_AUTO) {
                // bounded when the probability array would not fit beside what is allocated
                size_t free_b = 0, total_b = 0;
                QG_HIP(hipMemGetInfo(&free_b, &total_b));
                const size_t avail = free_b + dev_cached_bytes();
                const size_t headroom = size_t(1) << 30;
                bounded = pool_bytes(n_lanes, true) + headroom > avail;
            }
            by_prec(g->prec, [&](auto r) {
                typedef decltype(r) R;
                if (direct) {
                    pool_build_from_state<R>(pool->pool, static_cast<const Cx<R> *>(views[0].ptr), n_lanes,
                                             bounded);
                    if (bounded) {
                        pool->src = qstates_list[0];
                        pool->src->readers.push_back(pool);
                    }
                } else {
                    pool_alloc(pool->pool, n_lanes);
                    try {
                        launch_prob_array<R>(pool->pool.p, true, views, n_lanes, n_hidden_lanes);
                        pool_build_from_probs(pool->pool, n_lanes);
                    } catch (...) {
                        pool_free(pool->pool);
                        throw;
                    }
                }
            });
        } catch (...) {
            delete pool;
            throw;
        }
        *out = pool;
    });
}

int qghip_sampling_pool_sample(qghip_sampling_pool *pool, int64_t *observations, int64_t array_size,
                               int64_t n_samples, const double *randnum) {
    return guarded([&] {
        if (pool == nullptr || observations == nullptr || randnum == nullptr)
            value_error("null argument");
        if (n_samples < 0 || array_size < n_samples) value_error("array size too small.");
        if (pool->stale)
            value_error("the state vector this bounded sampling pool draws from has changed or been freed; "
                        "create the pool again");
        pool_sample(pool->pool, reinterpret_cast<i64 *>(observations), n_samples, randnum,
                    pool->empty_lanes);
    });
}

int qghip_sampling_pool_delete(qghip_sampling_pool *pool) {
    return guarded([&] {
        if (pool == nullptr) return;
        if (pool->src != nullptr) {
            auto &r = pool->src->readers;
            r.erase(std::remove(r.begin(), r.end(), pool), r.end());
        }
        pool_free(pool->pool);
        delete pool;
    });
}

int qghip_profile_gate(qghip_qubit_processor *p, qghip_qubit_states *qs, const double *mat8,
                       const int *local_control_lanes, int n_controls, int local_target_lane,
                       int reps, double *avg_ms) {
    return guarded([&] {
        if (p == nullptr || mat8 == nullptr || avg_ms == nullptr || reps < 1)
            value_error("bad argument");
        GateMatrix m = gate_matrix_from_array(mat8);
        flush(qs);
        hipStream_t s = stream();
        hipEvent_t e0, e1;
        QG_HIP(hipEventCreate(&e0));
        QG_HIP(hipEventCreate(&e1));
        QG_HIP(hipEventRecord(e0, s));
        for (int i = 0; i < reps; ++i)
            apply_matrix(nullptr, p->prec, qs, m, local_control_lanes, n_controls, local_target_lane);
        QG_HIP(hipEventRecord(e1, s));
        QG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        QG_HIP(hipEventElapsedTime(&ms, e0, e1));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        *avg_ms = ms / reps;
    });
}

int qghip_profile_copy(int64_t bytes, int reps, double *avg_ms) {
    return guarded([&] {
        if (avg_ms == nullptr || reps < 1 || bytes < 16 || bytes % 16 != 0) value_error("bad argument");
        void *src = dev_alloc((size_t)bytes), *dst = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        try {
            dst = dev_alloc((size_t)bytes);
            hipStream_t s = stream();
            QG_HIP(hipMemsetAsync(src, 0, (size_t)bytes, s));
            launch_stream_copy(dst, src, (u64)bytes);   // warm-up
            QG_HIP(hipEventCreate(&e0));
            QG_HIP(hipEventCreate(&e1));
            QG_HIP(hipEventRecord(e0, s));
            for (int i = 0; i < reps; ++i) launch_stream_copy(dst, src, (u64)bytes);
            QG_HIP(hipEventRecord(e1, s));
            QG_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            QG_HIP(hipEventElapsedTime(&ms, e0, e1));
            *avg_ms = ms / reps;
        } catch (...) {
            sync();
            if (e0) hipEventDestroy(e0);
            if (e1) hipEventDestroy(e1);
            dev_free(src, (size_t)bytes);
            if (dst) dev_free(dst, (size_t)bytes);
            throw;
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        dev_free(src, (size_t)bytes);
        dev_free(dst, (size_t)bytes);
    });
}

}  // extern "C"
