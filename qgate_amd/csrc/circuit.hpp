// Parameterised circuits held by the library (adjoint differentiation and batches).
#pragma once

#include <vector>

#include "gate_matrix.hpp"
#include "qg_common.hpp"

namespace qg {

// One gate of a parameterised circuit.
struct CircuitOp {
    int kind;              // 0 fixed matrix, 1 rotation exp(-i theta G / 2)
    int pauli;             // 1 X, 2 Y, 3 Z (rotations)
    int param;             // parameter index (rotations): theta = params[param]
    GateMatrix mat;        // fixed gates
    std::vector<int> ctrl;
    int target;
};

}  // namespace qg

struct qghip_circuit {
    qg::Prec prec;
    int n_qubits;
    std::vector<qg::CircuitOp> ops;
    int n_params = 0;
    void *psi = nullptr, *lam = nullptr;   // device state vectors
    size_t state_bytes = 0;
    // state left by the last expectation call: psi = final state, lam = H psi
    bool forward_valid = false;
    std::vector<double> forward_params;
    std::vector<double> obs_coefs;
    std::vector<uint64_t> obs_xmask, obs_zmask;
    int fusion = 1;                        // QGHIP_FUSION_TILE: tile passes; 0: one launch per gate
};
