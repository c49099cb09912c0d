// FP64 2x2 gate matrices (host).
#pragma once

#include <complex>

#include "qg_common.hpp"

namespace qg {

struct GateMatrix {
    std::complex<double> m[2][2];
};

inline bool is_diagonal(const GateMatrix &g) { return g.m[0][1] == 0. && g.m[1][0] == 0.; }

// Factory id for a gate type name ("U", "U2", "U1", "ID", "X", "Y", "Z", "H", "S", "T",
// "RX", "RY", "RZ", "ExpiI", "ExpiZ", "SH"), or -1.
int matrix_factory_id(const char *name);
int matrix_factory_n_args(int id);
GateMatrix make_gate_matrix(int id, const double *args, int n_args, bool adjoint);
// Row-major (m00, m01, m10, m11) as interleaved (re, im) pairs.
GateMatrix gate_matrix_from_array(const double *mat8);
void gate_matrix_to_array(const GateMatrix &m, double *mat8);
// exp(-i theta G / 2) for G = X (1), Y (2), Z (3): RX, RY, RZ.
GateMatrix rotation(int pauli, double theta);
// Conjugate transpose.
GateMatrix dagger(const GateMatrix &g);

// Rounded to the state precision, as the reference does before applying a gate
// (qgate/simulator/src/CPUQubitProcessor.cpp:312, DeviceTypes.h:105-112).
template <class R>
Mat2<R> round_matrix(const GateMatrix &g) {
    Mat2<R> m;
    m.m00 = cx<R>((R)g.m[0][0].real(), (R)g.m[0][0].imag());
    m.m01 = cx<R>((R)g.m[0][1].real(), (R)g.m[0][1].imag());
    m.m10 = cx<R>((R)g.m[1][0].real(), (R)g.m[1][0].imag());
    m.m11 = cx<R>((R)g.m[1][1].real(), (R)g.m[1][1].imag());
    return m;
}

}  // namespace qg
