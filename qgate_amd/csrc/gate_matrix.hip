// FP64 2x2 gate matrices, built on the host.
// Same formulas and operation order as the reference CPU/CUDA runtimes' matrix factories
// (qgate/simulator/src/GateMatrix.cpp:13-158, selected by name as in
// glue.cpp:92-134 register_matrix_factory); the adjoint is taken in FP64 before the matrix
// is rounded to the state precision (glue.cpp:390-398).
#include <cmath>
#include <complex>
#include <cstring>

#include "gate_matrix.hpp"

namespace qg {

namespace {

typedef std::complex<double> cd;
const cd J(0., 1.);

void set(GateMatrix &m, cd a00, cd a01, cd a10, cd a11) {
    m.m[0][0] = a00;
    m.m[0][1] = a01;
    m.m[1][0] = a10;
    m.m[1][1] = a11;
}

struct Factory {
    const char *name;
    int n_args;
};

const Factory kFactories[] = {
    {"U", 3},  {"U2", 2}, {"U1", 1}, {"ID", 0}, {"X", 0},  {"Y", 0},     {"Z", 0},     {"H", 0},
    {"S", 0},  {"T", 0},  {"RX", 1}, {"RY", 1}, {"RZ", 1}, {"ExpiI", 1}, {"ExpiZ", 1}, {"SH", 0},
};
const int kNFactories = sizeof(kFactories) / sizeof(kFactories[0]);

}  // namespace

int matrix_factory_id(const char *name) {
    for (int i = 0; i < kNFactories; ++i)
        if (std::strcmp(kFactories[i].name, name) == 0) return i;
    return -1;
}

int matrix_factory_n_args(int id) {
    return (0 <= id && id < kNFactories) ? kFactories[id].n_args : -1;
}

GateMatrix make_gate_matrix(int id, const double *a, int n_args, bool adjoint) {
    QG_REQUIRE(0 <= id && id < kNFactories, "Unknown gate type (factory id %d).", id);
    QG_REQUIRE(n_args == kFactories[id].n_args, "gate %s takes %d argument(s), %d given",
               kFactories[id].name, kFactories[id].n_args, n_args);
    GateMatrix m;
    const double r = std::sqrt(0.5);
    switch (id) {
    case 0: {  // U(theta, phi, lambda), global phase as in qgate (not OpenQASM)
        double c = std::cos(a[0] * 0.5), s = std::sin(a[0] * 0.5);
        double l2 = a[2] * 0.5, p2 = a[1] * 0.5;
        set(m, std::exp(J * (-l2 - p2)) * c, -std::exp(J * (l2 - p2)) * s,
            std::exp(J * (-l2 + p2)) * s, std::exp(J * (l2 + p2)) * c);
        break;
    }
    case 1: {  // U2(phi, lambda) = U(pi/2, phi, lambda)
        double l2 = a[1] * 0.5, p2 = a[0] * 0.5;
        set(m, std::exp(J * (-l2 - p2)) * r, -std::exp(J * (l2 - p2)) * r,
            std::exp(J * (-l2 + p2)) * r, std::exp(J * (l2 + p2)) * r);
        break;
    }
    case 2: set(m, 1., 0., 0., std::exp(J * a[0])); break;             // U1
    case 3: set(m, 1., 0., 0., 1.); break;                             // ID
    case 4: set(m, 0., 1., 1., 0.); break;                             // X
    case 5: set(m, 0., -J, J, 0.); break;                              // Y
    case 6: set(m, 1., 0., 0., -1.); break;                            // Z
    case 7: set(m, 1. * r, 1. * r, 1. * r, -1. * r); break;            // H
    case 8: set(m, 1., 0., 0., J); break;                              // S
    case 9: set(m, 1., 0., 0., std::exp(M_PI * 0.25 * J)); break;      // T
    case 10: {                                                         // RX
        double c = std::cos(a[0] / 2.), s = std::sin(a[0] / 2.);
        set(m, c, -J * s, -J * s, c);
        break;
    }
    case 11: {                                                         // RY
        double c = std::cos(a[0] / 2.), s = std::sin(a[0] / 2.);
        set(m, c, -s, s, c);
        break;
    }
    case 12: {                                                         // RZ
        double t2 = a[0] / 2.;
        set(m, std::exp(-J * t2), 0., 0., std::exp(J * t2));
        break;
    }
    case 13: {                                                         // ExpiI
        cd d = std::exp(a[0] * J);
        set(m, d, 0., 0., d);
        break;
    }
    case 14: set(m, std::exp(a[0] * J), 0., 0., std::exp(-a[0] * J)); break;  // ExpiZ
    case 15: set(m, 1. * r, 1. * r, J * r, -J * r); break;                     // SH
    }
    if (adjoint) {
        cd t = m.m[0][1];
        m.m[0][0] = std::conj(m.m[0][0]);
        m.m[0][1] = std::conj(m.m[1][0]);
        m.m[1][0] = std::conj(t);
        m.m[1][1] = std::conj(m.m[1][1]);
    }
    return m;
}

GateMatrix gate_matrix_from_array(const double *mat8) {
    GateMatrix m;
    for (int i = 0; i < 4; ++i) m.m[i / 2][i % 2] = cd(mat8[2 * i], mat8[2 * i + 1]);
    return m;
}

void gate_matrix_to_array(const GateMatrix &m, double *mat8) {
    for (int i = 0; i < 4; ++i) {
        mat8[2 * i] = m.m[i / 2][i % 2].real();
        mat8[2 * i + 1] = m.m[i / 2][i % 2].imag();
    }
}

GateMatrix rotation(int pauli, double theta) {
    GateMatrix m;
    const double c = std::cos(theta / 2.), s = std::sin(theta / 2.);
    const std::complex<double> J(0., 1.);
    if (pauli == 1) {          // RX
        m.m[0][0] = c; m.m[0][1] = -J * s; m.m[1][0] = -J * s; m.m[1][1] = c;
    } else if (pauli == 2) {   // RY
        m.m[0][0] = c; m.m[0][1] = -s; m.m[1][0] = s; m.m[1][1] = c;
    } else {                   // RZ
        m.m[0][0] = std::exp(-J * (theta / 2.)); m.m[0][1] = 0.; m.m[1][0] = 0.;
        m.m[1][1] = std::exp(J * (theta / 2.));
    }
    return m;
}

GateMatrix dagger(const GateMatrix &g) {
    GateMatrix d;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) d.m[r][c] = std::conj(g.m[c][r]);
    return d;
}

}  // namespace qg
