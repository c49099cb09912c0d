"""Expectation values of Pauli sums and their gradients by adjoint differentiation on the GPU.

    circ = ParametricCircuit(n_qubits, dtype=np.float32)
    circ.ry(q, param_index); circ.rz(q, param_index); circ.gate(gate_type.Z(), q, controls=[c])
    H = PauliSum.zz_x(a, b)                     # sum a_i Z_i Z_{i+1} + sum b_i X_i
    value = circ.expectation(params, H)
    value, grad = circ.gradient(params, H)
    loss = Expectation.apply(torch_params, circ, H)   # torch.autograd.Function

Rotations follow qgate's gate definitions (qgate/simulator/src/GateMatrix.cpp:96-124):
RX/RY/RZ(theta) = exp(-i theta G / 2), so d/dtheta of a rotation is -i/2 G times it. The
work runs in the HIP library (csrc/adjoint.hip): forward pass on the streaming gate kernels,
lambda = H psi in one pass, then a backward sweep that un-computes psi and lambda together and
reduces the gradient terms in FP64; only two state vectors are kept.
"""
import ctypes

import numpy as np

from .simulator import hipruntime
from .model import gate_type as gtype

_PAULI = {'X': 1, 'Y': 2, 'Z': 3}
_bound = False


def _lib():
    global _bound
    L = hipruntime.lib()
    if not _bound:
        H = ctypes.c_void_p
        P_I = ctypes.POINTER(ctypes.c_int)
        P_D = ctypes.POINTER(ctypes.c_double)
        P_U = ctypes.POINTER(ctypes.c_uint64)
        sig = {
            'qghip_circuit_new': [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
            'qghip_circuit_delete': [H],
            'qghip_circuit_add_gate': [H, P_D, P_I, ctypes.c_int, ctypes.c_int],
            'qghip_circuit_add_rotation': [H, ctypes.c_int, ctypes.c_int, P_I, ctypes.c_int, ctypes.c_int],
            'qghip_circuit_n_params': [H, P_I],
            'qghip_circuit_set_fusion': [H, ctypes.c_int],
            'qghip_circuit_expectation': [H, P_D, ctypes.c_int, ctypes.c_int, P_D, P_U, P_U, P_D],
            'qghip_circuit_backward': [H, P_D, ctypes.c_int, P_D],
            'qghip_circuit_gradient': [H, P_D, ctypes.c_int, ctypes.c_int, P_D, P_U, P_U, P_D, P_D],
            'qghip_circuit_state': [H, ctypes.c_void_p, ctypes.c_int64],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = args
        _bound = True
    return L


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _uptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class PauliSum:
    """H = sum_k c_k P_k; each term is (coef, {qubit: 'X' | 'Y' | 'Z'})."""

    def __init__(self, terms):
        coefs, xm, zm = [], [], []
        for coef, ops in terms:
            x = z = 0
            for q, p in ops.items():
                if p == 'X':
                    x |= 1 << q
                elif p == 'Z':
                    z |= 1 << q
                elif p == 'Y':
                    x |= 1 << q
                    z |= 1 << q
                elif p != 'I':
                    raise ValueError('unknown Pauli {}'.format(p))
            coefs.append(float(coef))
            xm.append(x)
            zm.append(z)
        self.terms = list(terms)
        self.coefs = np.ascontiguousarray(coefs, np.float64)
        self.xmask = np.ascontiguousarray(xm, np.uint64)
        self.zmask = np.ascontiguousarray(zm, np.uint64)

    def key(self):
        """Value identity of the observable (coefficients and masks)."""
        return (self.coefs.tobytes(), self.xmask.tobytes(), self.zmask.tobytes())

    @classmethod
    def zz_x(cls, a, b):
        """sum_i a_i Z_i Z_{i+1} + sum_i b_i X_i (open chain)."""
        terms = [(a[i], {i: 'Z', i + 1: 'Z'}) for i in range(len(a))]
        terms += [(b[i], {i: 'X'}) for i in range(len(b))]
        return cls(terms)

    def matrix(self, n):
        """Dense 2^n x 2^n matrix (small n, for tests)."""
        paulis = {'I': np.eye(2), 'X': np.array([[0, 1], [1, 0]]), 'Y': np.array([[0, -1j], [1j, 0]]),
                  'Z': np.diag([1., -1.])}
        out = np.zeros((1 << n, 1 << n), np.complex128)
        for coef, ops in self.terms:
            m = np.ones((1, 1))
            for q in range(n - 1, -1, -1):
                m = np.kron(m, paulis[ops.get(q, 'I')])
            out += coef * m
        return out


class ParametricCircuit:
    """Gate list with rotation parameters, held by the HIP library.

    `ops` records what was appended: ('RX' | 'RY' | 'RZ', target, param_index, controls) and
    ('G', target, gate type or 2x2 matrix, controls, adjoint)."""

    def __init__(self, n_qubits, dtype=np.float64):
        self.n_qubits = n_qubits
        self.dtype = dtype
        L = _lib()
        h = ctypes.c_void_p()
        prec = hipruntime.FP64 if dtype == np.float64 else hipruntime.FP32
        if not hipruntime.initialized:
            hipruntime.module_init()
        hipruntime.check(L.qghip_circuit_new(prec, n_qubits, ctypes.byref(h)))
        self.ptr = h.value
        self.ops = []
        self.n_params = 0
        self._last = None   # (params, observable key) of the forward pass the library holds

    def __del__(self):
        self.delete()

    def delete(self):
        rt = hipruntime   # None while the interpreter shuts down (the process frees the memory)
        if getattr(self, 'ptr', None) is not None and rt is not None and rt._lib is not None:
            rt._lib.qghip_circuit_delete(self.ptr)
        self.ptr = None

    def set_fusion(self, mode):
        """'tile' (default): tile passes for the forward pass and the backward sweep; 'off': one
        launch per gate and per backward step."""
        modes = {'off': 0, 'tile': 1}
        if mode not in modes:
            raise ValueError('fusion mode must be one of {}'.format(sorted(modes)))
        hipruntime.check(_lib().qghip_circuit_set_fusion(self.ptr, modes[mode]))
        return self

    def _rot(self, pauli, qubit, param, controls):
        ctrl = list(controls)
        hipruntime.check(_lib().qghip_circuit_add_rotation(
            self.ptr, _PAULI[pauli], int(param), hipruntime._ints(ctrl), len(ctrl), int(qubit)))
        self.ops.append(('R' + pauli, qubit, int(param), tuple(ctrl)))
        self.n_params = max(self.n_params, int(param) + 1)
        return self

    def rx(self, qubit, param, controls=()):
        return self._rot('X', qubit, param, controls)

    def ry(self, qubit, param, controls=()):
        return self._rot('Y', qubit, param, controls)

    def rz(self, qubit, param, controls=()):
        return self._rot('Z', qubit, param, controls)

    def gate(self, gate, qubit, controls=(), adjoint=False):
        """Fixed gate: a qgate gate type (e.g. gate_type.H()) or a 2x2 complex matrix."""
        if isinstance(gate, gtype._Unary):
            mat = hipruntime.gate_matrix(gate, adjoint)
        else:
            mat = np.asarray(gate, np.complex128)
            if adjoint:
                mat = mat.conj().T
        m = mat.reshape(4)
        flat = np.empty(8)
        flat[0::2], flat[1::2] = m.real, m.imag
        ctrl = list(controls)
        hipruntime.check(_lib().qghip_circuit_add_gate(self.ptr, hipruntime._doubles(flat),
                                                       hipruntime._ints(ctrl), len(ctrl), int(qubit)))
        self.ops.append(('G', qubit, gate, tuple(ctrl), bool(adjoint)))
        return self

    def cz(self, a, b):
        return self.gate(gtype.Z(), b, controls=[a])

    def _params(self, params):
        p = np.ascontiguousarray(np.asarray(params, np.float64).reshape(-1))
        if len(p) < self.n_params:
            raise ValueError('expected {} parameters, got {}'.format(self.n_params, len(p)))
        return p

    def expectation(self, params, observable):
        p = self._params(params)
        out = ctypes.c_double()
        self._last = None
        hipruntime.check(_lib().qghip_circuit_expectation(
            self.ptr, _dptr(p), len(p), len(observable.coefs), _dptr(observable.coefs),
            _uptr(observable.xmask), _uptr(observable.zmask), ctypes.byref(out)))
        self._last = (p.tobytes(), observable.key())
        return out.value

    def backward(self, params):
        """Gradient for the observable of the last expectation() call."""
        p = self._params(params)
        grad = np.zeros(max(1, self.n_params), np.float64)
        hipruntime.check(_lib().qghip_circuit_backward(self.ptr, _dptr(p), len(p), _dptr(grad)))
        return grad[:self.n_params]

    def gradient(self, params, observable):
        p = self._params(params)
        out = ctypes.c_double()
        grad = np.zeros(max(1, self.n_params), np.float64)
        self._last = None
        hipruntime.check(_lib().qghip_circuit_gradient(
            self.ptr, _dptr(p), len(p), len(observable.coefs), _dptr(observable.coefs),
            _uptr(observable.xmask), _uptr(observable.zmask), ctypes.byref(out), _dptr(grad)))
        self._last = (p.tobytes(), observable.key())
        return out.value, grad[:self.n_params]

    def gradient_for(self, params, observable):
        """d<H>/dparams for `observable`: the backward sweep alone when the library still
        holds the forward pass of these parameters and this observable, else forward and
        backward."""
        p = self._params(params)
        if self._last == (p.tobytes(), observable.key()):
            return self.backward(p)
        return self.gradient(p, observable)[1]

    def state(self):
        """Final state of the last expectation call (qubit q = bit q)."""
        ctype = np.complex128 if self.dtype == np.float64 else np.complex64
        out = np.empty(1 << self.n_qubits, ctype)
        hipruntime.check(_lib().qghip_circuit_state(self.ptr, out.ctypes.data, len(out)))
        return out


def hea_circuit(n, layers, dtype=np.float64):
    """Hardware-efficient ansatz: per layer RY then RZ on every qubit, then a CZ ring;
    parameter order as qgate_amd.workloads.hea_ansatz."""
    c = ParametricCircuit(n, dtype)
    k = 0
    for _ in range(layers):
        for q in range(n):
            c.ry(q, k)
            k += 1
        for q in range(n):
            c.rz(q, k)
            k += 1
        for q in range(n):
            c.cz(q, (q + 1) % n)
    return c


try:
    import torch

    class Expectation(torch.autograd.Function):
        """<H>(params) with an adjoint-differentiation backward; params is a 1-D tensor."""

        @staticmethod
        def forward(ctx, params, circuit, observable):
            p = params.detach().to('cpu', torch.float64).numpy()
            value = circuit.expectation(p, observable)
            ctx.circuit = circuit
            ctx.observable = observable
            ctx.params = p.copy()
            ctx.param_dtype, ctx.param_device = params.dtype, params.device
            return torch.tensor(value, dtype=params.dtype, device=params.device)

        @staticmethod
        def backward(ctx, grad_output):
            # the observable of this forward: another expectation call on the same circuit
            # between forward and backward (or a second observable in the same loss) re-runs
            # the forward pass for this one
            g = ctx.circuit.gradient_for(ctx.params, ctx.observable)
            grad = torch.as_tensor(g, dtype=ctx.param_dtype).to(ctx.param_device)
            return grad_output * grad, None, None
except ImportError:  # torch is only needed for the autograd wrapper
    Expectation = None
