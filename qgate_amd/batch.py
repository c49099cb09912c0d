"""Batches of parameterised circuits on the GPU: <H> and adjoint gradients of many circuits
of uneven qubit counts in one replayed HIP graph (csrc/k_batch.hip).

    batch = AdjointBatch(np.float32)
    for n in sizes:
        batch.add(circuit_for[n], params, observable)     # circuit: a ParametricCircuit
    values, grads = batch.run()                            # grads: one array per item

Circuits built from the same ParametricCircuit object with observables of the same terms
share a group: their state vectors are stacked into one array and every tile-pass kernel
covers the whole group. The first run times the tile shape (register-resident or LDS passes,
tile size) and the LDS passes' workgroup and grid size per group and keeps the fastest
(`tuning` shows the choice), then captures the launch sequence of all groups into a graph
that later runs replay.
"""
import ctypes
import json

import numpy as np

from . import adjoint
from .simulator import hipruntime

_bound = False


def _lib():
    global _bound
    L = adjoint._lib()
    if not _bound:
        H = ctypes.c_void_p
        P_D = ctypes.POINTER(ctypes.c_double)
        P_U = ctypes.POINTER(ctypes.c_uint64)
        sig = {
            'qghip_batch_new': [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
            'qghip_batch_delete': [H],
            'qghip_batch_add': [H, H, P_D, ctypes.c_int, ctypes.c_int, P_D, P_U, P_U, ctypes.POINTER(ctypes.c_int)],
            'qghip_batch_set_params': [H, ctypes.c_int, P_D, ctypes.c_int],
            'qghip_batch_set_tuning': [H, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
            'qghip_batch_run': [H, ctypes.c_int, P_D, P_D, ctypes.c_int64],
            'qghip_batch_describe': [H, ctypes.c_char_p, ctypes.c_int],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = args
        _bound = True
    return L


class AdjointBatch:
    def __init__(self, dtype=np.float32):
        self.dtype = dtype
        L = _lib()
        if not hipruntime.initialized:
            hipruntime.module_init()
        h = ctypes.c_void_p()
        prec = hipruntime.FP64 if dtype == np.float64 else hipruntime.FP32
        hipruntime.check(L.qghip_batch_new(prec, ctypes.byref(h)))
        self.ptr = h.value
        self.n_params = []
        self._circuits = []   # keep the circuits alive with the batch

    def __del__(self):
        self.delete()

    def delete(self):
        rt = hipruntime   # None while the interpreter shuts down (the process frees the memory)
        if getattr(self, 'ptr', None) is not None and rt is not None and rt._lib is not None:
            rt._lib.qghip_batch_delete(self.ptr)
        self.ptr = None

    def __len__(self):
        return len(self.n_params)

    def add(self, circuit, params, observable):
        if circuit.dtype != self.dtype:
            raise ValueError('circuit precision does not match the batch')
        p = circuit._params(params)
        out = ctypes.c_int()
        hipruntime.check(_lib().qghip_batch_add(
            self.ptr, circuit.ptr, adjoint._dptr(p), len(p), len(observable.coefs), adjoint._dptr(observable.coefs),
            adjoint._uptr(observable.xmask), adjoint._uptr(observable.zmask), ctypes.byref(out)))
        self.n_params.append(circuit.n_params)
        self._circuits.append(circuit)
        return out.value

    def set_params(self, item, params):
        p = np.ascontiguousarray(np.asarray(params, np.float64).reshape(-1))
        hipruntime.check(_lib().qghip_batch_set_params(self.ptr, int(item), adjoint._dptr(p), len(p)))

    def set_tuning(self, tile_bits=0, threads=0, grid_per_cu=0, registers=-1):
        """Fix tuning knobs (0, or -1 for registers: let the autotuner search them).

        registers=1 keeps forward and sweep tiles in registers where a group's qubit count
        allows full-size tiles; 0 uses LDS passes only."""
        hipruntime.check(_lib().qghip_batch_set_tuning(self.ptr, tile_bits, threads, grid_per_cu, registers))

    def run(self, use_graph=True):
        """(values, grads): <H> per item and each item's gradient."""
        values = np.empty(len(self.n_params), np.float64)
        total = int(sum(self.n_params))
        flat = np.empty(max(1, total), np.float64)
        hipruntime.check(_lib().qghip_batch_run(self.ptr, 1 if use_graph else 0, adjoint._dptr(values),
                                                adjoint._dptr(flat), len(flat)))
        grads, off = [], 0
        for k in self.n_params:
            grads.append(flat[off:off + k].copy())
            off += k
        return values, grads

    @property
    def tuning(self):
        buf = ctypes.create_string_buffer(1 << 16)
        hipruntime.check(_lib().qghip_batch_describe(self.ptr, buf, len(buf)))
        return json.loads(buf.value.decode())
