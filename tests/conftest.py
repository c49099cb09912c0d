"""Shared fixtures. GPU tests are marked `gpu`; the CPU suite runs with -m "not gpu".

`backend` parametrises a test over the runtimes it runs on:
  port  the CPU oracle (C restatement of the reference CPU runtime) — CPU suite,
  hip   the HIP runtime on the GPU — marked gpu.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP runtime)')
    config.addinivalue_line('markers', 'slow: large-state GPU cases')


def _oracle_built():
    from oracle import cpuruntime
    if not os.path.exists(cpuruntime.PORT_LIB):
        cpuruntime.build(reference=False)
    return cpuruntime


@pytest.fixture(scope='session')
def oracle():
    return _oracle_built()


@pytest.fixture(scope='session')
def port_module(oracle):
    return oracle.port()


@pytest.fixture(scope='session')
def hip_module():
    from qgate_amd.simulator import hipruntime
    hipruntime.lib()          # fails loudly when the library is not built
    return hipruntime


def make_module(name):
    if name == 'port':
        return _oracle_built().port()
    if name == 'reference':
        return _oracle_built().reference()
    if name == 'hip':
        from qgate_amd.simulator import hipruntime
        hipruntime.lib()
        return hipruntime
    raise ValueError(name)


BACKENDS = ['port', pytest.param('hip', marks=pytest.mark.gpu)]


@pytest.fixture(params=BACKENDS)
def backend(request):
    return request.param


@pytest.fixture
def simulator(backend):
    """Factory: simulator(**prefs) on the parametrised backend."""
    from qgate_amd.simulator.simulator import Simulator
    module = make_module(backend)

    def factory(**prefs):
        return Simulator(module, **prefs)
    factory.runtime = backend
    return factory


@pytest.fixture(autouse=True)
def _seed_numpy():
    np.random.seed(12345)
