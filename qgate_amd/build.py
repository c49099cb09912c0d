"""Build libqgate_hip.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

    python -m qgate_amd.build          # or qgate_amd.build.build()
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIBDIR = os.path.join(HERE, 'lib')
LIBNAME = 'libqgate_hip.so'
SOURCES = ['qg_runtime.hip', 'gate_matrix.hip', 'k_gates.hip', 'k_state.hip', 'k_getter.hip',
           'capi.hip', 'adjoint.hip', 'k_tile.hip', 'k_batch.hip', 'k_marginal.hip']
ARCH = os.environ.get('QGATE_HIP_ARCH', 'gfx950')

# -packed-fp32-ops: no v_pk_*_f32 instructions are generated.
HIPFLAGS = ['-std=c++17', '-O3', '-fPIC', '--offload-arch=' + ARCH, '-munsafe-fp-atomics',
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-Wall', '-Wno-unused-function', '-Wno-unused-command-line-argument']


def lib_path():
    return os.path.join(LIBDIR, LIBNAME)


def _hipcc():
    exe = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(exe):
        raise RuntimeError('hipcc not found')
    return exe


def _compile(src, obj, verbose):
    cmd = [_hipcc()] + HIPFLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for {}:\n{}{}'.format(src, r.stdout, r.stderr))
    return r.stderr


def build(verbose=False, jobs=None):
    """Compile every source and link the shared library; returns its path."""
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + '.o') for s in SOURCES]
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with ThreadPoolExecutor(jobs) as ex:
        warnings = list(ex.map(lambda so: _compile(so[0], so[1], verbose), zip(SOURCES, objs)))
    for w in warnings:
        if w.strip() and verbose:
            print(w, file=sys.stderr)
    tmp = lib_path() + '.tmp'
    cmd = [_hipcc(), '-shared', '-fPIC', '--offload-arch=' + ARCH] + objs + ['-o', tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n{}{}'.format(r.stdout, r.stderr))
    os.replace(tmp, lib_path())
    return lib_path()


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
