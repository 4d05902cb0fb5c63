"""Build step for the native library libxomc.so (gfx950).

Compiles the HIP kernels and the C ABI with hipcc, the host seed code with
g++, and links ``pyxopto_amd/libxomc.so`` in place. Then fills the
safe-prime multiplier cache ``pyxopto_amd/data/safeprimes_base32.u32`` (the
table the MWC seeds are drawn from) if it is missing.

Run ``python -m pyxopto_amd.build_native`` or call :func:`build`.
"""
import os
import subprocess
import sys
import tempfile

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, 'csrc')
LIB_PATH = os.path.join(PKG_DIR, 'libxomc.so')
OFFLOAD_ARCH = os.environ.get('XO_OFFLOAD_ARCH', 'gfx950')

# Device code: IEEE binary32 semantics that the host oracle can reproduce
# (no contraction, correctly rounded division and square root, denormals
# kept) and no packed-FP32 instructions.
HIP_FLAGS = [
    '-O3', '-std=c++17', '-fPIC', f'--offload-arch={OFFLOAD_ARCH}',
    '-ffp-contract=off',
    '-fhip-fp32-correctly-rounded-divide-sqrt',
    '-fno-gpu-flush-denormals-to-zero',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function',
]

HIP_SOURCES = ['xo_kernels.hip', 'xo_batch.hip', 'xo_api.hip', 'xo_serve.hip',
               'xo_probe.hip', 'xo_mlp.hip', 'xo_sv.hip', 'xo_noalloc.hip',
               'xo_plan.hip']
CXX_SOURCES = ['xo_seeds.cpp']


def _run(cmd, verbose):
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True)
    out = '\n'.join(line for line in res.stdout.splitlines()
                    if 'not a recognized feature' not in line
                    and 'argument unused during compilation' not in line)
    if res.returncode != 0:
        raise RuntimeError('command failed ({}):\n{}\n{}'.format(
            res.returncode, ' '.join(cmd), out))
    if verbose and out.strip():
        print(out, flush=True)


def build_library(verbose: bool = False) -> str:
    hipcc = os.environ.get('HIPCC', 'hipcc')
    with tempfile.TemporaryDirectory(prefix='xomc_build_') as tmp:
        objs = []
        for src in HIP_SOURCES:
            obj = os.path.join(tmp, src + '.o')
            _run([hipcc] + HIP_FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj],
                 verbose)
            objs.append(obj)
        for src in CXX_SOURCES:
            obj = os.path.join(tmp, src + '.o')
            _run(['g++', '-O2', '-std=c++17', '-fPIC', '-Wall', '-pthread',
                  '-c', os.path.join(CSRC, src), '-o', obj], verbose)
            objs.append(obj)
        tmp_lib = LIB_PATH + '.tmp'
        _run([hipcc, '-shared', '-fPIC', f'--offload-arch={OFFLOAD_ARCH}',
              '-o', tmp_lib] + objs + ['-pthread'], verbose)
        os.replace(tmp_lib, LIB_PATH)
    return LIB_PATH


def build(verbose: bool = False) -> str:
    path = build_library(verbose=verbose)
    # Fill the seed-table cache with the freshly built library.
    from . import _native
    _native.reload_library()
    _native.safeprime_table(ensure_cache=True)
    return path


if __name__ == '__main__':
    build(verbose='-v' in sys.argv)
    print(LIB_PATH)
