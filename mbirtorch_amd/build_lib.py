"""Build libmbir_hip.so (the gfx950 projector / VCD library) in-tree.

Usage: ``python -m mbirtorch_amd.build_lib`` (or ``build()`` from Python).
Each ``csrc/*.hip`` compiles to an object in ``_build/`` (parallel jobs) and
the objects link into ``mbirtorch_amd/libmbir_hip.so``, which the package
loads with ctypes.  Always rebuilds.
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libmbir_hip.so')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')

ARCH = os.environ.get('MBIR_OFFLOAD_ARCH', 'gfx950')

# -ffp-contract=off: every float32 geometry op rounds on its own, as the
# reference's eager chain does; accumulations use fmaf() explicitly.
# -packed-fp32-ops is switched off for all device code.
HIP_FLAGS = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
             '-ffp-contract=off', '-Xclang', '-target-feature', '-Xclang',
             '-packed-fp32-ops', '-I', INCLUDE, '-Wno-unused-result']


def _hipcc():
    for cand in (shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found; the HIP library cannot be built')


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def build(verbose=False):
    hipcc = _hipcc()
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.basename(src)[:-4] + '.o')
        cmd = [hipcc, *HIP_FLAGS, '-c', src, '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f'hipcc failed for {src}:\n{res.stdout}\n{res.stderr}')
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + '.tmp'
    cmd = [hipcc, f'--offload-arch={ARCH}', '-shared', '-fPIC', '-pthread', '-o', tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'link failed:\n{res.stdout}\n{res.stderr}')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
