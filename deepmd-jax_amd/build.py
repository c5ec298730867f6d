"""Build libdpmd_hip.so for gfx950 with hipcc (in-tree, next to the Python package).

Usage: python deepmd-jax_amd/build.py [--jobs N]
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
OUT_DIR = os.path.join(HERE, 'deepmd_jax')
BUILD_DIR = os.path.join(HERE, '_build')
SOURCES = ['api.hip', 'embed.hip', 'fit.hip', 'nlist.hip', 'md.hip', 'pppm.hip', 'dplr.hip']
LIB_NAME = 'libdpmd_hip.so'
# the issue-overlap probe (bench.py --probe-issue): a library of its own, outside the product
PROBE_SOURCE = 'probe.hip'
PROBE_LIB_NAME = 'libdpmd_probe.so'

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = [
    '-std=c++17', '-O3', '-fPIC', '--offload-arch=gfx950',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-munsafe-fp-atomics', '-Wno-unused-result',
    '-I', os.path.join(ROOT, 'include'), '-I', CSRC,
]


def _compile(src):
    obj = os.path.join(BUILD_DIR, src.replace('.hip', '.o'))
    cmd = [HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s\n%s' % (src, ' '.join(cmd), r.stderr[-8000:]))
    return obj


def build(jobs=4, verbose=True):
    os.makedirs(BUILD_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES + [PROBE_SOURCE]))
    probe_obj = objs.pop()
    _link([probe_obj], os.path.join(OUT_DIR, PROBE_LIB_NAME), [])
    lib = os.path.join(OUT_DIR, LIB_NAME)
    _link(objs, lib, ['-L/opt/rocm/lib', '-lhipfft', '-Wl,-rpath,/opt/rocm/lib'])
    if verbose:
        print('built', lib)
    return lib


def _link(objs, lib, extra):
    tmp = lib + '.tmp'
    cmd = [HIPCC, '-shared', '--offload-arch=gfx950', '-o', tmp] + objs + extra
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n%s\n%s' % (' '.join(cmd), r.stderr[-8000:]))
    os.replace(tmp, lib)


if __name__ == '__main__':
    ap = argparse.ArgumentParser()
    ap.add_argument('--jobs', type=int, default=4)
    args = ap.parse_args()
    try:
        build(args.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
