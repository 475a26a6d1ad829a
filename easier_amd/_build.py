"""Builds libesrhip.so in-tree: HIP sources with hipcc for gfx950, host C++
with g++, linked by hipcc against hipRTC.

Output: easier_amd/_lib/libesrhip.so (git-ignored; it travels with the tree).
"""
import concurrent.futures as _cf
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, 'csrc')
INCLUDE = os.path.join(REPO_DIR, 'include')
LIB_DIR = os.path.join(PKG_DIR, '_lib')
OBJ_DIR = os.path.join(REPO_DIR, '_build', 'obj')
LIB_PATH = os.path.join(LIB_DIR, 'libesrhip.so')

ARCH = 'gfx950'
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
CXX = os.environ.get('CXX', 'g++')

# Every HIP translation unit: gfx950 only, no FMA contraction (results must
# match the CPU reference op by op), no packed-FP32 instruction selection.
HIP_FLAGS = [
    f'--offload-arch={ARCH}', '-O3', '-fPIC', '-std=c++17',
    '-ffp-contract=off',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
    f'-I{INCLUDE}', f'-I{CSRC}',
]
CXX_FLAGS = ['-O3', '-fPIC', '-std=c++17', '-ffp-contract=off', f'-I{INCLUDE}',
             f'-I{CSRC}', '-Wall', '-Wno-unused-function']

HIP_SOURCES = ['runtime.hip', 'k_basic.hip', 'k_solver.hip', 'k_sell.hip', 'k_batched.hip',
               'k_mrhs.hip', 'k_gnn.hip', 'k_fcg.hip', 'k_mg.hip']
CXX_SOURCES = ['host_mesh.cpp', 'host_fem.cpp', 'host_amg.cpp']


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    if proc.returncode != 0:
        raise RuntimeError('build step failed:\n  ' + ' '.join(cmd) + '\n'
                           + proc.stdout)
    return proc.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ_DIR, src + '.o')
    if src.endswith('.hip'):
        cmd = [HIPCC, '-c', path, '-o', obj] + HIP_FLAGS
    else:
        cmd = [CXX, '-c', path, '-o', obj] + CXX_FLAGS
    _run(cmd)
    return obj


def build(verbose=False, jobs=None):
    """Compile every source and link libesrhip.so (always rebuilds)."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    sources = [s for s in HIP_SOURCES + CXX_SOURCES
               if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(8, len(sources))
    with _cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB_PATH + '.tmp'
    _run([HIPCC, '-shared', '-fPIC', f'--offload-arch={ARCH}', '-o', tmp]
         + objs + ['-L/opt/rocm/lib', '-lhiprtc', '-lamdhip64',
                   '-Wl,-rpath,/opt/rocm/lib'])
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f'built {LIB_PATH} from {len(sources)} sources', file=sys.stderr)
    return LIB_PATH


if __name__ == '__main__':
    build(verbose=True)
