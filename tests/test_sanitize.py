"""The host C++ of libesrhip.so (mesh generators, stable CSR build, RCM,
block order, bandwidth, P1 assembly, smoothed-aggregation levels) built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++, host code only) and driven over small and
odd sizes and the error paths by tests/sanitize/host_driver.cpp; any memory
error or undefined behaviour aborts the driver."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, 'easier_amd', 'csrc')


@pytest.mark.skipif(shutil.which('g++') is None, reason='needs g++')
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / 'host_driver')
    cmd = ['g++', '-std=c++17', '-O1', '-g', '-fsanitize=address,undefined',
           '-fno-sanitize-recover=all', '-fno-omit-frame-pointer',
           f'-I{REPO}/include', f'-I{CSRC}',
           os.path.join(REPO, 'tests', 'sanitize', 'host_driver.cpp'),
           os.path.join(CSRC, 'host_mesh.cpp'), os.path.join(CSRC, 'host_fem.cpp'),
           os.path.join(CSRC, 'host_amg.cpp'),
           '-o', exe]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    env = dict(os.environ,
               ASAN_OPTIONS='detect_leaks=1:abort_on_error=0:verify_asan_link_order=0',
               UBSAN_OPTIONS='print_stacktrace=1')
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert 'host entry points clean' in run.stdout
