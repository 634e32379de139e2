"""Host code under AddressSanitizer, UndefinedBehaviorSanitizer and LeakSanitizer (no GPU needed).

Builds the library's host sources (parser, program compiler, symptom planning, DEM parser, reference-
sample tableau, sampler/batch/server front ends, m2d column build) with the sanitizers into a temporary
directory, links them with the already-built kernel objects (never launched here) and runs
tests/sanitize/host_driver.cc over the fixture circuits, the fixture DEMs, the reference-sample
circuits and truncated prefixes of them. Any sanitizer report fails the test.
"""

import concurrent.futures
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, 'stim_amd', 'csrc')
KERNEL_OBJS = [os.path.join(ROOT, 'stim_amd', '_build', f'{k}.o')
               for k in ('frame_kernel', 'symptom_kernel', 'm2d_kernel')]
HOST_SOURCES = ['circuit', 'compiler', 'symptom', 'dem', 'tableau', 'expand', 'sampler', 'batch', 'server', 'm2d']
CLANG = '/opt/rocm/llvm/bin/clang++'
FLAGS = ['-x', 'c++', '-O1', '-g', '-fPIC', '-std=c++17', '-D__HIP_PLATFORM_AMD__', '-I/opt/rocm/include',
         '-I' + SRC, '-I' + os.path.join(ROOT, 'include'), '-fsanitize=address,undefined',
         '-fno-omit-frame-pointer', '-fno-sanitize-recover=undefined']


def test_host_code_is_clean_under_sanitizers(tmp_path, golden):
    if not os.path.exists(CLANG) or not all(os.path.exists(o) for o in KERNEL_OBJS):
        pytest.skip('needs ROCm clang++ and the built kernel objects (run the library build first)')
    sources = [(os.path.join(SRC, f + '.cc'), str(tmp_path / (f + '.o'))) for f in HOST_SOURCES]
    sources.append((os.path.join(ROOT, 'tests', 'sanitize', 'host_driver.cc'), str(tmp_path / 'driver.o')))

    def compile_one(pair):
        src, obj = pair
        r = subprocess.run([CLANG] + FLAGS + ['-c', src, '-o', obj], capture_output=True, text=True)
        return src, r.returncode, r.stderr

    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        for src, rc, err in pool.map(compile_one, sources):
            assert rc == 0, (src, err[-2000:])
    exe = str(tmp_path / 'driver')
    link = [CLANG, '-fsanitize=address,undefined', '-o', exe] + [o for _, o in sources] + KERNEL_OBJS + [
        '-L/opt/rocm/lib', '-lamdhip64', '-Wl,-rpath,/opt/rocm/lib']
    r = subprocess.run(link, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]

    inputs = []
    circuits = os.path.join(golden.dir, 'circuits')
    for name in sorted(os.listdir(circuits)):
        inputs.append(os.path.join(circuits, name))
    for k, case in enumerate(json.load(open(os.path.join(golden.dir, 'ref_samples.json')))):
        text = case.get('circuit') if isinstance(case, dict) else None
        if text:
            p = tmp_path / f'ref_{k}.stim'
            p.write_text(text)
            inputs.append(str(p))
    dems = os.path.join(golden.dir, 'dems')
    inputs += [os.path.join(dems, n) for n in sorted(os.listdir(dems)) if n.endswith('.dem')]
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=1:abort_on_error=0:exitcode=23',
               UBSAN_OPTIONS='print_stacktrace=1:halt_on_error=1:exitcode=24')
    r = subprocess.run([exe] + inputs, capture_output=True, text=True, env=env, timeout=600)
    report = r.stderr[-4000:]
    assert r.returncode == 0, report
    assert 'AddressSanitizer' not in r.stderr and 'runtime error' not in r.stderr, report
    assert r.stdout.startswith('files %d' % len(inputs)), r.stdout
    shutil.rmtree(tmp_path, ignore_errors=True)
