"""CPU sanitizer runs of the host C and the oracle.

`make -C oracle asan` and `make -C myers-gpu_amd san` build the oracle, the
host library (myers_host.c, myers_serve.c) and the data-preparation library
(myers_synth.c) with gcc under AddressSanitizer and UndefinedBehaviorSanitizer
(-fno-sanitize-recover: the first report aborts).  The CPU test files that
exercise them -- the oracle against the golden fixtures written by the
reference's own programs, the converters and formats, the C-ABI checks that
need no GPU -- then run again in a child pytest that loads those builds
(libasan preloaded into the interpreter, leak checks off: CPython keeps its
allocations to the end by design).  GPU code is not involved.
"""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_ORACLE = os.path.join(REPO, "oracle", "_san", "liboracle.so")
SAN_LIB = os.path.join(REPO, "myers-gpu_amd", "build", "san", "libmyers_amd.so")
SAN_TOOLS = os.path.join(REPO, "myers-gpu_amd", "build", "san", "libmyers_tools.so")
TESTS = ["tests/test_oracle.py", "tests/test_tools.py", "tests/test_formats.py", "tests/test_abi.py"]


def _runtime(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def sanitized_builds():
    if _runtime("libasan.so") is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "asan"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(["make", "-C", os.path.join(REPO, "myers-gpu_amd"), "san"], check=True,
                   stdout=subprocess.DEVNULL)
    return SAN_ORACLE, SAN_LIB, SAN_TOOLS


def _env(oracle, lib, tools):
    return dict(os.environ,
                LD_PRELOAD=_runtime("libasan.so"),
                ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:verify_asan_link_order=0",
                UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
                MYERS_ORACLE_SO=oracle, MYERS_AMD_LIB=lib, MYERS_TOOLS_LIB=tools)


PROBE = """
import sys
sys.path[:0] = ['tests', 'myers-gpu_amd/python']
import myers_amd as M, oracle_lib as O
M.lib(); M.tools(); O.load()
maps = open('/proc/self/maps').read()
print('asan' if 'libasan' in maps else 'no-asan')
for p in (M.LIB_PATH, M.TOOLS_PATH, O.ORACLE_SO):
    print(p in maps, p)
"""


def test_child_loads_the_sanitized_builds(sanitized_builds):
    run = subprocess.run([sys.executable, "-c", PROBE], cwd=REPO, env=_env(*sanitized_builds),
                         capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stderr[-2000:]
    lines = run.stdout.split("\n")
    assert lines[0] == "asan", run.stdout
    assert all(line.startswith("True ") for line in lines[1:4]), run.stdout
    assert sorted(x.split(" ", 1)[1] for x in lines[1:4]) == sorted(sanitized_builds), run.stdout


def test_cpu_suite_under_asan_ubsan(sanitized_builds):
    env = _env(*sanitized_builds)
    run = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                          *TESTS], cwd=REPO, env=env, capture_output=True, text=True, timeout=900)
    report = run.stdout[-3000:] + run.stderr[-3000:]
    assert run.returncode == 0, report
    assert "ERROR: AddressSanitizer" not in run.stderr, report
    assert "runtime error:" not in run.stderr, report
    assert " passed" in run.stdout, report
