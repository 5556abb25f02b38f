"""The compiled C++ drop-in test (tests/cpp/hip_solver_test.cpp, built with the library): the reference-shaped
wrapper include/drape/sim/hip_solver.hpp over libdrape_hip.so, driven as a reference build would drive a Solver
(libs/sim/include/drape/sim/solver.hpp:85-98)."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "drape-2_amd", "build",
                   "hip_solver_test")


def _run(args, env=None):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (drape-2_amd/build.py builds it with the library)")
    p = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=300, env=env)
    print(p.stdout + p.stderr)
    return p


def test_without_a_device_construction_raises_runtime_error():
    """No visible GPU: isAvailable() is false and constructing the solver raises std::runtime_error naming hip
    (solver_factory.cpp:57)."""
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    p = _run(["no-device"], env)
    assert p.returncode == 0 and "boundary test passed" in p.stdout


@pytest.mark.gpu
def test_cpp_wrapper_drives_the_solver_through_the_c_abi():
    """build, step, readPositions, stats, counters, range, snapshot -> restore bit-identical replay, and the
    reference's exception types for misuse, from a compiled C++ caller."""
    p = _run([])
    assert p.returncode == 0 and "boundary test passed" in p.stdout
