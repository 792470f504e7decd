"""The multipole field sum divides by n! through a correctly rounded
reciprocal and one residual step (st_rfact in st_maps.h); inside the
kernels' window that must equal the division the reference performs
(simtrack.cpp:4103-4107, `/fac`).  tests/cpp/kdiv_check.c checks it on the
CPU, whose fma and division are IEEE like the GPU's; dropping the residual
step must fail the same check."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cc = shutil.which("cc") or shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path_factory.mktemp("kdiv") / "kdiv_check")
    subprocess.run([cc, "-O2", "-ffp-contract=off", os.path.join(HERE, "cpp", "kdiv_check.c"), "-o", exe, "-lm"],
                   check=True)
    return exe


def test_reciprocal_step_equals_division(checker):
    r = subprocess.run([checker, "exact", "1000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "tested" in r.stdout and " 0 differ" in r.stdout


def test_missing_residual_step_is_caught(checker):
    r = subprocess.run([checker, "nocorrect", "200000"], capture_output=True, text=True)
    assert r.returncode == 1, r.stdout
