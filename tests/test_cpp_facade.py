"""The C++ facade (include/mdft/*.hpp) through its own test program
(tests/cpp/test_facade.cpp): host-only checks on the CPU, the transforms,
Convolution (from a Luc file), Solvent functions and the evaluation against the
FP64 oracle on the GPU."""
import os
import subprocess

import pytest

from mini_mdft_amd import luc

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
BIN = os.path.join(CPP, "_build", "test_facade")


def _binary():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-C", CPP], check=True, capture_output=True)
    return BIN


def test_facade_host_only():
    r = subprocess.run([_binary(), "--host-only"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout


@pytest.mark.gpu
def test_facade_gpu(tmp_path):
    # nmax = 3 table (252 tuples) with rows past the largest |k| of the
    # program's grids (8^3 in a 10 A box and smaller)
    tab = luc.synthetic_c(3, 2, luc.nq_for(5.0, luc.DQ_TIP3P), seed=21)
    path = str(tmp_path / "c_nmax3.luc")
    luc.write_luc(path, tab)
    r = subprocess.run([_binary(), path, "252"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
