"""Pinning the oracle's LASD, SGS-stress, convection, pressure and
time-advance functions, and showing that each pin catches a mistake.

test_pin runs every closed-form check of oracle_pins.py on the oracle the
other tests and the benchmark use.  test_planted_mistake_is_caught builds
les_oracle.cpp once per planted mistake (-DORC_PLANT=n; the default build
has none) into a temporary directory and checks that at least one of the
checks named for that mistake fails on it.
"""
import concurrent.futures
import pathlib
import subprocess

import pytest

from oracle_pins import PINS, PLANTED

ORACLE = pathlib.Path(__file__).resolve().parents[1] / "oracle"
FLAGS = ["-O1", "-std=c++17", "-fopenmp", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-mfma", "-shared"]


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(name):
    PINS[name](None)


@pytest.fixture(scope="session")
def planted_libs(tmp_path_factory):
    out = tmp_path_factory.mktemp("planted")

    def build(n):
        so = out / f"liblesoracle_plant{n}.so"
        r = subprocess.run(["g++", *FLAGS, f"-DORC_PLANT={n}", "-o", str(so), str(ORACLE / "les_oracle.cpp")],
                           capture_output=True, text=True)
        return n, so if r.returncode == 0 else r.stderr

    with concurrent.futures.ThreadPoolExecutor(8) as ex:
        return dict(ex.map(build, sorted(PLANTED)))


def test_every_planted_mistake_has_a_check():
    for n, (_, pins) in PLANTED.items():
        assert pins and all(p in PINS for p in pins), n


@pytest.mark.parametrize("n", sorted(PLANTED))
def test_planted_mistake_is_caught(n, planted_libs):
    so = planted_libs[n]
    assert isinstance(so, pathlib.Path), f"build of planted mistake {n} failed:\n{so}"
    what, pins = PLANTED[n]
    failed = []
    for p in pins:
        try:
            PINS[p](so)
        except AssertionError as e:
            failed.append(f"{p}: {e}")
    assert failed, f"planted mistake {n} ({what}) passes {pins}"
