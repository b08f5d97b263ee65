"""The library's host-side shape arithmetic and walk programs under AddressSanitizer and
UndefinedBehaviorSanitizer.

``csrc/layout.hpp`` (tiling, kernel choice, workspace layouts of hx_loglik* and of a job
evaluation) and ``csrc/walk.hpp`` (post- and pre-order walk programs, cherry rows) are free
of HIP and are what the plan and workspace host code run. They are built here with g++,
``-fsanitize=address,undefined -fno-sanitize-recover=all``, from the library's own headers:
``native/layout_check.cpp`` sweeps shapes and trees checking the layouts' invariants, and
the two program dump drivers run on random, caterpillar and balanced trees. Any sanitizer
report fails the test.
"""
import shutil
import subprocess

import numpy as np
import pytest

from torchtree_amd import synthetic
from torchtree_amd.build import CSRC

HERE = __import__("pathlib").Path(__file__).parent
INCLUDE = HERE.parent / "include"
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def _build(tmp_path_factory, src, name):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("san") / name
    r = subprocess.run(["g++", "-O1", "-g", "-std=c++17", *SAN, f"-I{CSRC}", f"-I{INCLUDE}",
                        str(HERE / "native" / src), "-o", str(out)], capture_output=True, text=True)
    if r.returncode != 0 and "asan" in (r.stderr or "").lower():
        pytest.skip("g++ has no sanitizer runtime here: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr
    return out


def _env():
    return {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0",
            "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1", "PATH": "/usr/bin:/bin"}


@pytest.fixture(scope="module")
def layout_bin(tmp_path_factory):
    return _build(tmp_path_factory, "layout_check.cpp", "layout_check")


@pytest.fixture(scope="module")
def dump_bins(tmp_path_factory):
    return (_build(tmp_path_factory, "walk_dump.cpp", "walk_dump"),
            _build(tmp_path_factory, "pre_walk_dump.cpp", "pre_walk_dump"))


def test_layouts_and_walks_sanitized(layout_bin):
    r = subprocess.run([str(layout_bin)], capture_output=True, text=True, env=_env(), timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert r.stdout.startswith("ok "), r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert int(r.stdout.split()[1]) > 100000


def _trees():
    rng = np.random.default_rng(11)
    for T in (2, 3, 7, 100, 777):
        topo = synthetic.random_topology(T, rng)
        yield T, [tuple(int(x) for x in r) for r in topo.postorder]
    T = 65
    yield T, [(T + t - 1, 0 if t == 1 else T + t - 2, t) for t in range(1, T)]  # caterpillar


@pytest.mark.parametrize("which", [0, 1])
def test_walk_dumps_sanitized(dump_bins, which):
    for T, post in _trees():
        inp = f"{T}\n" + "\n".join(f"{v} {a} {b}" for v, a, b in post) + "\n"
        r = subprocess.run([str(dump_bins[which])], input=inp, capture_output=True, text=True,
                           env=_env(), timeout=120)
        assert r.returncode == 0, (T, r.stderr[-4000:])
        assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
        assert len(r.stdout.splitlines()) >= T  # the header line and T - 1 entries at least
