"""The C++ host API (include/fastlem.hpp): the reference's error behaviour on
the CPU, and a full generate() on the GPU checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, graph_of, init_of

SRC = os.path.join(ROOT, "tests", "cpp", "generate_example.cpp")


@pytest.fixture(scope="module")
def example_bin(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cpp") / "generate_example")
    libdir = os.path.join(ROOT, "fastlem_amd", "lib")
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), SRC, "-o", out, "-L", libdir,
           "-lfastlem_hip", f"-Wl,-rpath,{libdir}"]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


def test_cpp_generation_errors(example_bin):
    r = subprocess.run([example_bin, "errors"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "errors 4/4" in r.stdout


@pytest.mark.gpu
def test_cpp_generate_matches_oracle(example_bin, plain10k, tmp_path):
    from oracle import flo

    w = plain10k
    rp, col, ln = w.model.graph()
    n = w.n
    ms = np.full(n, np.nan)
    ms[: n // 2] = 0.3  # half the sites have a max slope
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([n, len(col)], np.uint64).tobytes())
        for a, t in ((rp, np.uint32), (col, np.uint32), (ln, np.float64), (w.model.areas(), np.float64),
                     (w.is_outlet, np.uint8), (w.erodibility, np.float64), (ms, np.float64)):
            f.write(np.ascontiguousarray(a, t).tobytes())
    out = tmp_path / "out.bin"
    r = subprocess.run([example_bin, "run", str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    iters = int(np.frombuffer(raw[:4], np.uint32)[0])
    got = np.frombuffer(raw[4:], np.float64)
    ref, ref_iters, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init_of(w), w.outlets(),
                                     max_slope=ms)
    assert iters == ref_iters
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
