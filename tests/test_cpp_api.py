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


def _write_input(path, w, ms):
    rp, col, ln = w.model.graph()
    with open(path, "wb") as f:
        f.write(np.array([w.n, len(col)], np.uint64).tobytes())
        for a, t in ((rp, np.uint32), (col, np.uint32), (ln, np.float64), (w.model.areas(), np.float64),
                     (w.is_outlet, np.uint8), (w.erodibility, np.float64), (ms, np.float64)):
            f.write(np.ascontiguousarray(a, t).tobytes())


def _run_example(example_bin, tmp_path, w, ms, cap=None):
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(inp, w, ms)
    args = [example_bin, "run", str(inp), str(out)] + ([str(cap)] if cap is not None else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    return np.frombuffer(raw[4:], np.float64), int(np.frombuffer(raw[:4], np.uint32)[0])


@pytest.mark.gpu
def test_cpp_generate_matches_oracle(example_bin, tmp_path):
    """A 2000-site terrain, half the sites with a 0.3 rad maximum slope: the
    reference converges (34 iterations)."""
    from fastlem_amd.workloads import config_plain
    from oracle import flo

    w = config_plain(2000)
    ms = np.full(w.n, np.nan)
    ms[: w.n // 2] = 0.3
    got, iters = _run_example(example_bin, tmp_path, w, ms)
    ref, ref_iters, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init_of(w), w.outlets(),
                                     max_slope=ms)
    assert iters == ref_iters
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


@pytest.mark.gpu
def test_cpp_generate_capped_oscillation(example_bin, plain10k, tmp_path):
    """On the 10000-site terrain with the same slopes the reference never
    converges: from iteration 82 the elevations alternate with period 2, so
    generate() without max_iteration would not return.  With
    set_max_iteration(100) both stop after 100 iterations, bit-identical."""
    from oracle import flo

    w = plain10k
    n = w.n
    ms = np.full(n, np.nan)
    ms[: n // 2] = 0.3  # half the sites have a max slope
    got, iters = _run_example(example_bin, tmp_path, w, ms, cap=100)
    ref, ref_iters, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init_of(w), w.outlets(),
                                     max_slope=ms, max_iteration=100)
    assert iters == ref_iters == 100
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def _build_example(out, libdir, libname):
    cmd = ["g++", "-std=c++17", "-O2", "-pthread", "-I", os.path.join(ROOT, "include"), SRC, "-o", out, "-L",
           libdir, f"-l{libname}", f"-Wl,-rpath,{libdir}"]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


def _run_two(binary, tmp_path, w, ms):
    inp, out = tmp_path / "in2.bin", tmp_path / "out2.bin"
    _write_input(inp, w, ms)
    r = subprocess.run([binary, "run2", str(inp), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    its = np.frombuffer(raw[:8], np.uint32)
    h = np.frombuffer(raw[8:], np.float64).reshape(2, w.n)
    return its, h


def _check_two(its, h, w, ms):
    from oracle import flo

    ref, rit, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init_of(w), w.outlets(),
                               max_slope=ms)
    for k in range(2):
        assert its[k] == rit
        assert np.array_equal(h[k].view(np.uint64), ref.view(np.uint64))


def test_cpp_generate_from_two_threads_emulated(tmp_path):
    """Two generate() calls at once from two threads (each thread its own
    default context) through the CPU emulation of the library: both results
    equal the oracle's bit for bit."""
    from cpu_emu import build as emu_build
    from fastlem_amd.workloads import config_plain

    lib = emu_build.build()
    binary = _build_example(str(tmp_path / "gen_emu"), os.path.dirname(lib), "fastlem_emu")
    w = config_plain(1500, 39.0, 1.5)
    ms = np.full(w.n, np.nan)
    _check_two(*_run_two(binary, tmp_path, w, ms), w, None)


@pytest.mark.gpu
def test_cpp_generate_from_two_threads(example_bin, fbm20k, tmp_path):
    w = fbm20k
    ms = np.full(w.n, np.nan)
    binary = _build_example(str(tmp_path / "gen2"), os.path.join(ROOT, "fastlem_amd", "lib"), "fastlem_hip")
    _check_two(*_run_two(binary, tmp_path, w, ms), w, None)
